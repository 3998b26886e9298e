// ReadData.cpp -- the input deck and everything derived from it, host side.
//
// dem_derive restates, for the engine, the derived quantities of
// ReadData.cpp:216-375 (diameters, dt, size levels, nNebors, level
// membership, domain, skin, cells), the reduced units of
// UnitReduce.cpp:14-104 and the contact constants of main.cpp:328-348.
// dem_read_deck reads the heyi.in keyword file (ReadData.cpp:7-188): a key
// is matched against the start of a line and its values follow on the next
// lines.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>

#include "../../include/dem_capi.h"
#include "dem_host.h"

static const double kPI = 3.1415926;  // heydem.h:15

static void set_law(dem_calpara *c, double emod, double pois, double sfrc, double ystress) {
    c->emod = emod;
    c->pois = pois;
    c->sfrc = sfrc;
    c->ystress = ystress;
    c->ydsp_t = pow((kPI * c->ystress / (2.0 * c->emod)), 2.0);
    c->yforce_t = 4.0 / 3.0 * c->emod * pow(c->ydsp_t, 1.5);
    c->ystiff_t = kPI * c->ystress;
}

void dem_calpara_from_material(dem_setup *S) {
    set_law(&S->pp, (S->emod[0] * S->emod[0]) / (S->emod[0] + S->emod[0]),
            2.0 * (S->pois[0] * S->pois[0]) / (S->pois[0] + S->pois[0]), S->sfrc[0],
            S->yldp[0] * S->yldp[0] / (S->yldp[0] + S->yldp[0]));
    set_law(&S->pw, S->emod[0], 2.0 * (S->pois[0] * S->pois[1]) / (S->pois[0] + S->pois[1]),
            pow(S->sfrc[1] * S->sfrc[1], 0.5), S->yldp[0]);
}

extern "C" int dem_derive(const dem_input *in, const double *dia_m, dem_setup *S, double *rad,
                          int *pLevel, int *neborIdx) {
    if (!in || !S || !rad || !pLevel || !neborIdx) return dem_fail(DEM_ERR_ARG, "null argument");
    memset(S, 0, sizeof(*S));
    const int NP = in->NP, nL = in->nLevels;
    if (NP <= 0) return dem_fail(DEM_ERR_ARG, "NP must be positive");
    if (nL <= 0 || nL > DEM_MAX_LEVELS) return dem_fail(DEM_ERR_ARG, "nLevels out of range");
    if (in->realSize <= 0) return dem_fail(DEM_ERR_ARG, "REALSIZE must be positive");
    const double g = in->gravity > 0 ? in->gravity : (double)9.8f;  // heydem.h:16

    for (int m = 0; m < 2; ++m) {
        S->density[m] = in->mat[m][0];
        S->ymod[m] = in->mat[m][1];
        S->pois[m] = in->mat[m][2];
        S->sfrc[m] = in->mat[m][3];
        S->rfrc[m] = in->mat[m][4];
        S->dmpn[m] = in->mat[m][5];
        S->yldp[m] = in->mat[m][6];
        S->emod[m] = S->ymod[m] / (1 - S->pois[m] * S->pois[m]);
    }
    S->stension = in->stension;
    S->vol = in->vol;
    S->gapmn = in->gapmn;
    S->cirX = in->cirX;
    S->cirY = in->cirY;
    S->Bw = in->Bw;
    S->R = in->R;
    S->Tw = in->Bw + in->H;
    S->topv = 0.0;  // left uninitialised by the reference; a fixed lid
    S->btmv = 0.0;
    S->feedLmin = in->feedLmin;
    S->feedLmax = in->feedLmax;
    S->tPacking = in->tPacking;
    S->forceCap = in->forceCap;
    S->loadStage = in->loadStage;
    S->theta = in->theta;
    S->ruptureMode = in->ruptureMode;
    S->gridClampQuirk = in->gridClampQuirk;
    S->nLevels = nL;

    for (int i = 0; i < NP; ++i) {
        rad[i] = (dia_m ? dia_m[i] : in->realSize) / 2;
        if (!(rad[i] > 0)) return dem_fail(DEM_ERR_ARG, "diameters must be positive");
    }

    // time step from the smallest particle's Hertz stiffness
    double kmod = 2.0 * S->emod[0] * S->emod[0] / (S->emod[0] + S->emod[0]);
    double rmin = rad[0], rmax = rad[0];
    for (int i = 1; i < NP; ++i) {
        if (rad[i] <= rmin) rmin = rad[i];
        if (rad[i] >= rmax) rmax = rad[i];
    }
    double dmin = 2.0 * rmin;
    double kstiff = 1.333 * kmod * sqrt(dmin * 0.25);
    double mmin = S->density[0] * kPI * pow(dmin, 3.0) / 6.0;
    double dt = in->dtFactor * kPI * sqrt(0.5 * mmin / kstiff) * pow(10.0, -0.2);
    if (in->dtOverride > 0) dt = in->dtOverride;

    // size levels and their neighbour-slot budget
    for (int l = 0; l < nL; ++l)
        S->levelSize[l] = in->sizeLevelInput ? in->levelDia[l] / 2
                                             : (rmax - rmin) / nL * (l + 1) + rmin;
    for (int l = 0; l < nL; ++l) {
        double q = S->levelSize[l] / rmin;
        S->levelNNebors[l] = (int)ceil(1.2 * (pow((q + 2 + in->cutGap), 3) - pow(q, 3)));
        S->levelNPars[l] = 0;
    }
    for (int i = 0; i < NP; ++i) {
        int lv = -1;
        if (rad[i] <= S->levelSize[0]) {
            lv = 0;
        } else {
            for (int l = 1; l < nL; ++l)
                if (rad[i] > S->levelSize[l - 1] && rad[i] <= S->levelSize[l]) lv = l;
        }
        if (lv < 0) return dem_fail(DEM_ERR_LEVEL, "particle %d (r=%g m) fits no size level", i, rad[i]);
        pLevel[i] = lv;
        neborIdx[i] = S->levelNPars[lv]++;
    }

    // computational box and cells
    double ex = in->expandSize * rmax;
    double lo[3] = {(S->cirX - S->R) - ex, (S->cirY - S->R) - ex, S->Bw - ex};
    double hi[3] = {(S->cirX + S->R) + ex, (S->cirY + S->R) + ex, S->Tw + ex};
    double skin = in->cutGap * (rmin / rmax);
    for (int l = 0; l < nL; ++l) {
        S->cellSize[l] = (1.0 + skin * 1.4) * 2.0 * S->levelSize[l];
        for (int d = 0; d < 3; ++d) S->gridSize[l][d] = (int)ceil((hi[d] - lo[d]) / S->cellSize[l]);
        double nc = (double)S->gridSize[l][0] * S->gridSize[l][1] * S->gridSize[l][2];
        if (nc > 2.0e9) return dem_fail(DEM_ERR_ARG, "grid of level %d too large", l);
        S->numCells[l] = S->gridSize[l][0] * S->gridSize[l][1] * S->gridSize[l][2];
    }

    // reduced units: length = d_max, mass = its mass, force = m g
    const double L0 = in->realSize;
    S->rlunit = L0;
    S->rmunit = S->density[0] * kPI * pow(L0, 3.0) / 6.0;
    S->rfunit = S->rmunit * g;
    S->rsunit = S->rfunit / pow(L0, 2.0);
    S->rtunit = pow(L0 / g, 0.5);
    S->rvunit = S->rlunit / S->rtunit;
    S->reunit = S->rmunit * g * S->rlunit;
    S->cirX /= S->rlunit;
    S->cirY /= S->rlunit;
    S->Tw /= S->rlunit;
    S->Bw /= S->rlunit;
    S->R /= S->rlunit;
    for (int l = 0; l < nL; ++l) S->levelSize[l] /= S->rlunit;
    S->feedLmax /= S->rlunit;
    S->feedLmin /= S->rlunit;
    S->tPacking /= S->rtunit;
    for (int i = 0; i < NP; ++i) rad[i] /= S->rlunit;
    for (int m = 0; m < 2; ++m) {
        S->emod[m] = S->emod[m] * pow(S->rlunit, 2.0) / S->rfunit;
        S->yldp[m] = S->yldp[m] * pow(S->rlunit, 2.0) / S->rfunit;
        S->dmpn[m] = S->dmpn[m] / S->rtunit;
    }
    S->gapmn = S->gapmn / L0;
    S->layer = 0.5 * (pow((1.0 + S->vol), 1.0 / 3.0) - 1.0);  // from the unscaled fraction
    S->vol = S->vol * kPI / 6.0;
    S->stension = S->stension * S->rlunit / S->rfunit;
    S->brknmx = 0.99 * pow((S->vol / 6.0), 0.34);
    S->dt = dt / S->rtunit;
    for (int d = 0; d < 3; ++d) {
        S->origin[d] = lo[d] / S->rlunit;
        S->endPoint[d] = hi[d] / S->rlunit;
    }
    for (int l = 0; l < nL; ++l) S->cellSize[l] /= S->rlunit;
    S->searchGap = skin;
    S->radMin = rmin / S->rlunit;
    S->radMax = rmax / S->rlunit;

    dem_calpara_from_material(S);

    long long slots = 0, cells = 0;
    for (int l = 0; l < nL; ++l) {
        S->levelNebSId[l] = (int)slots;
        slots += (long long)S->levelNPars[l] * S->levelNNebors[l];
        cells += S->numCells[l];
        if (cells > 2000000000LL) return dem_fail(DEM_ERR_ARG, "too many cells");
        S->ncellIdx[l] = (int)cells;
    }
    if (slots > 2000000000LL) return dem_fail(DEM_ERR_ARG, "too many neighbour slots");
    S->nPPCnt = (int)slots;
    S->sysNCells = (int)cells;
    return DEM_OK;
}

// ---------------------------------------------------------------------
// heyi.in keyword deck (ReadData.cpp:7-188, main.cpp:65-68)
// ---------------------------------------------------------------------
namespace {
struct Deck {
    FILE *f;
    // positions the stream after the first line that starts with key
    bool seek(const char *key) {
        rewind(f);
        char line[512];
        size_t n = strlen(key);
        while (fgets(line, sizeof line, f))
            if (strncmp(line, key, n) == 0) return true;
        return false;
    }
    bool nums(double *v, int n) {
        for (int k = 0; k < n; ++k)
            if (fscanf(f, "%lf", &v[k]) != 1) return false;
        return true;
    }
    bool yes() {
        char w[64] = {0};
        if (fscanf(f, "%63s", w) != 1) return false;
        return w[0] == 'Y' || w[0] == 'y';
    }
};
}  // namespace

extern "C" int dem_read_deck(const char *path, dem_input *in) {
    if (!path || !in) return dem_fail(DEM_ERR_ARG, "null argument");
    FILE *f = fopen(path, "r");
    if (!f) return dem_fail(DEM_ERR_ARG, "cannot open deck %s", path);
    Deck d{f};
    memset(in, 0, sizeof(*in));
    in->loadStage = 1;
    double v[16];
    int rc = DEM_OK;
    auto need = [&](const char *key, int n) -> bool {
        if (!d.seek(key) || !d.nums(v, n)) {
            rc = dem_fail(DEM_ERR_ARG, "deck %s: missing or short keyword %s", path, key);
            return false;
        }
        return true;
    };
    do {
        if (!need("NUMPARTICLE", 1)) break;
        in->NP = (int)v[0];
        if (!need("SIZELEVEL", 1)) break;
        in->nLevels = (int)v[0];
        if (!need("LOADMODE", 1)) break;
        in->loadStage = (int)v[0];
        if (!need("packing_TIME", 1)) break;
        in->tPacking = v[0];
        if (!d.seek("FoceCapillary")) { rc = dem_fail(DEM_ERR_ARG, "deck: FoceCapillary missing"); break; }
        in->forceCap = d.yes() ? 1 : 0;
        if (!need("REALSIZE", 1)) break;
        in->realSize = v[0];
        if (!need("CUTGAPFORNEIGHBORLIST", 1)) break;
        in->cutGap = v[0];
        if (!need("EXPANDSIZE", 1)) break;
        in->expandSize = v[0];
        if (!d.seek("ContainerType")) { rc = dem_fail(DEM_ERR_ARG, "deck: ContainerType missing"); break; }
        char ct[64] = {0};
        if (fscanf(f, "%63s", ct) != 1 || strncmp(ct, "Cylinder", 8) != 0) {
            rc = dem_fail(DEM_ERR_ARG, "deck: only ContainerType Cylinder is supported");
            break;
        }
        if (!need("CylinderBC", 5)) break;
        in->cirX = v[0]; in->cirY = v[1]; in->Bw = v[2]; in->R = v[3]; in->H = v[4];
        if (!need("FeedZone", 2)) break;
        in->feedLmin = v[0]; in->feedLmax = v[1];
        if (!need("MATERIAL", 14)) break;
        for (int m = 0; m < 2; ++m)
            for (int k = 0; k < 7; ++k) in->mat[m][k] = v[7 * m + k];
        if (!need("LIQUID", 3)) break;
        in->stension = v[0]; in->vol = v[1]; in->gapmn = v[2];
        if (!need("SIMULATION", 2)) break;
        in->dtFactor = v[0]; in->nOutPacking = (int)v[1];
        if (d.seek("SIZELEVEL_INPUT") && d.yes()) {
            in->sizeLevelInput = 1;
            if (in->nLevels > DEM_MAX_LEVELS || !d.nums(in->levelDia, in->nLevels)) {
                rc = dem_fail(DEM_ERR_ARG, "deck: SIZELEVEL_INPUT sizes missing");
                break;
            }
        }
        // optional overrides of this engine
        if (d.seek("GRAVITY") && d.nums(v, 1)) in->gravity = v[0];
        if (d.seek("TIMESTEP") && d.nums(v, 1)) in->dtOverride = v[0];
        if (d.seek("CONTACTANGLE") && d.nums(v, 1)) in->theta = v[0];
        if (d.seek("RUPTUREMODE") && d.nums(v, 1)) in->ruptureMode = (int)v[0];
    } while (0);
    fclose(f);
    return rc;
}
