// Allocation.cpp -- context creation and device allocation.
//
// One dem_ctx replaces the process globals and the allocation helpers of
// the reference: ParticleInfoAlloc (Allocation.cpp:22-53), WallCntHstry_Alloc
// (57-110), PPCntHstry_Alloc (115-174), NeborListAlloc (181-235), ParamAlloc
// (251-287), SearchCellAlloc (291-303), and the constant uploads of
// kernel.cu:22-50.  The p-p history is sized cap * max(nNebors) slots (a
// slot-major layout over the cell-sorted order) instead of the reference's
// level-blocked sum_l nPars[l]*nNebors[l].
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <string>

#include "dem_host.h"
#include "dem_launch.h"

static thread_local std::string g_err;

int dem_fail(int code, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

extern "C" const char *dem_last_error(void) { return g_err.c_str(); }

static void willett_coeffs(double th, double *a) {
    // Willett et al. (2000), force between equal spheres bridged by a
    // liquid of contact angle th (rad); th = 0 gives kernel.cu:73-76.
    a[0] = -0.44507 + th * (0.050832 + th * (-1.1466));
    a[1] = -0.1119 + th * (-0.000411 + th * (-0.1490));
    a[2] = -0.012101 + th * (-0.0036456 + th * (-0.01255));
    a[3] = -0.0005 + th * (-0.0003505 + th * (-0.00029076));
    a[4] = 1.9222 + th * (-0.57473 + th * (-1.2918));
    a[5] = -0.0668 + th * (-0.1201 + th * (-0.22574));
    a[6] = -0.0013375 + th * (-0.0068988 + th * (-0.01137));
    a[7] = 1.268 + th * (-0.01396 + th * (-0.23566));
    a[8] = 0.198 + th * (0.092 + th * (-0.06418));
    a[9] = 0.02232 + th * (0.02238 + th * (-0.009853));
    a[10] = 0.0008585 + th * (0.001318 + th * (-0.00053));
    a[11] = -0.010703 + th * (0.073776 + th * (-0.34742));
    a[12] = 0.03345 + th * (0.04543 + th * (-0.09056));
    a[13] = 0.0018574 + th * (0.004456 + th * (-0.006257));
}

static void fill_law(const dem_calpara &c, double dmpn_a, double dmpn_b, int damp_on, DevLaw &L) {
    L.c43E = 4.0 / 3.0 * c.emod;
    L.ydsp_t = c.ydsp_t;
    L.yforce_t = c.yforce_t;
    L.ystiff_t = c.ystiff_t;
    L.dsmxA = 3.0 * c.sfrc;
    L.dsmxB = 16.0 * (1 - c.pois) * c.emod / (4.0 - 2.0 * c.pois);
    L.sfrc = c.sfrc;
    L.damp = -2.0 * sqrt(dmpn_a * dmpn_b) * c.emod;
    L.damp_on = damp_on;
    L.pad_ = 0;
}

void dem_fill_inst(const dem_setup &S, DevInst &P) {
    memset(&P, 0, sizeof(P));
    fill_law(S.pp, S.dmpn[0], S.dmpn[0], S.loadStage == 1, P.pp);
    fill_law(S.pw, S.dmpn[0], S.dmpn[1], S.loadStage == 1 || S.loadStage == 7, P.pw);
    P.rolling = 0.5 * (S.rfrc[0] + S.rfrc[0]);
    P.dt = S.dt;
    P.searchGap = S.searchGap;
    P.rebuildThr = 0.3 * S.searchGap;
    P.vol = S.vol;
    P.gapmn = S.gapmn;
    P.layer2 = 2 * S.layer;
    P.twoPI = 2 * 3.1415926;
    P.stension = S.stension;
    willett_coeffs(S.theta, P.wc);
    for (int k = 0; k < 14; ++k) P.wcf[k] = (float)P.wc[k];
    P.theta = S.theta;
    P.forceCap = S.forceCap;
    P.ruptureMode = S.ruptureMode;
    P.cirX = S.cirX;
    P.cirY = S.cirY;
    P.R = S.R;
    P.Bw = S.Bw;
    P.topv = S.topv;
    for (int d = 0; d < 3; ++d) P.origin[d] = S.origin[d];
    P.nLevels = S.nLevels;
    for (int l = 0; l < S.nLevels; ++l) {
        P.levelSize[l] = S.levelSize[l];
        P.cellSize[l] = S.cellSize[l];
        P.gx[l] = S.gridSize[l][0];
        P.gy[l] = S.gridSize[l][1];
        P.gz[l] = S.gridSize[l][2];
        P.cellOff[l] = l ? S.ncellIdx[l - 1] : 0;
        P.levelK[l] = S.levelNNebors[l];
    }
    long long tab = 0;
    for (int l = 0; l < S.nLevels; ++l) {
        P.bbx[l] = (P.gx[l] + 3) / 4;
        P.bby[l] = (P.gy[l] + 3) / 4;
        P.blkOff[l] = (int)tab;
        tab += (long long)P.bbx[l] * P.bby[l] * ((P.gz[l] + 3) / 4) * 64;
        if (tab > 0x7fffffffLL) tab = 0x7fffffffLL;  // refused by dem_create
    }
    P.ncellTab = (int)tab;
}

// Interleaved-level Morton keys (DevInst::morton = 2): the level-0 cell
// holding the centre of each cell of every level, by the device's own
// expression, must fit 10 bits per axis, and with the level bits the key 31.
static bool interleaved_layout(const dem_setup &S, int *level_bits, int *key_bits) {
    const double c0 = S.cellSize[0];
    int m = 1;
    for (int l = 0; l < S.nLevels; ++l)
        for (int a = 0; a < 3; ++a) {
            const double top = floor(((double)(S.gridSize[l][a] - 1) + 0.5) * S.cellSize[l] / c0);
            if (!(top < 1023.0)) return false;
            m = (int)top + 1 > m ? (int)top + 1 : m;
        }
    int b = 0;
    while ((1 << b) < m) ++b;
    int lb = 0;
    while ((1 << lb) < S.nLevels) ++lb;
    if (3 * b + lb > 31) return false;
    *level_bits = lb;
    *key_bits = 3 * b + lb < 1 ? 1 : 3 * b + lb;
    return true;
}

// Sort-key layout: Morton codes of the cells when every level grid fits 10
// bits per axis and the key space 31 bits, else the row-major hash.
static bool morton_layout(const dem_setup &S, unsigned off[DEM_MAXL], int *key_bits) {
    unsigned long long acc = 0;
    for (int l = 0; l < S.nLevels; ++l) {
        int m = S.gridSize[l][0];
        if (S.gridSize[l][1] > m) m = S.gridSize[l][1];
        if (S.gridSize[l][2] > m) m = S.gridSize[l][2];
        int b = 0;
        while ((1 << b) < m) ++b;
        if (b > 10) return false;
        off[l] = (unsigned)acc;
        acc += 1ull << (3 * b);
    }
    if (acc > (1ull << 31)) return false;
    int bits = 1;
    while (bits < 31 && (1ull << bits) < acc) ++bits;
    *key_bits = bits;
    return true;
}

void dem_refresh_inst(dem_ctx *c) {
    dem_fill_inst(c->S, c->P);
    int bits = 1;
    while (bits < 31 && ((long long)1 << bits) <= (long long)c->S.sysNCells) ++bits;
    c->key_bits = bits;
    c->P.morton = 0;
    c->P.levelBits = 0;
    int lb = 0, ib = 0;
    if (c->want_morton && c->S.nLevels > 1 && interleaved_layout(c->S, &lb, &ib)) {
        c->P.morton = 2;
        c->P.levelBits = lb;
        c->key_bits = ib;
    } else if (c->want_morton) {
        unsigned off[DEM_MAXL] = {0};
        int mb = 0;
        if (morton_layout(c->S, off, &mb)) {
            c->P.morton = 1;
            for (int l = 0; l < DEM_MAXL; ++l) c->P.mortonOff[l] = off[l];
            c->key_bits = mb;
        }
    }
}

extern "C" int dem_set_order(dem_ctx *c, int morton) {
    if (!c) return dem_fail(DEM_ERR_ARG, "null context");
    c->want_morton = morton != 0;
    dem_refresh_inst(c);
    c->rebuild_pending = true;
    return DEM_OK;
}

template <class T>
static int dalloc(dem_ctx *c, T **p, size_t n) {
    size_t b = (n ? n : 1) * sizeof(T);
    HIPCHK(hipMalloc((void **)p, b));
    HIPCHK(hipMemsetAsync(*p, 0, b, c->stream));
    c->bytes += (long long)b;
    return DEM_OK;
}

static int alloc_all(dem_ctx *c) {
    const size_t cap = (size_t)c->cap, K = (size_t)c->K;
    int rc;
    for (int b = 0; b < 2; ++b) {
        if ((rc = dalloc(c, &c->P4[b], cap)) || (rc = dalloc(c, &c->VW[b], 2 * cap)) ||
            (rc = dalloc(c, &c->lq[b], cap)) || (rc = dalloc(c, &c->level[b], cap)) ||
            (rc = dalloc(c, &c->id[b], cap)) || (rc = dalloc(c, &c->nbr[b], cap * K)) ||
            (rc = dalloc(c, &c->nnb[b], cap)) || (rc = dalloc(c, &c->wl[b], cap * DEM_WALLS)) ||
            (rc = dalloc(c, &c->nwl[b], cap)) || (rc = dalloc(c, &c->hp[b], 2 * cap * K)) ||
            (rc = dalloc(c, &c->amask[b], cap)) || (rc = dalloc(c, &c->actoff[b], cap + 1)) ||
            (rc = dalloc(c, &c->hw[b], cap * DEM_WALLS)))
            return rc;
    }
    if ((rc = dalloc(c, &c->cstart, (size_t)c->P.ncellTab)) ||
        (rc = dalloc(c, &c->cend, (size_t)c->P.ncellTab)) || (rc = dalloc(c, &c->keys, cap)) ||
        (rc = dalloc(c, &c->vals, cap)) || (rc = dalloc(c, &c->keys2, cap)) ||
        (rc = dalloc(c, &c->vals2, cap)) || (rc = dalloc(c, &c->nmask, cap)) ||
        (rc = dalloc(c, &c->sort_tmp, (size_t)radix_tmp_elems((long)cap))))
        return rc;
    for (int q = 0; q < 7; ++q)
        if ((rc = dalloc(c, &c->F[q], cap))) return rc;
    if ((rc = dalloc(c, &c->stop_at, 1)) || (rc = dalloc(c, &c->overflow, 1)) ||
        (rc = dalloc(c, &c->upd_count, 1)) || (rc = dalloc(c, &c->stats, 8)))
        return rc;
    return DEM_OK;
}

static void free_all(dem_ctx *c) {
    dem_pair_free(c);
    for (int b = 0; b < 2; ++b) {
        (void)hipFree(c->P4[b]); (void)hipFree(c->VW[b]);
        (void)hipFree(c->lq[b]); (void)hipFree(c->level[b]);
        (void)hipFree(c->id[b]); (void)hipFree(c->nbr[b]); (void)hipFree(c->nnb[b]);
        (void)hipFree(c->wl[b]); (void)hipFree(c->nwl[b]);
        (void)hipFree(c->hp[b]); (void)hipFree(c->hw[b]);
        (void)hipFree(c->amask[b]); (void)hipFree(c->actoff[b]);
    }
    (void)hipFree(c->cstart); (void)hipFree(c->cend);
    (void)hipFree(c->keys); (void)hipFree(c->vals); (void)hipFree(c->keys2); (void)hipFree(c->vals2);
    (void)hipFree(c->sort_tmp); (void)hipFree(c->nmask);
    for (int q = 0; q < 7; ++q) (void)hipFree(c->F[q]);
    (void)hipFree(c->stop_at); (void)hipFree(c->overflow); (void)hipFree(c->upd_count);
    for (hipEvent_t &e : c->rb_ev)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t &e : c->nf_ev)
        if (e) (void)hipEventDestroy(e);
    (void)hipFree(c->stats);
}

int dem_create_on(const dem_input *in, const double *dia_m, int precision, hipStream_t stream,
                  dem_ctx **out) {
    if (!in || !out) return dem_fail(DEM_ERR_ARG, "null argument");
    if (precision != DEM_PREC_FP64 && precision != DEM_PREC_MIXED)
        return dem_fail(DEM_ERR_ARG, "unknown precision %d", precision);
    if (in->NP < 1) return dem_fail(DEM_ERR_ARG, "NP must be positive");
    // kernel.cu:627-629 clamps cell coordinates to gridSize, one past the
    // last cell, so a particle outside the box hashes into another cell's
    // slot or past the cell table.  That behaviour is not reproduced: cells
    // clamp to gridSize - 1 (the same cell for every particle inside the box)
    if (in->gridClampQuirk)
        return dem_fail(DEM_ERR_ARG, "gridClampQuirk is not supported: cell coordinates clamp to "
                        "gridSize - 1");
    *out = nullptr;
    dem_ctx *c = new dem_ctx();
    c->in = *in;
    c->precision = precision;
    c->cap = in->NP;
    c->rad_h.resize(in->NP);
    c->level_h.resize(in->NP);
    std::vector<int> nidx(in->NP);
    int rc = dem_derive(in, dia_m, &c->S, c->rad_h.data(), c->level_h.data(), nidx.data());
    if (rc) {
        delete c;
        return rc;
    }
    for (int l = 0; l < c->S.nLevels; ++l) c->K = c->S.levelNNebors[l] > c->K ? c->S.levelNNebors[l] : c->K;
    dem_refresh_inst(c);
    if (c->P.ncellTab >= 0x7fffffff) {
        delete c;
        return dem_fail(DEM_ERR_ARG, "cell grids too large");
    }
    c->Tw = c->S.Tw;
    if (stream) {
        c->stream = stream;
        c->own_stream = false;
    } else {
        hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete c;
            return dem_fail(DEM_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(e));
        }
    }
    if ((rc = alloc_all(c)) != DEM_OK ||
        (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess)) {
        if (rc == DEM_OK) rc = dem_fail(DEM_ERR_HIP, "hipEventCreate failed");
        dem_destroy(c);
        return rc;
    }
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        dem_destroy(c);
        return dem_fail(DEM_ERR_HIP, "allocation sync: %s", hipGetErrorString(e));
    }
    *out = c;
    return DEM_OK;
}

extern "C" int dem_create(const dem_input *in, const double *dia_m, int precision, dem_ctx **out) {
    return dem_create_on(in, dia_m, precision, nullptr, out);
}

extern "C" int dem_destroy(dem_ctx *c) {
    if (!c) return DEM_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    free_all(c);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream && c->own_stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return DEM_OK;
}

extern "C" int dem_get_setup(dem_ctx *c, dem_setup *out) {
    if (!c || !out) return dem_fail(DEM_ERR_ARG, "null argument");
    *out = c->S;
    return DEM_OK;
}

extern "C" long long dem_device_bytes(dem_ctx *c) { return c ? c->bytes : 0; }

// --- constant setters (kernel.cu:22-50), per context -----------------------
extern "C" int dem_set_material(dem_ctx *c, const double mat[2][7]) {
    if (!c || !mat) return dem_fail(DEM_ERR_ARG, "null argument");
    for (int m = 0; m < 2; ++m) {
        c->S.density[m] = mat[m][0];
        c->S.ymod[m] = mat[m][1];
        c->S.pois[m] = mat[m][2];
        c->S.sfrc[m] = mat[m][3];
        c->S.rfrc[m] = mat[m][4];
        c->S.dmpn[m] = mat[m][5];
        c->S.yldp[m] = mat[m][6];
        c->S.emod[m] = c->S.ymod[m] / (1 - c->S.pois[m] * c->S.pois[m]);
    }
    dem_calpara_from_material(&c->S);
    dem_refresh_inst(c);
    return DEM_OK;
}

extern "C" int dem_set_liquid(dem_ctx *c, double stension, double vol, double gapmn, double layer,
                              double theta, int ruptureMode, int forceCap) {
    if (!c) return dem_fail(DEM_ERR_ARG, "null argument");
    c->S.stension = stension;
    c->S.vol = vol;
    c->S.gapmn = gapmn;
    c->S.layer = layer;
    c->S.theta = theta;
    c->S.ruptureMode = ruptureMode;
    c->S.forceCap = forceCap;
    dem_refresh_inst(c);
    return DEM_OK;
}

extern "C" int dem_set_parameters(dem_ctx *c, double dt, double searchGap) {
    if (!c || !(dt > 0) || !(searchGap >= 0)) return dem_fail(DEM_ERR_ARG, "bad dt or SearchGap");
    if (searchGap > c->S.searchGap)
        return dem_fail(DEM_ERR_ARG, "SearchGap beyond the cell size the grid was built for");
    c->S.dt = dt;
    c->S.searchGap = searchGap;
    dem_refresh_inst(c);
    return DEM_OK;
}
