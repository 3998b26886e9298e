/*
 * dem_oracle.c -- FP64 CPU restatement of the reference DEM time step.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).  See
 * dem_oracle.h for the map from each function to the reference file:line it
 * follows.
 *
 * Parity pins (the reference ships no tests, fixtures or golden vectors, and
 * cannot be compiled here: it needs nvcc, CUDA-only headers and CUDA-8
 * thrust).  Each function is pinned by an independent check in
 * tests/test_oracle.py:
 *   orc_contact_pp / _pw  Hertz closed form below yield; continuity of the
 *                         plastic branch at the yield overlap; unloading
 *                         identities F(delta_max) = sqrt(F_H F_p) and
 *                         F(plstDfm) = 0; elastic load/unload reversibility;
 *                         |F_t| <= mu F_n; slip cap |xi| = xi_max; Mindlin
 *                         ratio at |xi| = xi_max/2; damping linear in v_n;
 *                         p-w law equals the p-p law with R* = r_i and the
 *                         wall constants.
 *   orc_fcap              against Young-Laplace solutions of the pendular
 *                         bridge (tests/golden/capillary_yl.py, pinned by
 *                         virtual work and the 2 pi R gamma cos(theta)
 *                         contact limit): within the regression's accuracy
 *                         for theta <= 40 degrees, tight enough that a 5 %
 *                         error in any of the 14 theta = 0 coefficients or
 *                         the theta terms fails (tests/test_capillary_pin.py).
 *   orc_rupture           both rules against the Young-Laplace rupture gap.
 *   orc_calc_hash/sort/cell_range  integer closed forms, a stable sort and
 *                         brute-force cell occupancy counts.
 *   orc_neighbor_find     brute-force O(N^2) pair search (set equality for
 *                         one size level, subset for several).
 *   orc_wall_list         the closed-form wall predicates.
 *   orc_hstry_copy(_pw)   surviving (i,j) pairs keep identical history
 *                         across a rebuild; new pairs start at zero.
 *   orc_update            discrete free-fall closed form; rolling decay.
 *   orc_step              Newton's third law over directed pairs; settled
 *                         floor overlap (m g / (4/3 E* sqrt r))^(2/3).
 *   orc_derive            hand-computed values of dt, nNebors, cell size,
 *                         grid, units and contact constants.
 *   orc_exp/_log/_pow     within 4 ulp of the C library over their ranges,
 *                         x^p within 5e-15 at the law's exponents, log 1 = 0
 *                         (they replace libm so that host and GPU round
 *                         alike; the GPU copies are dm_exp/dm_log/dm_pow).
 */
#include "dem_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PI 3.1415926                 /* heydem.h:15 */
#define ORC_EMPTY (-1)                   /* 0xffffffff cell sentinel */

/* ------------------------------------------------------------------ */
/* double3 arithmetic in the evaluation order of double_math.h         */
/* ------------------------------------------------------------------ */
typedef struct { double x, y, z; } v3;
static inline v3 mk(double x, double y, double z) { v3 r = {x, y, z}; return r; }
static inline v3 vadd(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 vsub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 vmul(v3 a, double s) { return mk(a.x * s, a.y * s, a.z * s); }   /* a*s */
static inline v3 smul(double s, v3 a) { return mk(s * a.x, s * a.y, s * a.z); }   /* s*a */
static inline v3 vdiv(v3 a, double s) { return mk(a.x / s, a.y / s, a.z / s); }
static inline double vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline double vlen(v3 a) { return sqrt(vdot(a, a)); }
static inline v3 vcross(v3 a, v3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline v3 ld3(const double *p) { return mk(p[0], p[1], p[2]); }
static inline void st3(double *p, v3 a) { p[0] = a.x; p[1] = a.y; p[2] = a.z; }

/* ------------------------------------------------------------------ */
/* Rounding-reproducible transcendentals.                              */
/* The reference calls libm pow/exp/log; libm results are within an ulp */
/* of the true value but differ between libraries in the last bit, and */
/* the contact law is discontinuous at the stick/slip and yield        */
/* boundaries, so an ulp can change a branch.  The oracle therefore    */
/* evaluates pow(x, 0.5) as sqrt(x), pow(x, 1.5) as x*sqrt(x), integer */
/* powers as products, and exp/log/other powers with the functions     */
/* below, built only from correctly rounded +, -, *, /, sqrt and exact */
/* scaling, so any IEEE machine reproduces them bit for bit.           */
/*   exp: Cody-Waite reduction x = k ln2 + r, |r| <= ln2/2, Taylor      */
/*        series to r^14 (truncation < 1e-17), scale by 2^k.            */
/*   log: x = m 2^e, m in [sqrt(1/2), sqrt(2)), s = (m-1)/(m+1),        */
/*        log m = 2 atanh s = 2 s sum z^n/(2n+1), z = s^2 <= 0.0295.    */
/* Both are within a few ulp of libm (tests/test_oracle.py).           */
/* ------------------------------------------------------------------ */
static const double DM_LN2_HI = 0x1.62e42feep-1;          /* ln 2 to 32 bits: k*hi exact */
static const double DM_LN2_LO = 0x1.a39ef35793c76p-33;    /* ln 2 - hi */
static const double DM_INV_LN2 = 1.4426950408889634;

double orc_exp(double x) {
    if (x != x) return x;
    if (x > 709.0) return HUGE_VAL;
    if (x < -708.0) return 0.0;
    double k = rint(x * DM_INV_LN2);
    double r = (x - k * DM_LN2_HI) - k * DM_LN2_LO;
    double p = 1.0 / 87178291200.0;           /* 1/14! */
    p = 1.0 / 6227020800.0 + r * p;           /* 1/13! */
    p = 1.0 / 479001600.0 + r * p;
    p = 1.0 / 39916800.0 + r * p;
    p = 1.0 / 3628800.0 + r * p;
    p = 1.0 / 362880.0 + r * p;
    p = 1.0 / 40320.0 + r * p;
    p = 1.0 / 5040.0 + r * p;
    p = 1.0 / 720.0 + r * p;
    p = 1.0 / 120.0 + r * p;
    p = 1.0 / 24.0 + r * p;
    p = 1.0 / 6.0 + r * p;
    p = 0.5 + r * p;
    p = 1.0 + r * p;
    p = 1.0 + r * p;
    return ldexp(p, (int)k);
}

double orc_log(double x) {
    if (!(x > 0.0)) return x == 0.0 ? -HUGE_VAL : (x - x) / 0.0;
    int e;
    double m = frexp(x, &e);
    if (m < 0.70710678118654752) {
        m = m * 2.0;
        e = e - 1;
    }
    double f = m - 1.0;
    double s = f / (2.0 + f);
    double z = s * s;
    double q = 1.0 / 25.0;
    q = 1.0 / 23.0 + z * q;
    q = 1.0 / 21.0 + z * q;
    q = 1.0 / 19.0 + z * q;
    q = 1.0 / 17.0 + z * q;
    q = 1.0 / 15.0 + z * q;
    q = 1.0 / 13.0 + z * q;
    q = 1.0 / 11.0 + z * q;
    q = 1.0 / 9.0 + z * q;
    q = 1.0 / 7.0 + z * q;
    q = 1.0 / 5.0 + z * q;
    q = 1.0 / 3.0 + z * q;
    double lm = 2.0 * s + 2.0 * s * (z * q);
    double de = (double)e;
    return de * DM_LN2_HI + (de * DM_LN2_LO + lm);
}

/* x^y for x > 0 */
double orc_pow(double x, double y) { return orc_exp(y * orc_log(x)); }

/* ------------------------------------------------------------------ */
/* state                                                               */
/* ------------------------------------------------------------------ */
typedef struct hist_arr {
    double *dispt, *nDsp, *nDsp_mx, *plstRad, *plstDfm;
    int *stage, *lqdBrg;
} hist_arr;

struct orc_state {
    orc_setup S;
    int NP, total;
    double *pos, *vel, *angv, *force, *mom, *rmom, *rad, *pmove;
    int *pLevel, *neborIdx;
    int *lqdCnt, *lqdCntOld;
    int *numNebor, *numNeborOld, *nebor, *neborOld;
    hist_arr pp, ppOld, pw, pwOld;
    int *numWall, *numWallOld, *wallList, *wallListOld;
    int *gridHash, *gridIndex, *cellStart, *cellEnd;
    int reNebld, iter, restart, feedFlag, nRebuild, overflow;
    double ctime;
};

static void *zalloc(size_t n) { return calloc(n ? n : 1, 1); }

static void hist_alloc(hist_arr *h, size_t n) {
    h->dispt = zalloc(n * 3 * sizeof(double));
    h->nDsp = zalloc(n * sizeof(double));
    h->nDsp_mx = zalloc(n * sizeof(double));
    h->plstRad = zalloc(n * sizeof(double));
    h->plstDfm = zalloc(n * sizeof(double));
    h->stage = zalloc(n * sizeof(int));
    h->lqdBrg = zalloc(n * sizeof(int));
}
static void hist_free(hist_arr *h) {
    free(h->dispt); free(h->nDsp); free(h->nDsp_mx); free(h->plstRad);
    free(h->plstDfm); free(h->stage); free(h->lqdBrg);
}
static void hist_copy_clear(hist_arr *dst, hist_arr *src, size_t n) {
    memcpy(dst->dispt, src->dispt, n * 3 * sizeof(double));
    memcpy(dst->nDsp, src->nDsp, n * sizeof(double));
    memcpy(dst->nDsp_mx, src->nDsp_mx, n * sizeof(double));
    memcpy(dst->plstRad, src->plstRad, n * sizeof(double));
    memcpy(dst->plstDfm, src->plstDfm, n * sizeof(double));
    memcpy(dst->stage, src->stage, n * sizeof(int));
    memcpy(dst->lqdBrg, src->lqdBrg, n * sizeof(int));
    memset(src->dispt, 0, n * 3 * sizeof(double));
    memset(src->nDsp, 0, n * sizeof(double));
    memset(src->nDsp_mx, 0, n * sizeof(double));
    memset(src->plstRad, 0, n * sizeof(double));
    memset(src->plstDfm, 0, n * sizeof(double));
    memset(src->stage, 0, n * sizeof(int));
    memset(src->lqdBrg, 0, n * sizeof(int));
}
static void hist_load(const hist_arr *h, size_t s, orc_hist1 *o) {
    o->dispt[0] = h->dispt[3 * s]; o->dispt[1] = h->dispt[3 * s + 1]; o->dispt[2] = h->dispt[3 * s + 2];
    o->nDsp = h->nDsp[s]; o->nDsp_mx = h->nDsp_mx[s];
    o->plstRad = h->plstRad[s]; o->plstDfm = h->plstDfm[s];
    o->stage = h->stage[s]; o->lqdBrg = h->lqdBrg[s];
}
static void hist_store(hist_arr *h, size_t s, const orc_hist1 *o) {
    h->dispt[3 * s] = o->dispt[0]; h->dispt[3 * s + 1] = o->dispt[1]; h->dispt[3 * s + 2] = o->dispt[2];
    h->nDsp[s] = o->nDsp; h->nDsp_mx[s] = o->nDsp_mx;
    h->plstRad[s] = o->plstRad; h->plstDfm[s] = o->plstDfm;
    h->stage[s] = o->stage; h->lqdBrg[s] = o->lqdBrg;
}
static void hist_copy_slot(hist_arr *dst, size_t d, const hist_arr *src, size_t s) {
    orc_hist1 t;
    hist_load(src, s, &t);
    hist_store(dst, d, &t);
}

int orc_num_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

/* ------------------------------------------------------------------ */
/* derivation: ReadData.cpp:216-375, UnitReduce.cpp:14-104,            */
/* main.cpp:328-348                                                    */
/* ------------------------------------------------------------------ */
static void calpara(orc_calpara *c, double emod, double pois, double sfrc, double ystress) {
    c->emod = emod;
    c->pois = pois;
    c->sfrc = sfrc;
    c->ystress = ystress;
    c->ydsp_t = pow((ORC_PI * c->ystress / (2.0 * c->emod)), 2.0);
    c->yforce_t = 4.0 / 3.0 * c->emod * pow(c->ydsp_t, 1.5);
    c->ystiff_t = ORC_PI * c->ystress;
}

int orc_derive(const orc_input *in, const double *dia, orc_setup *S,
               double *rad, int *pLevel, int *neborIdx) {
    memset(S, 0, sizeof(*S));
    const int NP = in->NP, nL = in->nLevels;
    if (NP <= 0 || nL <= 0 || nL > ORC_MAX_LEVELS) return -1;
    const double g = in->gravity > 0 ? in->gravity : (double)9.8f;

    for (int m = 0; m < 2; ++m) {
        S->density[m] = in->mat[m][0];
        S->ymod[m] = in->mat[m][1];
        S->pois[m] = in->mat[m][2];
        S->sfrc[m] = in->mat[m][3];
        S->rfrc[m] = in->mat[m][4];
        S->dmpn[m] = in->mat[m][5];
        S->yldp[m] = in->mat[m][6];
        S->emod[m] = S->ymod[m] / (1 - S->pois[m] * S->pois[m]);   /* ReadData.cpp:165 */
    }
    S->stension = in->stension; S->vol = in->vol; S->gapmn = in->gapmn;
    S->cirX = in->cirX; S->cirY = in->cirY; S->Bw = in->Bw; S->R = in->R;
    S->Tw = in->Bw + in->H;                                        /* ReadData.cpp:129 */
    S->topv = 0.0; S->btmv = 0.0;                                  /* never set in the reference */
    S->feedLmin = in->feedLmin; S->feedLmax = in->feedLmax; S->tPacking = in->tPacking;
    S->forceCap = in->forceCap; S->loadStage = in->loadStage;
    S->theta = in->theta; S->ruptureMode = in->ruptureMode;
    S->gridClampQuirk = in->gridClampQuirk;
    S->nLevels = nL;

    /* diameters (ReadData.cpp:216-235) */
    for (int i = 0; i < NP; ++i) rad[i] = (dia ? dia[i] : in->realSize) / 2;

    /* time step (ReadData.cpp:243-263) */
    double pwemod = 2.0 * S->emod[0] * S->emod[0] / (S->emod[0] + S->emod[0]);
    double rmin = rad[0], rmax = rad[0];
    for (int i = 1; i < NP; ++i) {
        if (rad[i] <= rmin) rmin = rad[i];
        if (rad[i] >= rmax) rmax = rad[i];
    }
    double sizeMn = 2.0 * rmin;
    double stiffk = 1.333 * pwemod * sqrt(sizeMn * 0.25);
    double massMn = S->density[0] * ORC_PI * pow(sizeMn, 3.0) / 6.0;
    double dt = in->dtFactor * ORC_PI * sqrt(0.5 * massMn / stiffk) * pow(10.0, -0.2);
    if (in->dtOverride > 0) dt = in->dtOverride;

    /* size levels (ReadData.cpp:273-307) */
    if (in->sizeLevelInput) {
        for (int i = 0; i < nL; ++i) S->levelSize[i] = in->levelDia[i] / 2;
    } else {
        double lvl_gap = (rmax - rmin) / nL;
        for (int i = 0; i < nL; ++i) S->levelSize[i] = lvl_gap * (i + 1) + rmin;
    }
    for (int i = 0; i < nL; ++i) {
        double Sratio = S->levelSize[i] / rmin;
        S->levelNNebors[i] = (int)ceil(1.2 * (pow((Sratio + 2 + in->cutGap), 3) - pow(Sratio, 3)));
        S->levelNPars[i] = 0;
    }
    /* level membership (ReadData.cpp:311-334); unassigned radii are an error */
    for (int i = 0; i < NP; ++i) {
        pLevel[i] = -1;
        if (rad[i] <= S->levelSize[0]) {
            pLevel[i] = 0;
            neborIdx[i] = S->levelNPars[0]++;
        } else {
            for (int L = 1; L < nL; ++L)
                if (rad[i] > S->levelSize[L - 1] && rad[i] <= S->levelSize[L]) {
                    pLevel[i] = L;
                    neborIdx[i] = S->levelNPars[L]++;
                }
        }
        if (pLevel[i] < 0) return -2;
    }
    /* domain (ReadData.cpp:136-137, 350-357) */
    double ex = in->expandSize * rmax;
    S->origin[0] = (S->cirX - S->R) - ex;
    S->origin[1] = (S->cirY - S->R) - ex;
    S->origin[2] = S->Bw - ex;
    S->endPoint[0] = (S->cirX + S->R) + ex;
    S->endPoint[1] = (S->cirY + S->R) + ex;
    S->endPoint[2] = S->Tw + ex;
    /* skin and cells (ReadData.cpp:363-375) */
    double cutGap = in->cutGap * (rmin / rmax);
    for (int i = 0; i < nL; ++i) {
        S->cellSize[i] = (1.0 + cutGap * 1.4) * 2.0 * S->levelSize[i];
        for (int d = 0; d < 3; ++d)
            S->gridSize[i][d] = (int)ceil((S->endPoint[d] - S->origin[d]) / S->cellSize[i]);
        S->numCells[i] = S->gridSize[i][0] * S->gridSize[i][1] * S->gridSize[i][2];
    }

    /* reduced units (UnitReduce.cpp:14-104) */
    const double L0 = in->realSize;
    S->rlunit = L0;
    S->rmunit = S->density[0] * ORC_PI * pow(L0, 3.0) / 6.0;
    S->rfunit = S->rmunit * g;
    S->rsunit = S->rfunit / pow(L0, 2.0);
    S->rtunit = pow(L0 / g, 0.5);
    S->rvunit = S->rlunit / S->rtunit;
    S->reunit = S->rmunit * g * S->rlunit;
    S->cirX /= S->rlunit; S->cirY /= S->rlunit;
    S->Tw /= S->rlunit; S->Bw /= S->rlunit; S->R /= S->rlunit;
    for (int i = 0; i < nL; ++i) S->levelSize[i] /= S->rlunit;
    S->feedLmax /= S->rlunit; S->feedLmin /= S->rlunit;
    S->tPacking /= S->rtunit;
    for (int i = 0; i < NP; ++i) rad[i] /= S->rlunit;
    for (int m = 0; m < 2; ++m) {
        S->emod[m] = S->emod[m] * pow(S->rlunit, 2.0) / S->rfunit;
        S->yldp[m] = S->yldp[m] * pow(S->rlunit, 2.0) / S->rfunit;
        S->dmpn[m] = S->dmpn[m] / S->rtunit;
    }
    S->gapmn = S->gapmn / L0;
    S->layer = 0.5 * (pow((1.0 + S->vol), 1.0 / 3.0) - 1.0);
    S->vol = S->vol * ORC_PI / 6.0;
    S->stension = S->stension * S->rlunit / S->rfunit;
    S->brknmx = 0.99 * pow((S->vol / 6.0), 0.34);
    S->dt = dt / S->rtunit;
    for (int d = 0; d < 3; ++d) {
        S->origin[d] /= S->rlunit;
        S->endPoint[d] /= S->rlunit;
    }
    for (int i = 0; i < nL; ++i) S->cellSize[i] /= S->rlunit;
    S->searchGap = cutGap;
    S->radMin = rmin / S->rlunit;
    S->radMax = rmax / S->rlunit;

    /* contact constants (main.cpp:328-348) */
    calpara(&S->pp, (S->emod[0] * S->emod[0]) / (S->emod[0] + S->emod[0]),
            2.0 * (S->pois[0] * S->pois[0]) / (S->pois[0] + S->pois[0]), S->sfrc[0],
            S->yldp[0] * S->yldp[0] / (S->yldp[0] + S->yldp[0]));
    calpara(&S->pw, S->emod[0], 2.0 * (S->pois[0] * S->pois[1]) / (S->pois[0] + S->pois[1]),
            pow(S->sfrc[1] * S->sfrc[1], 0.5), S->yldp[0]);

    /* list sizing (main.cpp:275-279) and cell offsets (main.cpp:373-384) */
    int acc = 0, cells = 0;
    for (int i = 0; i < nL; ++i) {
        S->levelNebSId[i] = acc;
        acc += S->levelNPars[i] * S->levelNNebors[i];
        cells += S->numCells[i];
        S->ncellIdx[i] = cells;
    }
    S->nPPCnt = acc;
    S->sysNCells = cells;
    return 0;
}

/* ------------------------------------------------------------------ */
/* contact laws                                                        */
/* ------------------------------------------------------------------ */

/* load/unload/reload verdict and normal force, shared by the p-p law
 * (kernel.cu:104-144) and the p-w law (kernel.cu:250-290).  The two differ
 * only in the yield comparison: <= for p-p (kernel.cu:120), < for p-w
 * (kernel.cu:267). */
static double normal_law(const orc_calpara *c, orc_hist1 *h, double nrmDsp,
                         double ijRad, int yield_le) {
    double fN = 0.0;
    if (nrmDsp < h->plstDfm)
        h->stage = 1;
    else if (nrmDsp >= h->plstDfm && nrmDsp < h->nDsp)
        h->stage = 3;
    else if (nrmDsp >= h->nDsp && nrmDsp < h->nDsp_mx)
        h->stage = 4;
    else
        h->stage = 2;

    if (h->stage == 1) {
        fN = 0.0;
    } else if (h->stage == 2) {
        int elastic = yield_le ? (nrmDsp <= c->ydsp_t * ijRad) : (nrmDsp < (c->ydsp_t * ijRad));
        if (elastic) {
            fN = 4.0 / 3.0 * c->emod * sqrt(ijRad * nrmDsp) * nrmDsp;
            h->plstRad = ijRad;
            h->plstDfm = 0.0;
        } else {
            fN = (c->yforce_t * (ijRad * ijRad)) + (c->ystiff_t * ijRad) * (nrmDsp - (c->ydsp_t * ijRad));
            h->plstRad = ijRad * (4.0 / 3.0 * c->emod * sqrt(ijRad * nrmDsp) * nrmDsp) / fN;
            h->plstDfm = nrmDsp - orc_pow((fN / (4.0 / 3.0 * c->emod * sqrt(ijRad))), 2.0 / 3.0);
        }
        h->nDsp_mx = nrmDsp;
    } else {
        double u = nrmDsp - h->plstDfm;
        fN = 4.0 / 3.0 * c->emod * sqrt(h->plstRad) * (u * sqrt(u));
    }
    h->nDsp = nrmDsp;
    return fN;
}

/* tangential spring with history (kernel.cu:172-217 p-p, 306-350 p-w).
 * tol_rot/tol_ratio/tol_slip: the p-p law uses 1e-7 for the rotation, the
 * Mindlin ratio and the slip cap (kernel.cu:178, 193, 199); the p-w law uses
 * 1e-10 throughout (kernel.cu:310, 323, 329). */
static v3 tangential_law(const orc_calpara *c, orc_hist1 *h, double fN, double nrmDsp,
                         double ijRad, v3 n, v3 vt, double dt, double tol_rot,
                         double tol_ratio, double tol_slip) {
    v3 old = ld3(h->dispt);
    v3 d = vcross(smul(-1.0, n), vcross(n, old));
    if (vlen(d) > tol_rot)
        d = vadd(vdiv(vmul(d, vlen(old)), vlen(d)), vmul(vt, dt));
    else
        d = vadd(d, vmul(vt, dt));

    double ijDsmx = 3.0 * c->sfrc * fN /
                    (16.0 * (1 - c->pois) * c->emod / (4.0 - 2.0 * c->pois) * sqrt(ijRad * nrmDsp));
    double ratio = 1.0;
    if (vlen(d) < ijDsmx) {
        if (ijDsmx > tol_ratio) {
            double u = 1.0 - vlen(d) / ijDsmx;
            ratio = 1.0 - u * sqrt(u);
        }
    } else {
        ratio = 1.0;
        if (vlen(d) > tol_slip)
            d = vdiv(smul(ijDsmx, d), vlen(d));
        else
            d = smul(ijDsmx, d);
    }
    st3(h->dispt, d);

    v3 ft = mk(0.0, 0.0, 0.0);
    if (vlen(d) < ijDsmx) {
        if (vlen(d) > 1.0e-10) ft = vdiv(smul(-1.0 * c->sfrc * fN * ratio, d), vlen(d));
    } else {
        if (vlen(vt) > 1.0e-10) ft = vdiv(smul(-1.0 * c->sfrc * fN * ratio, vt), vlen(vt));
    }
    return ft;
}

void orc_contact_pp(const orc_setup *S, orc_hist1 *h, double nrmDsp,
                    const double unitVec[3], double iRad, const double iVel[3],
                    const double iAngv[3], double jRad, const double jVel[3],
                    const double jAngv[3], int stage, double F[3], double M[3],
                    double *Mr) {
    const orc_calpara *c = &S->pp;
    v3 n = ld3(unitVec);
    double ijRad = iRad * jRad / (iRad + jRad);
    double fN = normal_law(c, h, nrmDsp, ijRad, 1);

    /* damping (kernel.cu:159-165) */
    v3 cpv = vsub(vsub(ld3(iVel), ld3(jVel)),
                  vcross(vadd(vmul(ld3(iAngv), iRad), vmul(ld3(jAngv), jRad)), n));
    double cpvn = vdot(cpv, n);
    double fD = 0.0;
    if (stage == 1)
        fD = -2.0 * sqrt(S->dmpn[0] * S->dmpn[0]) * c->emod * sqrt(ijRad * nrmDsp) * cpvn;

    v3 vt = vsub(cpv, smul(cpvn, n));
    v3 ft = tangential_law(c, h, fN, nrmDsp, ijRad, n, vt, S->dt, 1.0e-7, 1e-7, 1.0e-7);

    /* kernel.cu:223-226 */
    st3(F, vadd(smul(fN + fD, n), ft));
    double arm = iRad - (nrmDsp) / 2.0;
    st3(M, vcross(smul(-arm, n), ft));
    *Mr = 0.5 * (S->rfrc[0] + S->rfrc[0]) * fN * iRad;
}

void orc_contact_pw(const orc_setup *S, orc_hist1 *h, double nrmDsp,
                    const double unitVec[3], double iRad, const double iVel[3],
                    const double iAngv[3], const double jVel[3], int stage,
                    double F[3], double M[3], double *Mr) {
    const orc_calpara *c = &S->pw;
    v3 n = ld3(unitVec);
    double ijRad = iRad;
    double fN = normal_law(c, h, nrmDsp, ijRad, 0);

    /* damping (kernel.cu:296-302) */
    v3 cpv = vsub(vsub(ld3(iVel), ld3(jVel)), vcross(vmul(ld3(iAngv), iRad), n));
    double cpvn = vdot(cpv, n);
    double fD = 0.0;
    if (stage == 1 || stage == 7)
        fD = -2.0 * sqrt(S->dmpn[0] * S->dmpn[1]) * c->emod * sqrt(ijRad * nrmDsp) * cpvn;

    v3 vt = vsub(cpv, smul(cpvn, n));
    v3 ft = tangential_law(c, h, fN, nrmDsp, ijRad, n, vt, S->dt, 1.0e-10, 1.0e-10, 1.0e-10);

    /* kernel.cu:354-357 */
    st3(F, vadd(smul(fN + fD, n), ft));
    double arm = iRad - (nrmDsp) / 2;
    st3(M, vcross(smul(-arm, n), ft));
    *Mr = 0.5 * (S->rfrc[0] + S->rfrc[0]) * fN * iRad;
}

/* Willett et al. (2000) regression for the bridge force between two spheres,
 * F = 2 pi R gamma exp(f1 - f2 exp(f3 ln S+ + f4 ln^2 S+)) with every
 * coefficient of f1..f4 a quadratic in the contact angle (radians):
 * a = c0 + theta (c1 + theta c2).  Rows: f1 (ln V)^0..3, f2 (ln V)^0..2,
 * f3 (ln V)^0..3, f4 (ln V)^0..2.  Column c0 (theta = 0) is exactly the
 * coefficients of kernel.cu:73-76; c1 and c2 are the published fit in the
 * contact angle. */
const double ORC_WILLETT[14][3] = {
    {-0.44507, 0.050832, -1.1466},       {-0.1119, -0.000411, -0.1490},
    {-0.012101, -0.0036456, -0.01255},   {-0.0005, -0.0003505, -0.00029076},
    {1.9222, -0.57473, -1.2918},         {-0.0668, -0.1201, -0.22574},
    {-0.0013375, -0.0068988, -0.01137},
    {1.268, -0.01396, -0.23566},         {0.198, 0.092, -0.06418},
    {0.02232, 0.02238, -0.009853},       {0.0008585, 0.001318, -0.00053},
    {-0.010703, 0.073776, -0.34742},     {0.03345, 0.04543, -0.09056},
    {0.0018574, 0.004456, -0.006257},
};

/* fCap (kernel.cu:65-81) with the coefficient table as an argument, so that
 * tests can show which coefficient errors the Young-Laplace check detects. */
double orc_fcap_coef(const double coef[14][3], const orc_setup *S, double Splus, double ijVol,
                     double effectR) {
    const double th = S->theta;
    double a[14];
    for (int k = 0; k < 14; ++k) a[k] = coef[k][0] + th * (coef[k][1] + th * coef[k][2]);
    Splus = fmax(S->gapmn, Splus);
    Splus = Splus / sqrt(ijVol);
    double LV = orc_log(ijVol);
    double L2 = LV * LV, L3 = L2 * LV;
    double f1 = a[0] + a[1] * LV + a[2] * L2 + a[3] * L3;
    double f2 = a[4] + a[5] * LV + a[6] * L2;
    double f3 = a[7] + a[8] * LV + a[9] * L2 + a[10] * L3;
    double f4 = a[11] + a[12] * LV + a[13] * L2;
    double lS = orc_log(Splus);
    return 2 * ORC_PI * effectR * S->stension * orc_exp(f1 - f2 * orc_exp(f3 * lS + f4 * (lS * lS)));
}

double orc_fcap(const orc_setup *S, double Splus, double ijVol, double effectR) {
    return orc_fcap_coef(ORC_WILLETT, S, Splus, ijVol, effectR);
}

/* bridge rupture distance: kernel.cu:522, or Lian et al. (1993)
 * (1 + theta/2) V^(1/3) when ruptureMode = 1 */
double orc_rupture(const orc_setup *S, double ijVol) {
    if (S->ruptureMode == 1) return (1.0 + 0.5 * S->theta) * orc_pow(ijVol, 1.0 / 3.0);
    return 0.99 * orc_pow(ijVol, 0.33);
}

/* ------------------------------------------------------------------ */
/* state management                                                    */
/* ------------------------------------------------------------------ */
orc_state *orc_create(const orc_input *in, const double *dia) {
    orc_state *st = zalloc(sizeof(orc_state));
    const int NP = in->NP;
    st->NP = NP;
    st->rad = zalloc(NP * sizeof(double));
    st->pLevel = zalloc(NP * sizeof(int));
    st->neborIdx = zalloc(NP * sizeof(int));
    if (orc_derive(in, dia, &st->S, st->rad, st->pLevel, st->neborIdx) != 0) {
        free(st->rad); free(st->pLevel); free(st->neborIdx); free(st);
        return NULL;
    }
    size_t n3 = (size_t)NP * 3;
    st->pos = zalloc(n3 * sizeof(double));
    st->vel = zalloc(n3 * sizeof(double));
    st->angv = zalloc(n3 * sizeof(double));
    st->force = zalloc(n3 * sizeof(double));
    st->mom = zalloc(n3 * sizeof(double));
    st->rmom = zalloc(NP * sizeof(double));
    st->pmove = zalloc(NP * sizeof(double));
    st->lqdCnt = zalloc(NP * sizeof(int));
    st->lqdCntOld = zalloc(NP * sizeof(int));
    size_t nPP = (size_t)st->S.nPPCnt, nW = (size_t)NP * ORC_WALLS;
    st->numNebor = zalloc(NP * sizeof(int));
    st->numNeborOld = zalloc(NP * sizeof(int));
    st->nebor = zalloc(nPP * sizeof(int));
    st->neborOld = zalloc(nPP * sizeof(int));
    hist_alloc(&st->pp, nPP);
    hist_alloc(&st->ppOld, nPP);
    hist_alloc(&st->pw, nW);
    hist_alloc(&st->pwOld, nW);
    st->numWall = zalloc(NP * sizeof(int));
    st->numWallOld = zalloc(NP * sizeof(int));
    st->wallList = zalloc(nW * sizeof(int));
    st->wallListOld = zalloc(nW * sizeof(int));
    st->gridHash = zalloc(NP * sizeof(int));
    st->gridIndex = zalloc(NP * sizeof(int));
    st->cellStart = zalloc((size_t)st->S.sysNCells * sizeof(int));
    st->cellEnd = zalloc((size_t)st->S.sysNCells * sizeof(int));
    st->feedFlag = 1; /* the first step builds the lists */
    return st;
}

void orc_free(orc_state *st) {
    if (!st) return;
    free(st->pos); free(st->vel); free(st->angv); free(st->force); free(st->mom);
    free(st->rmom); free(st->rad); free(st->pmove); free(st->pLevel); free(st->neborIdx);
    free(st->lqdCnt); free(st->lqdCntOld); free(st->numNebor); free(st->numNeborOld);
    free(st->nebor); free(st->neborOld);
    hist_free(&st->pp); hist_free(&st->ppOld); hist_free(&st->pw); hist_free(&st->pwOld);
    free(st->numWall); free(st->numWallOld); free(st->wallList); free(st->wallListOld);
    free(st->gridHash); free(st->gridIndex); free(st->cellStart); free(st->cellEnd);
    free(st);
}

const orc_setup *orc_get_setup(const orc_state *st) { return &st->S; }

void orc_set_particles(orc_state *st, int total, const double *pos,
                       const double *vel, const double *angv) {
    st->total = total;
    memcpy(st->pos, pos, (size_t)total * 3 * sizeof(double));
    if (vel) memcpy(st->vel, vel, (size_t)total * 3 * sizeof(double));
    if (angv) memcpy(st->angv, angv, (size_t)total * 3 * sizeof(double));
    st->feedFlag = 1;
}
/* main.cpp:439-448: ParticleFeed appends positions, FeedFlag forces a rebuild */
void orc_feed(orc_state *st, int n, const double *pos) {
    memcpy(&st->pos[3 * (size_t)st->total], pos, (size_t)n * 3 * sizeof(double));
    st->total += n;
    st->feedFlag = 1;
}
void orc_get_particles(const orc_state *st, double *pos, double *vel, double *angv) {
    size_t n = (size_t)st->total * 3 * sizeof(double);
    if (pos) memcpy(pos, st->pos, n);
    if (vel) memcpy(vel, st->vel, n);
    if (angv) memcpy(angv, st->angv, n);
}
void orc_get_forces(const orc_state *st, double *F, double *M, double *Mr) {
    size_t n = (size_t)st->total;
    if (F) memcpy(F, st->force, n * 3 * sizeof(double));
    if (M) memcpy(M, st->mom, n * 3 * sizeof(double));
    if (Mr) memcpy(Mr, st->rmom, n * sizeof(double));
}
void orc_get_pmove(const orc_state *st, double *pmove) {
    memcpy(pmove, st->pmove, (size_t)st->total * sizeof(double));
}
void orc_get_lqdcnt(const orc_state *st, int *lqdCnt) {
    memcpy(lqdCnt, st->lqdCnt, (size_t)st->total * sizeof(int));
}
int orc_total(const orc_state *st) { return st->total; }
int orc_iter(const orc_state *st) { return st->iter; }
double orc_ctime(const orc_state *st) { return st->ctime; }
double orc_tw(const orc_state *st) { return st->S.Tw; }
int orc_n_rebuild(const orc_state *st) { return st->nRebuild; }
int orc_overflow(const orc_state *st) { return st->overflow; }

static inline size_t slot_of(const orc_state *st, int i, int k) {
    int L = st->pLevel[i];
    return (size_t)st->S.levelNebSId[L] + (size_t)k * st->S.levelNPars[L] + st->neborIdx[i];
}

int orc_get_nebor(const orc_state *st, int *num, int *list, int maxK) {
    int bad = 0;
    for (int i = 0; i < st->total; ++i) {
        num[i] = st->numNebor[i];
        for (int k = 0; k < maxK; ++k)
            list[(size_t)i * maxK + k] = k < st->numNebor[i] ? st->nebor[slot_of(st, i, k)] : -1;
        if (st->numNebor[i] > maxK) bad = 1;
    }
    return bad ? -1 : 0;
}

int orc_get_hist_pp(const orc_state *st, int maxK, double *dispt, double *nDsp,
                    double *nDsp_mx, double *plstRad, double *plstDfm,
                    int *stage, int *lqdBrg) {
    for (int i = 0; i < st->total; ++i)
        for (int k = 0; k < maxK; ++k) {
            size_t o = (size_t)i * maxK + k;
            orc_hist1 h;
            memset(&h, 0, sizeof(h));
            if (k < st->numNebor[i]) hist_load(&st->pp, slot_of(st, i, k), &h);
            dispt[3 * o] = h.dispt[0]; dispt[3 * o + 1] = h.dispt[1]; dispt[3 * o + 2] = h.dispt[2];
            nDsp[o] = h.nDsp; nDsp_mx[o] = h.nDsp_mx; plstRad[o] = h.plstRad;
            plstDfm[o] = h.plstDfm; stage[o] = h.stage; lqdBrg[o] = h.lqdBrg;
        }
    return 0;
}

/* RESTART (main.cpp:57-85, 473-519): lists and history saved by the getters
 * (lists as original ids in slot order, num[total], list[total][maxK]; wall
 * lists [total][3]) become the current ones, and the next step rebuilds and
 * moves every surviving pair's history to its new slot.  Returns -1 if a
 * list is longer than its size level allows. */
int orc_set_contacts(orc_state *st, int maxK, const int *num, const int *list,
                     const double *dispt, const double *nDsp, const double *nDsp_mx,
                     const double *plstRad, const double *plstDfm, const int *stage,
                     const int *lqdBrg, const int *wnum, const int *wlist,
                     const double *wdispt, const double *wnDsp, const double *wnDsp_mx,
                     const double *wplstRad, const double *wplstDfm, const int *wstage,
                     const int *wlqdBrg) {
    for (int i = 0; i < st->total; ++i)
        if (num[i] > st->S.levelNNebors[st->pLevel[i]] || num[i] > maxK || wnum[i] > ORC_WALLS)
            return -1;
    for (int i = 0; i < st->total; ++i) {
        st->numNebor[i] = num[i];
        for (int k = 0; k < num[i]; ++k) {
            size_t o = (size_t)i * maxK + k, sl = slot_of(st, i, k);
            orc_hist1 h;
            h.dispt[0] = dispt[3 * o]; h.dispt[1] = dispt[3 * o + 1]; h.dispt[2] = dispt[3 * o + 2];
            h.nDsp = nDsp[o]; h.nDsp_mx = nDsp_mx[o]; h.plstRad = plstRad[o];
            h.plstDfm = plstDfm[o]; h.stage = stage[o]; h.lqdBrg = lqdBrg[o];
            st->nebor[sl] = list[o];
            hist_store(&st->pp, sl, &h);
        }
        st->numWall[i] = wnum[i];
        for (int k = 0; k < wnum[i]; ++k) {
            size_t o = (size_t)i * ORC_WALLS + k;
            orc_hist1 h;
            h.dispt[0] = wdispt[3 * o]; h.dispt[1] = wdispt[3 * o + 1]; h.dispt[2] = wdispt[3 * o + 2];
            h.nDsp = wnDsp[o]; h.nDsp_mx = wnDsp_mx[o]; h.plstRad = wplstRad[o];
            h.plstDfm = wplstDfm[o]; h.stage = wstage[o]; h.lqdBrg = wlqdBrg[o];
            st->wallList[o] = wlist[o];
            hist_store(&st->pw, (size_t)k * st->NP + i, &h);
        }
    }
    st->restart = 1;
    return 0;
}

/* the bridge counts of the last step (lqdCnt, which VarInit makes
 * lqdCnt_Old), as a restart restores them */
void orc_set_lqdcnt(orc_state *st, const int *lqdCnt) {
    memcpy(st->lqdCnt, lqdCnt, (size_t)st->total * sizeof(int));
}

void orc_get_walls(const orc_state *st, int *num, int *list) {
    for (int i = 0; i < st->total; ++i) {
        num[i] = st->numWall[i];
        for (int k = 0; k < ORC_WALLS; ++k)
            list[i * ORC_WALLS + k] = k < st->numWall[i] ? st->wallList[i * ORC_WALLS + k] : -1;
    }
}

void orc_get_hist_pw(const orc_state *st, double *dispt, double *nDsp,
                     double *nDsp_mx, double *plstRad, double *plstDfm,
                     int *stage, int *lqdBrg) {
    for (int i = 0; i < st->total; ++i)
        for (int k = 0; k < ORC_WALLS; ++k) {
            size_t o = (size_t)i * ORC_WALLS + k;
            orc_hist1 h;
            memset(&h, 0, sizeof(h));
            if (k < st->numWall[i]) hist_load(&st->pw, (size_t)k * st->NP + i, &h);
            dispt[3 * o] = h.dispt[0]; dispt[3 * o + 1] = h.dispt[1]; dispt[3 * o + 2] = h.dispt[2];
            nDsp[o] = h.nDsp; nDsp_mx[o] = h.nDsp_mx; plstRad[o] = h.plstRad;
            plstDfm[o] = h.plstDfm; stage[o] = h.stage; lqdBrg[o] = h.lqdBrg;
        }
}

void orc_get_hash(const orc_state *st, int *hash, int *index) {
    memcpy(hash, st->gridHash, (size_t)st->total * sizeof(int));
    memcpy(index, st->gridIndex, (size_t)st->total * sizeof(int));
}
void orc_get_cells(const orc_state *st, int *cellStart, int *cellEnd) {
    memcpy(cellStart, st->cellStart, (size_t)st->S.sysNCells * sizeof(int));
    memcpy(cellEnd, st->cellEnd, (size_t)st->S.sysNCells * sizeof(int));
}

/* ------------------------------------------------------------------ */
/* cell hash, sort, cell range, neighbour search                       */
/* ------------------------------------------------------------------ */
/* kernel.cu:617-641; the upper clamp is gridSize in the reference
 * (gridClampQuirk = 1) and gridSize-1 otherwise.  In-domain particles get
 * the same cell either way. */
static void grid_pos(const orc_setup *S, const double *p, int L, int g[3]) {
    for (int d = 0; d < 3; ++d) {
        int c = (int)floor((p[d] - S->origin[d]) / S->cellSize[L]);
        int hi = S->gridClampQuirk ? S->gridSize[L][d] : S->gridSize[L][d] - 1;
        if (c > hi) c = hi;
        if (c < 0) c = 0;
        g[d] = c;
    }
}
static inline int grid_hash(const orc_setup *S, const int g[3], int L) {
    return (g[2] * S->gridSize[L][1] * S->gridSize[L][0]) + (g[1] * S->gridSize[L][0]) + g[0];
}

void orc_calc_hash(orc_state *st) {
    const orc_setup *S = &st->S;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < st->total; ++i) {
        int L = st->pLevel[i], g[3];
        grid_pos(S, &st->pos[3 * (size_t)i], L, g);
        st->gridHash[i] = grid_hash(S, g, L) + (L != 0 ? S->ncellIdx[L - 1] : 0);
        st->gridIndex[i] = i;
    }
}

/* stable key/value sort (ParticleSystem.cu:61-69): a counting sort over the
 * hash keeps equal keys in index order. */
void orc_sort(orc_state *st) {
    const int n = st->total, nc = st->S.sysNCells + 1;
    int *count = zalloc(((size_t)nc + 1) * sizeof(int));
    int *key = malloc((size_t)(n ? n : 1) * sizeof(int));
    int *val = malloc((size_t)(n ? n : 1) * sizeof(int));
    for (int i = 0; i < n; ++i) count[st->gridHash[i] + 1]++;
    for (int c = 0; c < nc; ++c) count[c + 1] += count[c];
    for (int i = 0; i < n; ++i) {
        int p = count[st->gridHash[i]]++;
        key[p] = st->gridHash[i];
        val[p] = st->gridIndex[i];
    }
    memcpy(st->gridHash, key, (size_t)n * sizeof(int));
    memcpy(st->gridIndex, val, (size_t)n * sizeof(int));
    free(count); free(key); free(val);
}

/* ParticleSystem.cu:90-101 + kernel.cu:682-724 */
void orc_cell_range(orc_state *st) {
    for (int c = 0; c < st->S.sysNCells; ++c) st->cellStart[c] = st->cellEnd[c] = ORC_EMPTY;
    const int n = st->total;
    for (int k = 0; k < n; ++k) {
        int h = st->gridHash[k];
        if (k == 0 || h != st->gridHash[k - 1]) {
            st->cellStart[h] = k;
            if (k > 0) st->cellEnd[st->gridHash[k - 1]] = k;
        }
        if (k == n - 1) st->cellEnd[h] = k + 1;
    }
}

/* kernel.cu:729-855 */
void orc_neighbor_find(orc_state *st) {
    const orc_setup *S = &st->S;
    int ovf = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(| : ovf)
    for (int i = 0; i < st->total; ++i) {
        const double *op = &st->pos[3 * (size_t)i];
        int iL = st->pLevel[i];
        int cap = S->levelNNebors[iL];
        int num = 0;
        for (int L = 0; L < S->nLevels; ++L) {
            int g[3];
            grid_pos(S, op, L, g);
            int off = (int)ceil(S->levelSize[iL] / S->levelSize[L]);
            int cellSum = L != 0 ? S->ncellIdx[L - 1] : 0;
            const int gx = S->gridSize[L][0], gy = S->gridSize[L][1], gz = S->gridSize[L][2];
            for (int z = -off; z <= off; ++z)
                for (int y = -off; y <= off; ++y) {
                    int s[3] = {g[0] - off, g[1] + y, g[2] + z};
                    int e[3] = {g[0] + off, g[1] + y, g[2] + z};
                    if (s[0] < 0) s[0] = 0;
                    if (e[0] > gx - 1) e[0] = gx - 1;
                    if (s[1] < 0 || s[2] < 0 || s[1] > gy - 1 || s[2] > gz - 1) continue;
                    while (s[0] != e[0]) {
                        if (st->cellStart[grid_hash(S, s, L) + cellSum] == ORC_EMPTY)
                            s[0]++;
                        else
                            break;
                    }
                    int startIndex = st->cellStart[grid_hash(S, s, L) + cellSum];
                    while (s[0] != e[0]) {
                        if (st->cellEnd[grid_hash(S, e, L) + cellSum] == ORC_EMPTY)
                            e[0]--;
                        else
                            break;
                    }
                    int endIndex = st->cellEnd[grid_hash(S, e, L) + cellSum];
                    if (startIndex == ORC_EMPTY || endIndex == ORC_EMPTY) continue;
                    for (int k = startIndex; k < endIndex; ++k) {
                        int r = st->gridIndex[k];
                        if (r == i) continue;
                        v3 dist = vsub(ld3(op), ld3(&st->pos[3 * (size_t)r]));
                        if (vlen(dist) < st->rad[r] + st->rad[i] + S->searchGap) {
                            if (num < cap)
                                st->nebor[slot_of(st, i, num)] = r;
                            else
                                ovf = 1;
                            num++;
                        }
                    }
                }
        }
        st->numNebor[i] = num < cap ? num : cap;
    }
    st->overflow |= ovf;
}

/* kernel.cu:1034-1073 */
void orc_wall_list(orc_state *st) {
    const orc_setup *S = &st->S;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < st->total; ++i) {
        const double *p = &st->pos[3 * (size_t)i];
        double r = st->rad[i];
        int n = 0, *row = &st->wallList[i * ORC_WALLS];
        if (p[2] >= (S->Tw - (2.0 * r + S->searchGap))) row[n++] = 0;
        if (p[2] <= ((2.0 * r + S->searchGap) + S->Bw)) row[n++] = 1;
        double ax = S->cirX - p[0], ay = S->cirY - p[1];
        double dc = sqrt(ax * ax + ay * ay);
        if (dc >= S->R - (2.0 * r + S->searchGap)) row[n++] = 2;
        st->numWall[i] = n;
    }
}

/* ParticleSystem.cu:336-375 plus the list copies of main.cpp:479-485 */
void orc_hist_save_and_clear(orc_state *st) {
    const int NP = st->NP;
    memcpy(st->numNeborOld, st->numNebor, (size_t)NP * sizeof(int));
    memcpy(st->neborOld, st->nebor, (size_t)st->S.nPPCnt * sizeof(int));
    hist_copy_clear(&st->ppOld, &st->pp, (size_t)st->S.nPPCnt);
    memcpy(st->numWallOld, st->numWall, (size_t)NP * sizeof(int));
    memcpy(st->wallListOld, st->wallList, (size_t)NP * ORC_WALLS * sizeof(int));
    hist_copy_clear(&st->pwOld, &st->pw, (size_t)NP * ORC_WALLS);
}

/* kernel.cu:869-911 */
void orc_hstry_copy(orc_state *st) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < st->total; ++i) {
        for (int a = 0; a < st->numNebor[i]; ++a) {
            size_t ine = slot_of(st, i, a);
            int idx = st->nebor[ine];
            for (int b = 0; b < st->numNeborOld[i]; ++b) {
                size_t jne = slot_of(st, i, b);
                if (idx == st->neborOld[jne]) hist_copy_slot(&st->pp, ine, &st->ppOld, jne);
            }
        }
    }
}

/* kernel.cu:918-950 */
void orc_hstry_copy_pw(orc_state *st) {
    const int NP = st->NP;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < st->total; ++i)
        for (int a = 0; a < st->numWall[i]; ++a) {
            int idx = st->wallList[i * ORC_WALLS + a];
            for (int b = 0; b < st->numWallOld[i]; ++b)
                if (idx == st->wallListOld[i * ORC_WALLS + b])
                    hist_copy_slot(&st->pw, (size_t)a * NP + i, &st->pwOld, (size_t)b * NP + i);
        }
}

/* Allocation.cpp:308-317 */
void orc_var_init(orc_state *st) {
    const size_t NP = (size_t)st->NP;
    memcpy(st->lqdCntOld, st->lqdCnt, NP * sizeof(int));
    memset(st->lqdCnt, 0, NP * sizeof(int));
    memset(st->force, 0, NP * 3 * sizeof(double));
    memset(st->mom, 0, NP * 3 * sizeof(double));
    memset(st->rmom, 0, NP * sizeof(double));
}

/* ------------------------------------------------------------------ */
/* forces and integration                                              */
/* ------------------------------------------------------------------ */
/* kernel.cu:413-552 */
void orc_pp_force(orc_state *st) {
    const orc_setup *S = &st->S;
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < st->total; ++i) {
        v3 ip = ld3(&st->pos[3 * (size_t)i]);
        const double *iv = &st->vel[3 * (size_t)i], *iw = &st->angv[3 * (size_t)i];
        double ir = st->rad[i];
        int inl = st->lqdCntOld[i];
        if (inl < 12) inl = 12;
        double dv = st->rad[i] * 2;
        double iVol = S->vol * (dv * dv * dv) / inl;
        v3 F = ld3(&st->force[3 * (size_t)i]);
        v3 M = ld3(&st->mom[3 * (size_t)i]);
        double Mr = st->rmom[i];
        double di = 2.0 * st->rad[i];
        F.z = F.z - di * di * di; /* gravity, kernel.cu:439 */
        for (int k = 0; k < st->numNebor[i]; ++k) {
            size_t ine = slot_of(st, i, k);
            int j = st->nebor[ine];
            v3 jp = ld3(&st->pos[3 * (size_t)j]);
            double jr = st->rad[j];
            v3 d = vsub(ip, jp);
            v3 n = vdiv(d, vlen(d));
            double gap = vlen(d) - ir - jr;
            if (gap < 0.0) {
                orc_hist1 h;
                hist_load(&st->pp, ine, &h);
                double dF[3], dM[3], dMr, nn[3] = {n.x, n.y, n.z};
                orc_contact_pp(S, &h, -gap, nn, ir, iv, iw, jr, &st->vel[3 * (size_t)j],
                               &st->angv[3 * (size_t)j], S->loadStage, dF, dM, &dMr);
                hist_store(&st->pp, ine, &h);
                F = vadd(F, ld3(dF));
                M = vadd(M, ld3(dM));
                Mr = Mr + dMr;
            }
            if (S->forceCap == 1) {
                int jnl = st->lqdCntOld[j];
                if (jnl < 12) jnl = 12;
                double dj = jr * 2;
                double jVol = S->vol * (dj * dj * dj) / jnl;
                double ijVol = iVol + jVol;
                double rupture = orc_rupture(S, ijVol);
                if (gap < 2 * S->layer * (ir + jr))
                    st->pp.lqdBrg[ine] = 1;
                else if (gap > rupture)
                    st->pp.lqdBrg[ine] = 0;
                double Splus = gap < 0 ? 0.0 : gap;
                if (st->pp.lqdBrg[ine] == 1) {
                    double effectR = 2 * ir * jr / (ir + jr);
                    ijVol = ijVol * 8.0;
                    v3 fc = smul(-1.0 * orc_fcap(S, Splus, ijVol, effectR), n);
                    F = vadd(F, fc);
                    st->lqdCnt[i] = st->lqdCnt[i] + 1;
                }
            }
        }
        st3(&st->force[3 * (size_t)i], F);
        st3(&st->mom[3 * (size_t)i], M);
        st->rmom[i] = Mr;
    }
}

/* kernel.cu:958-1023 with MultiPWForceD (kernel.cu:370-403) */
void orc_pw_force(orc_state *st) {
    orc_setup *S = &st->S;
    S->Tw = S->Tw + S->topv * S->dt;
    const int NP = st->NP;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < st->total; ++i) {
        v3 ip = ld3(&st->pos[3 * (size_t)i]);
        const double *iv = &st->vel[3 * (size_t)i], *iw = &st->angv[3 * (size_t)i];
        double ir = st->rad[i];
        v3 jp = mk(0.0, 0.0, 0.0), jv = mk(0.0, 0.0, 0.0);
        for (int k = 0; k < st->numWall[i]; ++k) {
            int w = st->wallList[i * ORC_WALLS + k];
            if (w == 0) {
                jp = mk(ip.x, ip.y, S->Tw);
                jv = mk(0.0, 0.0, S->topv);
            } else if (w == 1) {
                jp = mk(ip.x, ip.y, S->Bw);
            } else if (w == 2) {
                double ax = S->cirX - ip.x, ay = S->cirY - ip.y;
                double dc = sqrt(ax * ax + ay * ay);
                double cs = (ip.x - S->cirX) / dc, sn = (ip.y - S->cirY) / dc;
                jp.x = S->R * cs + S->cirX;
                jp.y = S->R * sn + S->cirY;
                jp.z = ip.z;
            } else {
                continue;
            }
            v3 d = vsub(ip, jp);
            v3 n = vdiv(d, vlen(d));
            double gap = vlen(d) - ir;
            if (gap < 0.0) {
                size_t s = (size_t)k * NP + i;
                orc_hist1 h;
                hist_load(&st->pw, s, &h);
                double dF[3], dM[3], dMr, nn[3] = {n.x, n.y, n.z}, jvv[3] = {jv.x, jv.y, jv.z};
                orc_contact_pw(S, &h, -gap, nn, ir, iv, iw, jvv, S->loadStage, dF, dM, &dMr);
                hist_store(&st->pw, s, &h);
                st3(&st->force[3 * (size_t)i], vadd(ld3(&st->force[3 * (size_t)i]), ld3(dF)));
                st3(&st->mom[3 * (size_t)i], vadd(ld3(&st->mom[3 * (size_t)i]), ld3(dM)));
                st->rmom[i] = st->rmom[i] + dMr;
            }
        }
    }
}

/* kernel.cu:563-606; returns the number of particles over the threshold */
int orc_update(orc_state *st) {
    const orc_setup *S = &st->S;
    const double dt = S->dt;
    int flag = 0;
#pragma omp parallel for schedule(static) reduction(+ : flag)
    for (int i = 0; i < st->total; ++i) {
        double *v = &st->vel[3 * (size_t)i], *x = &st->pos[3 * (size_t)i];
        double *w = &st->angv[3 * (size_t)i];
        const double *F = &st->force[3 * (size_t)i], *M = &st->mom[3 * (size_t)i];
        double di = 2.0 * st->rad[i], d2 = di * di;
        double mass = di * di * di;
        v[0] = v[0] + F[0] * dt / mass;
        v[1] = v[1] + F[1] * dt / mass;
        v[2] = v[2] + F[2] * dt / mass;
        st3(x, vadd(ld3(x), vmul(ld3(v), dt)));
        st->pmove[i] = st->pmove[i] + vlen(ld3(v)) * dt;
        double inert = 0.1 * (d2 * d2 * di);
        w[0] = w[0] + M[0] * dt / inert;
        w[1] = w[1] + M[1] * dt / inert;
        w[2] = w[2] + M[2] * dt / inert;
        st->rmom[i] = st->rmom[i] * dt / inert;
        double al = vlen(ld3(w));
        if (al > 1.0e-7) st3(w, vmul(ld3(w), (1.0 - fmin(1.0, st->rmom[i] / al))));
        if (st->pmove[i] > (0.3 * S->searchGap)) flag += 1;
    }
    return flag;
}

/* main.cpp:473-527 */
int orc_rebuild(orc_state *st) {
    int copy = (st->iter != 0 || st->restart == 1);
    if (copy) orc_hist_save_and_clear(st);
    memset(st->pmove, 0, (size_t)st->NP * sizeof(double));
    orc_calc_hash(st);
    orc_sort(st);
    orc_cell_range(st);
    orc_neighbor_find(st);
    orc_wall_list(st);
    if (copy) {
        orc_hstry_copy(st);
        orc_hstry_copy_pw(st);
    }
    st->feedFlag = 0;
    st->restart = 0;
    st->nRebuild++;
    st->reNebld = 0;
    return st->overflow ? -1 : 0;
}

void orc_force_rebuild(orc_state *st) { st->feedFlag = 1; }

/* main.cpp:470-564 */
int orc_step(orc_state *st) {
    if (st->reNebld > 0 || st->feedFlag == 1 || st->restart == 1)
        if (orc_rebuild(st) != 0) return -1;
    orc_var_init(st);
    orc_pp_force(st);
    orc_pw_force(st);
    st->reNebld = 0;
    st->reNebld = orc_update(st);
    st->ctime = st->ctime + st->S.dt;
    st->iter = st->iter + 1;
    return 0;
}

int orc_run(orc_state *st, int nsteps) {
    for (int s = 0; s < nsteps; ++s)
        if (orc_step(st) != 0) return -1;
    return 0;
}

static double now_s(void) {
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return t.tv_sec + 1e-9 * t.tv_nsec;
}

double orc_time_force_steps(orc_state *st, int nsteps) {
    double t0 = now_s();
    for (int s = 0; s < nsteps; ++s) {
        orc_var_init(st);
        orc_pp_force(st);
        orc_pw_force(st);
        st->reNebld = orc_update(st);
        st->ctime = st->ctime + st->S.dt;
        st->iter = st->iter + 1;
    }
    return now_s() - t0;
}
