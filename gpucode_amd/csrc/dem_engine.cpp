// dem_engine.cpp -- the step driver and state import/export of the engine.
//
// The reference's time loop (main.cpp:425-623) syncs with the device on
// every step to read the rebuild flag (main.cpp:470).  dem_step launches a
// chunk of fused steps back to back instead; a step kernel that finds a
// rebuild request from an earlier step of the chunk exits before it writes
// anything, and the host rebuilds and re-runs the skipped steps after
// reading the request once per chunk.
#include <limits.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "dem_host.h"
#include "dem_launch.h"

static const int kChunk = 64;
static const int kPairMinInterval = 64;

// shared pair records: dry packings whose slots fit the 64-bit slot masks
static bool pair_fits(const dem_ctx *c) { return c->S.forceCap != 1 && c->K <= 64; }

StepArgs dem_step_args(dem_ctx *c, bool fused) {
    StepArgs A;
    memset(&A, 0, sizeof(A));
    if (fused) {
        A.in = c->buf(c->cur, c->Lidx);
        A.out = c->buf(1 - c->cur, 1 - c->Lidx);
    } else {
        // separate kernels: in place, lqdCnt_Old = lq[1-Lidx], lqdCnt = lq[Lidx]
        A.in = c->buf(c->cur, 1 - c->Lidx);
        A.out = c->buf(c->cur, c->Lidx);
    }
    A.nbr = c->nbr[c->hcur];
    A.nnb = c->nnb[c->hcur];
    A.h = c->hp[c->hl];
    A.amask = c->amask[c->hl];
    A.actoff = c->actoff[c->hl];
    A.nmask = c->nmask;
    A.wl = c->wl[c->hcur];
    A.nwl = c->nwl[c->hcur];
    A.hw = c->hw[c->hcur];
    A.Fx = c->F[0]; A.Fy = c->F[1]; A.Fz = c->F[2];
    A.Mx = c->F[3]; A.My = c->F[4]; A.Mz = c->F[5]; A.Mr = c->F[6];
    A.stop_at = c->stop_at;
    A.cap = c->cap;
    A.K = c->K;
    A.total = c->total;
    A.step = 0;
    A.store_forces = fused ? c->store_forces : 1;
    A.Tw = c->Tw;
    return A;
}

static int check_overflow(dem_ctx *c) {
    int ov = 0;
    HIPCHK(hipMemcpyAsync(&ov, c->overflow, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (c->nf_pending) {  // a timed neighbour search of this rebuild
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, c->nf_ev[0], c->nf_ev[1]));
        c->nf_ms[c->nf_timed] = ms;
        c->nf_pending = false;
        if (++c->nf_trials >= 2) {
            c->nf_choice = c->nf_ms[1] <= c->nf_ms[0] ? 1 : 0;
            c->nf_since = 0;
        }
    }
    c->lists_overflow = ov > 0 ? ov : 0;
    return dem_lists_ok(c);
}

// Refuses to step on lists that a rebuild truncated (kernel.cu:833-836
// writes past the list instead); the state stays refused until it is
// replaced (dem_set_particles, dem_set_contacts, dem_dump_read) or a
// rebuild fits.
int dem_lists_ok(dem_ctx *c) {
    if (c->lists_overflow > 0)
        return dem_fail(DEM_ERR_OVERFLOW,
                        "a neighbour list needed %d slots, more than its size level allows",
                        c->lists_overflow);
    return DEM_OK;
}

// ---------------------------------------------------------------------
// state import / export
// ---------------------------------------------------------------------
static void reset_lists(dem_ctx *c) {
    c->pair_packed = false;
    c->pair_bad = false;
    c->iter_at_rebuild = -1;
    c->rebuild_interval = -1;
    c->has_lists = false;
    c->saved = false;
    c->old_total = 0;
    c->rebuild_pending = true;
    c->forces_valid = false;
    c->lists_overflow = 0;
}

static int upload_particles(dem_ctx *c, int off, int n, const double *pos, const double *vel,
                            const double *angv) {
    if (n <= 0) return DEM_OK;
    // staging: pos, vel, angv (n x 3 each) and radii (n)
    double *tmp = nullptr;
    HIPCHK(hipMalloc((void **)&tmp, (size_t)n * 10 * sizeof(double)));
    double *dpos = tmp, *dvel = tmp + 3 * (size_t)n, *dang = tmp + 6 * (size_t)n,
           *drad = tmp + 9 * (size_t)n;
    std::vector<int> ids(n), lv(n);
    std::vector<double> r(n);
    for (int k = 0; k < n; ++k) {
        ids[k] = off + k;
        lv[k] = c->level_h[off + k];
        r[k] = c->rad_h[off + k];
    }
    int rc = DEM_OK;
    hipError_t e = hipMemcpyAsync(dpos, pos, (size_t)n * 3 * sizeof(double), hipMemcpyHostToDevice,
                                  c->stream);
    if (e == hipSuccess && vel)
        e = hipMemcpyAsync(dvel, vel, (size_t)n * 3 * sizeof(double), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess && angv)
        e = hipMemcpyAsync(dang, angv, (size_t)n * 3 * sizeof(double), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(drad, r.data(), n * sizeof(double), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(c->id[c->pcur] + off, ids.data(), n * sizeof(int), hipMemcpyHostToDevice,
                           c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(c->level[c->pcur] + off, lv.data(), n * sizeof(int),
                           hipMemcpyHostToDevice, c->stream);
    if (e != hipSuccess) rc = dem_fail(DEM_ERR_HIP, "upload: %s", hipGetErrorString(e));
    // both buffers get the records, so the radius is valid whichever is read
    launch_scatter_state(dpos, vel ? dvel : nullptr, angv ? dang : nullptr, drad,
                         c->buf(c->cur, c->Lidx), n, off, c->stream);
    launch_scatter_state(dpos, vel ? dvel : nullptr, angv ? dang : nullptr, drad,
                         c->buf(1 - c->cur, c->Lidx), n, off, c->stream);
    launch_fill_i32(c->lq[0] + off, 0, n, c->stream);
    launch_fill_i32(c->lq[1] + off, 0, n, c->stream);
    launch_fill_i32(c->nnb[c->hcur] + off, 0, n, c->stream);
    launch_fill_i32(c->nwl[c->hcur] + off, 0, n, c->stream);
    e = hipStreamSynchronize(c->stream);
    (void)hipFree(tmp);
    if (rc == DEM_OK && e != hipSuccess) rc = dem_fail(DEM_ERR_HIP, "upload: %s", hipGetErrorString(e));
    return rc;
}

extern "C" int dem_set_particles(dem_ctx *c, int total, const double *pos, const double *vel,
                                 const double *angv) {
    if (!c || !pos || total < 0 || total > c->cap) return dem_fail(DEM_ERR_ARG, "bad particle count");
    reset_lists(c);
    c->total = total;
    return upload_particles(c, 0, total, pos, vel, angv);
}

extern "C" int dem_feed(dem_ctx *c, int n, const double *pos) {
    if (!c || n < 0 || !pos || c->total + n > c->cap) return dem_fail(DEM_ERR_ARG, "feed exceeds NP");
    int off = c->total;
    int rc = dem_pair_unpack(c);
    if (rc) return rc;
    rc = upload_particles(c, off, n, pos, nullptr, nullptr);
    if (rc) return rc;
    c->total += n;
    c->rebuild_pending = true;  // FeedFlag (main.cpp:442, 473)
    return DEM_OK;
}

extern "C" int dem_total(dem_ctx *c) { return c ? c->total : 0; }

template <class T>
static int download(dem_ctx *c, T *host, const T *dev, size_t n) {
    if (!host || n == 0) return DEM_OK;
    HIPCHK(hipMemcpyAsync(host, dev, n * sizeof(T), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return DEM_OK;
}

// gathers a sorted-order device array into original order on the host
static int get3(dem_ctx *c, double *host, double *const *src) {
    if (!host) return DEM_OK;
    double *tmp = nullptr;
    size_t n = (size_t)c->total;
    HIPCHK(hipMalloc((void **)&tmp, (n ? n : 1) * 3 * sizeof(double)));
    launch_gather3(c->id[c->pcur], src[0], src[1], src[2], tmp, c->total, c->stream);
    int rc = download(c, host, tmp, n * 3);
    (void)hipFree(tmp);
    return rc;
}

static int get1(dem_ctx *c, double *host, const double *src) {
    if (!host) return DEM_OK;
    double *tmp = nullptr;
    size_t n = (size_t)c->total;
    HIPCHK(hipMalloc((void **)&tmp, (n ? n : 1) * sizeof(double)));
    launch_gather1(c->id[c->pcur], src, tmp, c->total, c->stream);
    int rc = download(c, host, tmp, n);
    (void)hipFree(tmp);
    return rc;
}

static int get1i(dem_ctx *c, int *host, const int *src) {
    if (!host) return DEM_OK;
    int *tmp = nullptr;
    size_t n = (size_t)c->total;
    HIPCHK(hipMalloc((void **)&tmp, (n ? n : 1) * sizeof(int)));
    launch_gather1i(c->id[c->pcur], src, tmp, c->total, c->stream);
    int rc = download(c, host, tmp, n);
    (void)hipFree(tmp);
    return rc;
}

// one field of the particle records, in original order: what = 0 position,
// 1 velocity, 2 spin (3 values each), 4 path length (one value)
static int get_state(dem_ctx *c, double *host, int what) {
    if (!host) return DEM_OK;
    double *tmp = nullptr;
    size_t n = (size_t)c->total, w = what < 3 ? 3 : 1;
    HIPCHK(hipMalloc((void **)&tmp, (n ? n : 1) * w * sizeof(double)));
    launch_gather_state(c->id[c->pcur], c->buf(c->cur, c->Lidx), what, tmp, c->total, c->stream);
    int rc = download(c, host, tmp, n * w);
    (void)hipFree(tmp);
    return rc;
}

extern "C" int dem_get_particles(dem_ctx *c, double *pos, double *vel, double *angv) {
    if (!c) return dem_fail(DEM_ERR_ARG, "null context");
    int rc;
    if ((rc = get_state(c, pos, 0)) || (rc = get_state(c, vel, 1)) || (rc = get_state(c, angv, 2)))
        return rc;
    return DEM_OK;
}

extern "C" int dem_get_forces(dem_ctx *c, double *F, double *M, double *Mr) {
    if (!c) return dem_fail(DEM_ERR_ARG, "null context");
    if (!c->forces_valid) return dem_fail(DEM_ERR_STATE, "no forces stored for the current state");
    int rc;
    if ((rc = get3(c, F, &c->F[0])) || (rc = get3(c, M, &c->F[3])) || (rc = get1(c, Mr, c->F[6])))
        return rc;
    return DEM_OK;
}

extern "C" int dem_set_store_forces(dem_ctx *c, int on) {
    if (!c) return dem_fail(DEM_ERR_ARG, "null context");
    c->store_forces = on ? 1 : 0;
    return DEM_OK;
}

extern "C" int dem_get_pmove(dem_ctx *c, double *pmove) {
    if (!c) return dem_fail(DEM_ERR_ARG, "null context");
    return get_state(c, pmove, 4);
}

extern "C" int dem_get_lqdcnt(dem_ctx *c, int *lqdCnt) {
    if (!c) return dem_fail(DEM_ERR_ARG, "null context");
    return get1i(c, lqdCnt, c->lq[c->Lidx]);
}

// sorted-order index -> original id, on the host
static int host_ids(dem_ctx *c, std::vector<int> &ids) {
    ids.resize(c->total);
    return download(c, ids.data(), c->id[c->pcur], (size_t)c->total);
}

extern "C" int dem_get_nebor(dem_ctx *c, int *num, int *list, int maxK) {
    if (!c || !num || !list || maxK <= 0) return dem_fail(DEM_ERR_ARG, "bad argument");
    const int n = c->total;
    std::vector<int> ids, cnt(n), nb((size_t)c->K * (n ? n : 1));
    int rc;
    if ((rc = host_ids(c, ids)) || (rc = download(c, cnt.data(), c->nnb[c->hcur], (size_t)n)))
        return rc;
    for (int k = 0; k < c->K; ++k)
        if ((rc = download(c, nb.data() + (size_t)k * n, c->nbr[c->hcur] + (size_t)k * c->cap, (size_t)n)))
            return rc;
    int bad = 0;
    for (int i = 0; i < n; ++i) {
        int o = ids[i];
        num[o] = cnt[i];
        if (cnt[i] > maxK) bad = 1;
        for (int k = 0; k < maxK; ++k)
            list[(size_t)o * maxK + k] = k < cnt[i] ? ids[nb[(size_t)k * n + i]] : -1;
    }
    return bad ? dem_fail(DEM_ERR_ARG, "maxK smaller than a list") : DEM_OK;
}

static int get_hist(dem_ctx *c, const HistRec *h, const int *dcnt, int slots, int maxK,
                    double *dispt, double *nDsp, double *nDsp_mx, double *plstRad,
                    double *plstDfm, int *stage, int *lqdBrg) {
    const int n = c->total;
    std::vector<int> ids, cnt(n);
    std::vector<HistRec> rec(n);
    int rc;
    if ((rc = host_ids(c, ids)) || (rc = download(c, cnt.data(), dcnt, (size_t)n))) return rc;
    for (int k = 0; k < maxK; ++k) {
        bool live = k < slots;
        if (live && (rc = download(c, rec.data(), h + (size_t)k * c->cap, (size_t)n))) return rc;
        for (int i = 0; i < n; ++i) {
            size_t o = (size_t)ids[i] * maxK + k;
            HistRec r;
            memset(&r, 0, sizeof r);
            if (live && k < cnt[i]) r = rec[i];
            if (dispt) {
                dispt[3 * o] = r.dx;
                dispt[3 * o + 1] = r.dy;
                dispt[3 * o + 2] = r.dz;
            }
            if (nDsp) nDsp[o] = r.nd;
            if (nDsp_mx) nDsp_mx[o] = r.ndm;
            if (plstRad) plstRad[o] = r.prad;
            if (plstDfm) plstDfm[o] = r.pdfm;
            if (stage) stage[o] = r.flags & 0xff;
            if (lqdBrg) lqdBrg[o] = (r.flags >> 8) & 1;
        }
    }
    return DEM_OK;
}

// a slot-major copy of the history (slot k of particle i at k*cap + i), for
// export and import
static int slot_view(dem_ctx *c, HistRec **out) {
    *out = nullptr;
    HIPCHK(hipMalloc((void **)out, (size_t)c->cap * c->K * sizeof(HistRec)));
    HIPCHK(hipMemsetAsync(*out, 0, (size_t)c->cap * c->K * sizeof(HistRec), c->stream));
    return DEM_OK;
}

extern "C" int dem_get_hist_pp(dem_ctx *c, int maxK, double *dispt, double *nDsp, double *nDsp_mx,
                               double *plstRad, double *plstDfm, int *stage, int *lqdBrg) {
    if (!c || maxK <= 0) return dem_fail(DEM_ERR_ARG, "bad argument");
    HistRec *v = nullptr;
    int rc = dem_pair_unpack(c);
    if (rc) return rc;
    rc = slot_view(c, &v);
    if (rc == DEM_OK) {
        launch_hist_slots(c->hp[c->hl], c->nnb[c->hcur], c->amask[c->hl], c->actoff[c->hl], v,
                          c->cap, (long)c->K * c->cap, c->total, true, c->stream);
        rc = get_hist(c, v, c->nnb[c->hcur], c->K, maxK, dispt, nDsp, nDsp_mx, plstRad, plstDfm, stage,
                      lqdBrg);
    }
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(v);
    return rc;
}

extern "C" int dem_get_walls(dem_ctx *c, int *num, int *list) {
    if (!c || !num || !list) return dem_fail(DEM_ERR_ARG, "bad argument");
    const int n = c->total;
    std::vector<int> ids, cnt(n), w((size_t)DEM_WALLS * (n ? n : 1));
    int rc;
    if ((rc = host_ids(c, ids)) || (rc = download(c, cnt.data(), c->nwl[c->hcur], (size_t)n)))
        return rc;
    for (int k = 0; k < DEM_WALLS; ++k)
        if ((rc = download(c, w.data() + (size_t)k * n, c->wl[c->hcur] + (size_t)k * c->cap, (size_t)n)))
            return rc;
    for (int i = 0; i < n; ++i) {
        num[ids[i]] = cnt[i];
        for (int k = 0; k < DEM_WALLS; ++k)
            list[(size_t)ids[i] * DEM_WALLS + k] = k < cnt[i] ? w[(size_t)k * n + i] : -1;
    }
    return DEM_OK;
}

extern "C" int dem_get_hist_pw(dem_ctx *c, double *dispt, double *nDsp, double *nDsp_mx,
                               double *plstRad, double *plstDfm, int *stage, int *lqdBrg) {
    if (!c) return dem_fail(DEM_ERR_ARG, "null context");
    return get_hist(c, c->hw[c->hcur], c->nwl[c->hcur], DEM_WALLS, DEM_WALLS, dispt, nDsp, nDsp_mx,
                    plstRad, plstDfm, stage, lqdBrg);
}

// restores lists and history in original order (identity permutation)
static int put_hist(dem_ctx *c, HistRec *h, int maxK, int n, const double *dispt,
                    const double *nDsp, const double *nDsp_mx, const double *plstRad,
                    const double *plstDfm, const int *stage, const int *lqdBrg) {
    std::vector<HistRec> rec(n);
    for (int k = 0; k < maxK; ++k) {
        for (int i = 0; i < n; ++i) {
            size_t o = (size_t)i * maxK + k;
            HistRec &r = rec[i];
            r.dx = dispt ? dispt[3 * o] : 0.0;
            r.dy = dispt ? dispt[3 * o + 1] : 0.0;
            r.dz = dispt ? dispt[3 * o + 2] : 0.0;
            r.nd = nDsp ? nDsp[o] : 0.0;
            r.ndm = nDsp_mx ? nDsp_mx[o] : 0.0;
            r.prad = plstRad ? plstRad[o] : 0.0;
            r.pdfm = plstDfm ? plstDfm[o] : 0.0;
            r.flags = (stage ? (stage[o] & 0xff) : 0) | ((lqdBrg && lqdBrg[o]) ? 256 : 0);
            r.pad = 0;
        }
        HIPCHK(hipMemcpyAsync(h + (size_t)k * c->cap, rec.data(), n * sizeof(HistRec),
                              hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    return DEM_OK;
}

extern "C" int dem_set_contacts(dem_ctx *c, int maxK, const int *num, const int *list,
                                const double *dispt, const double *nDsp, const double *nDsp_mx,
                                const double *plstRad, const double *plstDfm, const int *stage,
                                const int *lqdBrg, const int *wnum, const int *wlist,
                                const double *wdispt, const double *wnDsp, const double *wnDsp_mx,
                                const double *wplstRad, const double *wplstDfm, const int *wstage,
                                const int *wlqdBrg) {
    if (!c || !num || !list || !wnum || !wlist || maxK <= 0 || maxK > c->K)
        return dem_fail(DEM_ERR_ARG, "bad argument");
    c->pair_packed = false;  // the records are replaced
    c->pair_bad = false;
    const int n = c->total;
    // the restored lists must be in the current (identity) order: ids 0..n-1
    std::vector<int> ids;
    int rc = host_ids(c, ids);
    if (rc) return rc;
    for (int i = 0; i < n; ++i)
        if (ids[i] != i)
            return dem_fail(DEM_ERR_STATE, "dem_set_contacts needs state fresh from dem_set_particles");
    std::vector<int> col(n);
    for (int k = 0; k < maxK; ++k) {
        for (int i = 0; i < n; ++i) {
            int j = list[(size_t)i * maxK + k];
            col[i] = (k < num[i] && j >= 0 && j < n) ? j : 0;
        }
        HIPCHK(hipMemcpyAsync(c->nbr[c->hcur] + (size_t)k * c->cap, col.data(), n * sizeof(int),
                              hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    std::vector<int> cnt(num, num + n);
    for (int i = 0; i < n; ++i)
        if (cnt[i] < 0 || cnt[i] > maxK) return dem_fail(DEM_ERR_ARG, "list count out of range");
    HIPCHK(hipMemcpyAsync(c->nnb[c->hcur], cnt.data(), n * sizeof(int), hipMemcpyHostToDevice, c->stream));
    std::vector<int> wc(wnum, wnum + n);
    for (int k = 0; k < DEM_WALLS; ++k) {
        for (int i = 0; i < n; ++i) col[i] = k < wnum[i] ? wlist[(size_t)i * DEM_WALLS + k] : 0;
        HIPCHK(hipMemcpyAsync(c->wl[c->hcur] + (size_t)k * c->cap, col.data(), n * sizeof(int),
                              hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    HIPCHK(hipMemcpyAsync(c->nwl[c->hcur], wc.data(), n * sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    // restored history goes to the slot-major region (no compact records)
    launch_fill_i32((int *)c->amask[c->hl], 0, 2L * n, c->stream);
    launch_fill_i32(c->actoff[c->hl], 0, n, c->stream);
    // no bound on how far the restored lists' particles have moved: every
    // slot is tested
    launch_fill_i32((int *)c->nmask, -1, 2L * n, c->stream);
    HistRec *v = nullptr;
    if ((rc = slot_view(c, &v))) return rc;
    rc = put_hist(c, v, maxK, n, dispt, nDsp, nDsp_mx, plstRad, plstDfm, stage, lqdBrg);
    if (rc == DEM_OK)
        launch_hist_slots(c->hp[c->hl], c->nnb[c->hcur], c->amask[c->hl], c->actoff[c->hl], v,
                          c->cap, (long)c->K * c->cap, n, false, c->stream);
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(v);
    if (rc) return rc;
    if ((rc = put_hist(c, c->hw[c->hcur], DEM_WALLS, n, wdispt, wnDsp, wnDsp_mx, wplstRad, wplstDfm,
                       wstage, wlqdBrg)))
        return rc;
    c->has_lists = true;
    c->saved = false;
    c->old_total = n;
    c->rebuild_pending = true;  // RESTART (main.cpp:473)
    return DEM_OK;
}

// The device tables give each cell's range in storage order (row-major or
// Morton), in the blocked cell layout (cell_slot); the reference's tables
// (thrust-sorted (hash, index) pairs and CellStart/CellEnd into them)
// follow from walking the cells in row-major order, since a cell's
// particles are stored by original index either way.
static int ref_tables(dem_ctx *c, int *hash, int *index, int *cellStart, int *cellEnd) {
    const size_t nt = (size_t)c->P.ncellTab, nc = (size_t)c->S.sysNCells;
    std::vector<int> cs(nt), ce(nt), ids;
    int rc;
    if ((rc = download(c, cs.data(), c->cstart, nt)) || (rc = download(c, ce.data(), c->cend, nt)) ||
        (rc = host_ids(c, ids)))
        return rc;
    // the table entry of every row-major cell, level by level
    std::vector<int> slot(nc);
    for (int L = 0; L < c->S.nLevels; ++L) {
        const int gx = c->P.gx[L], gy = c->P.gy[L], gz = c->P.gz[L];
        size_t h = (size_t)c->P.cellOff[L];
        for (int z = 0; z < gz; ++z)
            for (int y = 0; y < gy; ++y)
                for (int x = 0; x < gx; ++x) slot[h++] = cell_slot(c->P, L, x, y, z);
    }
    int acc = 0;
    for (size_t cell = 0; cell < nc; ++cell) {
        const int t = slot[cell];
        if (cs[t] == DEM_EMPTY) {
            if (cellStart) cellStart[cell] = DEM_EMPTY;
            if (cellEnd) cellEnd[cell] = DEM_EMPTY;
            continue;
        }
        const int n = ce[t] - cs[t];
        if (cellStart) cellStart[cell] = acc;
        if (cellEnd) cellEnd[cell] = acc + n;
        for (int q = cs[t]; q < ce[t]; ++q, ++acc) {
            if (hash) hash[acc] = (int)cell;
            if (index) index[acc] = ids[q];
        }
    }
    return DEM_OK;
}

extern "C" int dem_get_hash(dem_ctx *c, int *hash, int *index) {
    if (!c) return dem_fail(DEM_ERR_ARG, "null context");
    return ref_tables(c, hash, index, nullptr, nullptr);
}

extern "C" int dem_get_cells(dem_ctx *c, int *cellStart, int *cellEnd) {
    if (!c) return dem_fail(DEM_ERR_ARG, "null context");
    return ref_tables(c, nullptr, nullptr, cellStart, cellEnd);
}

// ---------------------------------------------------------------------
// shared pair records of the wave kernel (PAIR_MIRROR in dem_internal.h)
// ---------------------------------------------------------------------
void dem_pair_free(dem_ctx *c) {
    for (int b = 0; b < 2; ++b) {
        (void)hipFree(c->hq[b]);
        (void)hipFree(c->hc[b]);
        c->hq[b] = nullptr;
        c->hc[b] = nullptr;
    }
    (void)hipFree(c->nl); (void)hipFree(c->pcnt); (void)hipFree(c->poff);
    (void)hipFree(c->pair_err);
    c->nl = nullptr; c->pcnt = nullptr; c->poff = nullptr; c->pair_err = nullptr;
    c->hq_cap = c->nl_cap = 0;
    c->pair_packed = false;
}

static int grow(dem_ctx *c, void **p, long *have, long need, size_t elem) {
    if (need <= *have) return DEM_OK;
    long n = need + need / 4 + 1024;
    (void)hipFree(*p);
    *p = nullptr;
    *have = 0;
    HIPCHK(hipMalloc(p, (size_t)n * elem));
    *have = n;
    return DEM_OK;
}

// The directed records of the near slots become pair records (both
// buffers), with the lists unchanged.  When the lists or the records are
// not symmetric (k_pair_fill_mirror), the context keeps to the directed
// records from then on.
int dem_pair_pack(dem_ctx *c) {
    if (c->pair_packed || c->pair_bad || !c->has_lists || c->total <= 0) return DEM_OK;
    const int n = c->total;
    const long cap = c->cap, Kcap = (long)c->K * c->cap;
    hipStream_t s = c->stream;
    if (!c->pcnt) {
        HIPCHK(hipMalloc((void **)&c->pcnt, 4 * (size_t)cap * sizeof(unsigned)));
        HIPCHK(hipMalloc((void **)&c->poff, (3 * (size_t)cap + 1) * sizeof(unsigned)));
        HIPCHK(hipMalloc((void **)&c->pair_err, sizeof(int)));
    }
    launch_pair_count(c->P4[c->cur], c->nbr[c->hcur], c->nnb[c->hcur], c->nmask, c->pcnt, c->cap, n, s);
    for (int q = 0; q < 3; ++q)
        launch_excl_scan(c->pcnt + q * cap, c->poff + q * cap, n, c->sort_tmp, s);
    // work of each chunk of 64 particles, for the XCD ranges of the wave kernel
    const int nch = (n + 63) / 64;
    std::vector<unsigned> work(nch);
    launch_chunk_work(c->pcnt + 3 * cap, c->keys2, n, s);
    HIPCHK(hipMemcpyAsync(work.data(), c->keys2, nch * sizeof(unsigned), hipMemcpyDeviceToHost, s));
    unsigned last[6];
    for (int q = 0; q < 3; ++q) {
        HIPCHK(hipMemcpyAsync(&last[2 * q], c->pcnt + q * cap + (n - 1), sizeof(unsigned),
                              hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(&last[2 * q + 1], c->poff + q * cap + (n - 1), sizeof(unsigned),
                              hipMemcpyDeviceToHost, s));
    }
    HIPCHK(hipStreamSynchronize(s));
    const long t1 = (long)last[0] + last[1], t2 = (long)last[2] + last[3], tn = (long)last[4] + last[5];
    {
        double W = 0.0;
        for (unsigned w : work) W += w;
        double acc = 0.0;
        int x = 1;
        c->xb_work[0] = 0;
        for (int ch = 0; ch < nch && x < 8; ++ch) {
            while (x < 8 && acc >= W * x / 8.0) c->xb_work[x++] = (unsigned)ch;
            acc += work[ch];
        }
        while (x < 8) c->xb_work[x++] = (unsigned)nch;
        c->xb_work[8] = (unsigned)nch;
        c->xb_work_ok = true;
    }
    if (t1 + t2 > 0x7fffffffL || tn > 0x7fffffffL)
        return dem_fail(DEM_ERR_NOMEM, "too many near slots for 32-bit pair record indices");
    long h[4] = {c->hq_cap, c->hq_cap, c->hq_cap, c->hq_cap};
    int rc;
    if ((rc = grow(c, (void **)&c->hq[0], &h[0], t1 + t2, sizeof(PairHot))) ||
        (rc = grow(c, (void **)&c->hq[1], &h[1], t1 + t2, sizeof(PairHot))) ||
        (rc = grow(c, (void **)&c->hc[0], &h[2], t1 + t2, sizeof(double2))) ||
        (rc = grow(c, (void **)&c->hc[1], &h[3], t1 + t2, sizeof(double2))) ||
        (rc = grow(c, (void **)&c->nl, &c->nl_cap, tn, sizeof(int2))))
        return rc;
    // the near lists end at noff[n]
    const unsigned tnu = (unsigned)tn;
    HIPCHK(hipMemcpyAsync(c->poff + 2 * cap + n, &tnu, sizeof(unsigned), hipMemcpyHostToDevice, s));
    c->hq_cap = h[0];
    for (int q = 1; q < 4; ++q) c->hq_cap = h[q] < c->hq_cap ? h[q] : c->hq_cap;
    launch_fill_i32(c->pair_err, 0, 1, s);
    launch_pair_fill(c->P4[c->cur], c->VW[c->cur], c->nbr[c->hcur], c->nnb[c->hcur], c->nmask, c->poff,
                     (unsigned)t1, c->hp[c->hl], c->amask[c->hl], c->actoff[c->hl], c->nl, c->hq[0],
                     c->hq[1], c->hc[0], c->hc[1], c->pair_err, c->cap, Kcap, n, s);
    HIPCHK(hipGetLastError());
    int bad = 0;
    HIPCHK(hipMemcpyAsync(&bad, c->pair_err, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (bad) {
        c->pair_bad = true;
        return DEM_OK;
    }
    c->pair_packed = true;
    c->hqc = 0;
    c->layout_age = 0;
    c->n_pack++;
    return DEM_OK;
}

int dem_pair_unpack(dem_ctx *c) {
    if (!c->pair_packed) return DEM_OK;
    c->pair_packed = false;
    launch_pair_unpack(c->P4[c->cur], c->nnb[c->hcur], c->nmask, c->poff + 2L * c->cap, c->nl,
                       c->hq[c->hqc], c->hc[c->hqc], c->hp[c->hl], c->amask[c->hl],
                       c->actoff[c->hl], c->cap, (long)c->K * c->cap, c->total, c->stream);
    HIPCHK(hipGetLastError());
    return DEM_OK;
}

extern "C" int dem_set_pair_records(dem_ctx *c, int on) {
    if (!c) return dem_fail(DEM_ERR_ARG, "null context");
    int rc;
    if (!on && (rc = dem_pair_unpack(c))) return rc;
    c->pair_on = on != 0;
    return DEM_OK;
}

extern "C" int dem_pair_records_used(dem_ctx *c) {
    return c && c->pair_on && !c->pair_bad && dem_step_kernel(c) == DEM_KERNEL_WAVE && pair_fits(c) &&
           (c->rebuild_interval < 0 || c->rebuild_interval >= kPairMinInterval);
}

// ---------------------------------------------------------------------
// rebuild pieces (main.cpp:473-527)
// ---------------------------------------------------------------------
// event after rebuild stage k (0 = start) while a profile is taken
static void rb_mark(dem_ctx *c, int k) {
    if (c->rb_prof) (void)hipEventRecord(c->rb_ev[k], c->stream);
}

extern "C" int dem_cnt_hstry_save(dem_ctx *c) {
    if (!c) return dem_fail(DEM_ERR_ARG, "null context");
    rb_mark(c, 0);
    int rc = dem_pair_unpack(c);
    if (rc) return rc;
    rb_mark(c, 1);
    if (c->has_lists) {
        c->hcur = 1 - c->hcur;
        c->hl = 1 - c->hl;
        c->saved = true;
    }
    return DEM_OK;
}

extern "C" int dem_bld_nebor_list(dem_ctx *c) {
    if (!c) return dem_fail(DEM_ERR_ARG, "null context");
    const int n = c->total;
    hipStream_t s = c->stream;
    c->old_total = c->saved ? c->old_total : 0;
    StateBuf a = c->buf(c->cur, c->Lidx), b = c->buf(1 - c->cur, 1 - c->Lidx);
    launch_calc_hash(a, c->level[c->pcur], c->id[c->pcur], c->P, c->keys, c->vals, n, s);
    rb_mark(c, 2);
    launch_radix_sort(c->keys, c->vals, c->keys2, c->vals2, n, c->key_bits, c->sort_tmp, s);
    rb_mark(c, 3);
    launch_permute(c->vals, a, b, c->level[c->pcur], c->level[1 - c->pcur], c->id[c->pcur],
                   c->id[1 - c->pcur], n, s);
    rb_mark(c, 4);
    c->cur = 1 - c->cur;
    c->Lidx = 1 - c->Lidx;
    c->pcur = 1 - c->pcur;
    launch_cell_range(c->keys, c->buf(c->cur, c->Lidx), c->level[c->pcur], c->P, c->cstart,
                      c->cend, c->P.ncellTab, n, s);
    rb_mark(c, 5);
    launch_fill_i32(c->overflow, 0, 1, s);
    // the LDS-staged (tiled) search pays where tiles hold many particles:
    // under AUTO each context times both forms and keeps the faster
    bool tiled = false;
    const bool can_tile = neighbor_tiles_apply(c->P);
    bool timing = false;
    if (can_tile) {
        if (c->nf_mode == DEM_NSEARCH_TILED) {
            tiled = true;
        } else if (c->nf_mode == DEM_NSEARCH_AUTO) {
            if (c->nf_choice >= 0 && ++c->nf_since >= 256) {
                c->nf_choice = -1;  // conditions change (a pour settles): time both again
                c->nf_trials = 0;
            }
            tiled = c->nf_choice >= 0 ? c->nf_choice == 1 : (c->nf_trials & 1) == 0;
            timing = c->nf_choice < 0;
        }
    }
    if (timing) {
        for (hipEvent_t &e : c->nf_ev)
            if (!e) HIPCHK(hipEventCreate(&e));
        HIPCHK(hipEventRecord(c->nf_ev[0], s));
    }
    launch_neighbor_find(c->buf(c->cur, c->Lidx), c->level[c->pcur], c->P, c->keys, c->cstart,
                         c->cend, c->nbr[c->hcur], c->nnb[c->hcur], c->overflow, tiled, c->cap, n, s);
    if (timing) {
        HIPCHK(hipEventRecord(c->nf_ev[1], s));
        c->nf_timed = tiled ? 1 : 0;
        c->nf_pending = true;
    }
    rb_mark(c, 6);
    launch_active_slots(c->P4[c->cur], c->nbr[c->hcur], c->nnb[c->hcur], c->amask[c->hl],
                        c->actoff[c->hl], c->nmask, near_gap_of(c->P), c->keys2, c->sort_tmp,
                        c->cap, n, s);
    HIPCHK(hipGetLastError());
    return DEM_OK;
}

extern "C" int dem_bld_cir_wall(dem_ctx *c) {
    if (!c) return dem_fail(DEM_ERR_ARG, "null context");
    rb_mark(c, 7);
    launch_wall_list(c->buf(c->cur, c->Lidx), c->P, c->Tw, c->wl[c->hcur], c->nwl[c->hcur], c->cap,
                     c->total, c->stream);
    rb_mark(c, 8);
    HIPCHK(hipGetLastError());
    return DEM_OK;
}

extern "C" int dem_hstry_copy(dem_ctx *c) {
    if (!c) return dem_fail(DEM_ERR_ARG, "null context");
    const int o = 1 - c->hcur, ol = 1 - c->hl;
    launch_hist_remap(c->vals, c->id[c->pcur], c->id[1 - c->pcur], c->nbr[c->hcur], c->nnb[c->hcur],
                      c->amask[c->hl], c->actoff[c->hl], c->nbr[o], c->nnb[o], c->amask[ol],
                      c->actoff[ol], c->hp[c->hl], c->hp[ol], c->wl[c->hcur],
                      c->nwl[c->hcur], c->wl[o], c->nwl[o], c->hw[c->hcur], c->hw[o],
                      c->cap, (long)c->K * c->cap, c->total, c->saved ? c->old_total : 0, c->stream);
    rb_mark(c, 9);
    HIPCHK(hipGetLastError());
    c->saved = false;
    c->has_lists = true;
    c->old_total = c->total;
    return DEM_OK;
}

int dem_do_rebuild(dem_ctx *c) {
    int rc;
    if ((rc = dem_cnt_hstry_save(c)) || (rc = dem_bld_nebor_list(c)) || (rc = dem_bld_cir_wall(c)) ||
        (rc = dem_hstry_copy(c)))
        return rc;
    c->rebuild_pending = false;
    c->n_rebuild++;
    // two rebuilds at one step (a restart's lists, then its RESTART rebuild;
    // a feed) say nothing about how far apart motion brings them
    if (c->iter_at_rebuild >= 0 && c->iter > c->iter_at_rebuild)
        c->rebuild_interval = c->iter - c->iter_at_rebuild;
    c->iter_at_rebuild = c->iter;
    c->forces_valid = false;
    c->layout_age = 0;
    return DEM_OK;
}

// Layout refresh: the compact region is rebuilt from the contacts of now,
// with the lists unchanged.  Contacts that formed since the layout was made
// move out of the slot-major region; records only move, so results are
// unchanged.
int dem_relayout(dem_ctx *c) {
    if (!c->has_lists || c->total <= 0) return DEM_OK;
    int rc = dem_pair_unpack(c);
    if (rc) return rc;
    const int o = c->hl, n = c->total;
    c->hl = 1 - c->hl;
    launch_active_slots(c->P4[c->cur], c->nbr[c->hcur], c->nnb[c->hcur], c->amask[c->hl],
                        c->actoff[c->hl], nullptr, 0.0, c->keys2, c->sort_tmp, c->cap, n,
                        c->stream);
    launch_hist_relayout(c->nnb[c->hcur], c->amask[o], c->actoff[o], c->amask[c->hl],
                         c->actoff[c->hl], c->hp[o], c->hp[c->hl], c->cap, (long)c->K * c->cap, n,
                         c->stream);
    HIPCHK(hipGetLastError());
    c->layout_age = 0;
    c->n_relayout++;
    return DEM_OK;
}

extern "C" int dem_rebuild(dem_ctx *c) {
    if (!c) return dem_fail(DEM_ERR_ARG, "null context");
    int rc = dem_do_rebuild(c);
    if (rc) return rc;
    return check_overflow(c);
}

extern "C" int dem_flush_rebuild(dem_ctx *c) {
    if (!c) return dem_fail(DEM_ERR_ARG, "null context");
    if (!c->rebuild_pending) return DEM_OK;
    return dem_rebuild(c);
}

extern "C" int dem_key_bits(dem_ctx *c) { return c ? c->key_bits : 0; }

extern "C" int dem_set_neighbor_search(dem_ctx *c, int mode) {
    if (!c || mode < DEM_NSEARCH_AUTO || mode > DEM_NSEARCH_TILED)
        return dem_fail(DEM_ERR_ARG, "bad argument");
    c->nf_mode = mode;
    c->nf_choice = -1;
    c->nf_trials = 0;
    c->nf_pending = false;
    return DEM_OK;
}

extern "C" int dem_neighbor_search_used(dem_ctx *c) {
    if (!c || !neighbor_tiles_apply(c->P) || c->nf_mode == DEM_NSEARCH_PLAIN) return DEM_NSEARCH_PLAIN;
    if (c->nf_mode == DEM_NSEARCH_TILED) return DEM_NSEARCH_TILED;
    return c->nf_choice == 1 ? DEM_NSEARCH_TILED : c->nf_choice == 0 ? DEM_NSEARCH_PLAIN
                                                                      : DEM_NSEARCH_AUTO;
}

extern "C" int dem_scale_velocities(dem_ctx *c, double factor) {
    if (!c || !(factor >= 0.0)) return dem_fail(DEM_ERR_ARG, "bad argument");
    launch_scale_velocities(c->buf(c->cur, c->Lidx), factor, c->total, c->stream);
    HIPCHK(hipGetLastError());
    return DEM_OK;
}

extern "C" int dem_profile_rebuild(dem_ctx *c, int reps, double *stage_ms) {
    if (!c || reps < 1 || !stage_ms) return dem_fail(DEM_ERR_ARG, "bad argument");
    if (!c->has_lists || c->total <= 0) return dem_fail(DEM_ERR_STATE, "no lists to rebuild");
    for (int k = 0; k <= dem_ctx::kRbStages; ++k)
        if (!c->rb_ev[k]) HIPCHK(hipEventCreate(&c->rb_ev[k]));
    for (int k = 0; k < dem_ctx::kRbStages; ++k) stage_ms[k] = 0.0;
    int rc = DEM_OK;
    for (int r = 0; r < reps && rc == DEM_OK; ++r) {
        c->rb_prof = true;
        rc = dem_rebuild(c);  // synchronises (overflow check)
        c->rb_prof = false;
        for (int k = 0; rc == DEM_OK && k < dem_ctx::kRbStages; ++k) {
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, c->rb_ev[k], c->rb_ev[k + 1]));
            stage_ms[k] += ms / reps;
        }
    }
    return rc;
}

// ---------------------------------------------------------------------
// the reference's separate kernels
// ---------------------------------------------------------------------
extern "C" int dem_var_init(dem_ctx *c) {
    if (!c) return dem_fail(DEM_ERR_ARG, "null context");
    c->Lidx = 1 - c->Lidx;  // lqdCnt_Old <- lqdCnt
    launch_fill_i32(c->lq[c->Lidx], 0, c->cap, c->stream);
    for (int q = 0; q < 7; ++q) launch_fill_f64(c->F[q], 0.0, c->cap, c->stream);
    HIPCHK(hipGetLastError());
    return DEM_OK;
}

extern "C" int dem_pp_inter_calc(dem_ctx *c) {
    if (!c) return dem_fail(DEM_ERR_ARG, "null context");
    int rc = dem_lists_ok(c);
    if (rc || (rc = dem_pair_unpack(c))) return rc;
    StepArgs A = dem_step_args(c, false);
    launch_pp_force(c->P, A, c->S.forceCap == 1, c->precision == DEM_PREC_MIXED, c->stream);
    HIPCHK(hipGetLastError());
    return DEM_OK;
}

extern "C" int dem_cir_pw_calc(dem_ctx *c) {
    if (!c) return dem_fail(DEM_ERR_ARG, "null context");
    int rc = dem_pair_unpack(c);
    if (rc) return rc;
    c->Tw = c->Tw + c->S.topv * c->S.dt;  // kernel.cu:966-969
    StepArgs A = dem_step_args(c, false);
    launch_pw_force(c->P, A, c->precision == DEM_PREC_MIXED, c->stream);
    HIPCHK(hipGetLastError());
    return DEM_OK;
}

extern "C" int dem_update_system(dem_ctx *c, int *reNebld) {
    if (!c) return dem_fail(DEM_ERR_ARG, "null context");
    int rc = dem_pair_unpack(c);
    if (rc) return rc;
    StepArgs A = dem_step_args(c, false);
    launch_fill_i32(c->upd_count, 0, 1, c->stream);
    launch_update(c->P, A, c->upd_count, c->stream);
    HIPCHK(hipGetLastError());
    int cnt = 0;
    HIPCHK(hipMemcpyAsync(&cnt, c->upd_count, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (reNebld) *reNebld = cnt;
    c->forces_valid = true;
    return DEM_OK;
}

extern "C" int dem_step_unfused(dem_ctx *c) {
    if (!c) return dem_fail(DEM_ERR_ARG, "null context");
    int rc;
    if (c->rebuild_pending && (rc = dem_rebuild(c))) return rc;
    if ((rc = dem_lists_ok(c))) return rc;
    int flag = 0;
    if ((rc = dem_var_init(c)) || (rc = dem_pp_inter_calc(c)) || (rc = dem_cir_pw_calc(c)) ||
        (rc = dem_update_system(c, &flag)))
        return rc;
    c->rebuild_pending = flag > 0;
    c->ctime = c->ctime + c->S.dt;
    c->iter = c->iter + 1;
    return DEM_OK;
}

// ---------------------------------------------------------------------
// the fused path
// ---------------------------------------------------------------------
// the step kernel that dem_step launches: DEM_KERNEL_LOOP, _COMPACT or _WAVE
// (the wave kernel takes any packing: wet ones with the capillary block, a
// particle with more than 64 slots stepped whole by its owner lane; the
// block-compaction kernel only dry packings with at most 64 slots)
int dem_step_kernel(const dem_ctx *c) {
    if (c->variant == DEM_KERNEL_COMPACT)
        return c->S.forceCap == 1 || c->K > 64 ? DEM_KERNEL_LOOP : DEM_KERNEL_COMPACT;
    return c->variant == DEM_KERNEL_AUTO ? DEM_KERNEL_WAVE : c->variant;
}

extern "C" int dem_set_relayout(dem_ctx *c, int every) {
    if (!c || every < 0) return dem_fail(DEM_ERR_ARG, "bad argument");
    c->relayout_every = every;
    return DEM_OK;
}

extern "C" int dem_n_relayout(dem_ctx *c) { return c ? c->n_relayout : 0; }

extern "C" int dem_set_wave_split(dem_ctx *c, int parts) {
    if (!c || (parts != 0 && parts != 1 && parts != 2 && parts != 4))
        return dem_fail(DEM_ERR_ARG, "sub-waves per chunk must be 0 (by size), 1, 2 or 4");
    c->wave_split = parts;
    return DEM_OK;
}

// Sub-waves per chunk of 64 particles for the wave kernel.  By default one
// wave per chunk: measured on the 200,000-sphere dense wet bed, whose 3,125
// chunks give each of the chip's 3,072 wave slots about one, two sub-waves
// per chunk ran at the same speed (147 us per step) and four were 8 %
// slower; on the wet pour four were 39 % slower (each sub-wave repeats the
// wave's fixed costs with a quarter of the owners).
int dem_wave_split(const dem_ctx *c) { return c->wave_split > 0 ? c->wave_split : 1; }

extern "C" int dem_set_kernel(dem_ctx *c, int variant) {
    if (!c) return dem_fail(DEM_ERR_ARG, "null context");
    if (variant < DEM_KERNEL_AUTO || variant > DEM_KERNEL_WAVE)
        return dem_fail(DEM_ERR_ARG, "unknown kernel variant %d", variant);
    if (variant == DEM_KERNEL_COMPACT && (c->S.forceCap == 1 || c->K > 64))
        return dem_fail(DEM_ERR_ARG, "the block-compaction kernel needs a dry packing with <= 64 slots");
    int rc = dem_pair_unpack(c);
    if (rc) return rc;
    c->variant = variant;
    return DEM_OK;
}

static int run_chunk(dem_ctx *c, int nsteps, int *executed) {
    const bool wet = c->S.forceCap == 1;
    const int kern = dem_step_kernel(c);
    const bool mixed = c->precision == DEM_PREC_MIXED;
    // A pack and its write-back cost about as much as 30 steps gain from
    // the pair records, so they are made only while rebuilds come at least
    // kPairMinInterval steps apart (or their spacing is not known yet).
    int rc;
    if (kern == DEM_KERNEL_WAVE && c->pair_on && pair_fits(c) &&
        (c->rebuild_interval < 0 || c->rebuild_interval >= kPairMinInterval))
        rc = dem_pair_pack(c);
    else
        rc = dem_pair_unpack(c);
    if (rc) return rc;
    const bool pair = c->pair_packed;
    int big = INT_MAX;
    HIPCHK(hipMemcpyAsync(c->stop_at, &big, sizeof(int), hipMemcpyHostToDevice, c->stream));
    StepArgs A = dem_step_args(c, true);
    A.wsplit = dem_wave_split(c);
    if (pair) {
        A.noff = (const int *)(c->poff + 2L * c->cap);
        A.nl = c->nl;
    }
    if (pair && c->xb_work_ok) {
        for (int x = 0; x < 9; ++x) A.xb[x] = c->xb_work[x];
    } else {
        const unsigned nch = (unsigned)(c->total + 63) / 64u, per = nch / 8u, rem = nch % 8u;
        for (unsigned x = 0; x < 9; ++x) A.xb[x] = x * per + (x < rem ? x : rem);
    }
    double tw = c->Tw;
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    for (int s = 0; s < nsteps; ++s) {
        const int b = (c->cur + s) & 1, l = (c->Lidx + s) & 1;
        A.in = c->buf(b, l);
        A.out = c->buf(1 - b, 1 - l);
        A.step = s;
        tw = tw + c->S.topv * c->S.dt;
        A.Tw = tw;
        if (pair) {
            A.hq_in = c->hq[(c->hqc + s) & 1];
            A.hq_out = c->hq[(c->hqc + s + 1) & 1];
            A.hc_in = c->hc[(c->hqc + s) & 1];
            A.hc_out = c->hc[(c->hqc + s + 1) & 1];
        }
        if (kern == DEM_KERNEL_WAVE)
            launch_step_wave(c->P, A, mixed, pair, wet, c->stream);
        else if (kern == DEM_KERNEL_COMPACT)
            launch_step_compact(c->P, A, c->K, mixed, c->stream);
        else
            launch_step_fused(c->P, A, wet, mixed, c->stream);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    int stop = INT_MAX;
    HIPCHK(hipMemcpyAsync(&stop, c->stop_at, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->fused_ms += ms;
    c->fused_launches += nsteps;
    *executed = dem_finish_chunk(c, nsteps, stop);
    return DEM_OK;
}

int dem_finish_chunk(dem_ctx *c, int nsteps, int stop) {
    const int done = stop < nsteps ? stop : nsteps;
    c->fused_steps += done;
    c->cur = (c->cur + done) & 1;
    c->Lidx = (c->Lidx + done) & 1;
    if (c->pair_packed) c->hqc = (c->hqc + done) & 1;
    for (int s = 0; s < done; ++s) c->Tw = c->Tw + c->S.topv * c->S.dt;
    for (int s = 0; s < done; ++s) c->ctime = c->ctime + c->S.dt;
    c->iter += done;
    c->layout_age += done;
    if (stop <= nsteps) c->rebuild_pending = true;
    return done;
}

extern "C" int dem_step(dem_ctx *c, int nsteps) {
    if (!c || nsteps < 0) return dem_fail(DEM_ERR_ARG, "bad argument");
    int done = 0, rc;
    if (!c->rebuild_pending && (rc = dem_lists_ok(c))) return rc;
    while (done < nsteps) {
        if (c->rebuild_pending) {
            HIPCHK(hipEventRecord(c->ev0, c->stream));
            if ((rc = dem_do_rebuild(c))) return rc;
            HIPCHK(hipEventRecord(c->ev1, c->stream));
            if ((rc = check_overflow(c))) return rc;
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
            c->rebuild_ms += ms;
        } else if (c->relayout_every > 0 && c->layout_age >= c->relayout_every) {
            // with pair records the refresh is a write-back and a new pack
            // (run_chunk), which puts the contacts of now first
            if (c->pair_packed) {
                if ((rc = dem_pair_unpack(c))) return rc;
                c->n_relayout++;
            } else if ((rc = dem_relayout(c))) {
                return rc;
            }
        }
        int chunk = nsteps - done < kChunk ? nsteps - done : kChunk;
        int ex = 0;
        if ((rc = run_chunk(c, chunk, &ex))) return rc;
        done += ex;
        c->forces_valid = c->store_forces != 0 && ex > 0;
    }
    return DEM_OK;
}

extern "C" int dem_time_steps(dem_ctx *c, int nsteps, double *ms_per_step, int *rebuild_needed) {
    if (!c || nsteps <= 0) return dem_fail(DEM_ERR_ARG, "bad argument");
    int rc;
    if (c->rebuild_pending && (rc = dem_rebuild(c))) return rc;
    if ((rc = dem_lists_ok(c))) return rc;
    double ms0 = c->fused_ms;
    long long s0 = c->fused_steps, l0 = c->fused_launches;
    int ex = 0;
    if ((rc = run_chunk(c, nsteps, &ex))) return rc;
    if (ms_per_step)
        *ms_per_step = (c->fused_ms - ms0) / (double)(c->fused_launches - l0 ? c->fused_launches - l0 : 1);
    if (rebuild_needed) *rebuild_needed = (c->fused_steps - s0) < nsteps || c->rebuild_pending;
    return DEM_OK;
}

extern "C" int dem_get_timing(dem_ctx *c, double *fused_ms, long long *fused_steps,
                              long long *fused_launches, double *rebuild_ms, int *n_rebuild) {
    if (!c) return dem_fail(DEM_ERR_ARG, "null context");
    if (fused_ms) *fused_ms = c->fused_ms;
    if (fused_steps) *fused_steps = c->fused_steps;
    if (fused_launches) *fused_launches = c->fused_launches;
    if (rebuild_ms) *rebuild_ms = c->rebuild_ms;
    if (n_rebuild) *n_rebuild = c->n_rebuild;
    return DEM_OK;
}

extern "C" const char *dem_step_kernel_name(dem_ctx *c) {
    if (!c) return "";
    switch (dem_step_kernel(c)) {
    case DEM_KERNEL_WAVE:
        if (c->S.forceCap == 1)
            return c->precision == DEM_PREC_MIXED ? "k_step_wave_wet_mixed" : "k_step_wave_wet";
        return c->precision == DEM_PREC_MIXED ? "k_step_wave_mixed" : "k_step_wave";
    case DEM_KERNEL_COMPACT: return "k_step_compact";
    default: return "k_step_fused";
    }
}

extern "C" int dem_sync(dem_ctx *c) {
    if (!c) return dem_fail(DEM_ERR_ARG, "null context");
    HIPCHK(hipStreamSynchronize(c->stream));
    return DEM_OK;
}

extern "C" int dem_iter(dem_ctx *c) { return c ? c->iter : 0; }
extern "C" double dem_ctime(dem_ctx *c) { return c ? c->ctime : 0.0; }
extern "C" int dem_n_rebuild(dem_ctx *c) { return c ? c->n_rebuild : 0; }

extern "C" int dem_stream_copy(long long nbytes, int iters, double *gbps) {
    if (nbytes < 16 || iters < 1 || !gbps) return dem_fail(DEM_ERR_ARG, "bad argument");
    void *a = nullptr, *b = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = DEM_OK;
    do {
        hipError_t e;
        if ((e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking)) != hipSuccess ||
            (e = hipMalloc(&a, (size_t)nbytes)) != hipSuccess ||
            (e = hipMalloc(&b, (size_t)nbytes)) != hipSuccess ||
            (e = hipEventCreate(&e0)) != hipSuccess || (e = hipEventCreate(&e1)) != hipSuccess ||
            (e = hipMemsetAsync(a, 0, (size_t)nbytes, s)) != hipSuccess) {
            rc = dem_fail(DEM_ERR_HIP, "stream copy setup: %s", hipGetErrorString(e));
            break;
        }
        // the best of a few launch shapes (blocks per CU) is the attainable rate
        double best = 0.0;
        for (int bpc : {16, 32, 64, 128}) {
            launch_stream_copy(a, b, nbytes, bpc, s);  // warm
            (void)hipEventRecord(e0, s);
            for (int k = 0; k < iters; ++k) launch_stream_copy(a, b, nbytes, bpc, s);
            (void)hipEventRecord(e1, s);
            if ((e = hipStreamSynchronize(s)) != hipSuccess) break;
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            best = std::max(best, 2.0 * (double)nbytes * iters / (ms * 1e-3) / 1e9);
        }
        if (e != hipSuccess) {
            rc = dem_fail(DEM_ERR_HIP, "stream copy: %s", hipGetErrorString(e));
            break;
        }
        *gbps = best;
    } while (0);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    return rc;
}

extern "C" int dem_contact_stats(dem_ctx *c, dem_stats *out) {
    if (!c || !out) return dem_fail(DEM_ERR_ARG, "bad argument");
    int rc;
    if (c->rebuild_pending && (rc = dem_rebuild(c))) return rc;
    // the counts come from positions and lists; only the bridge count reads
    // the directed records, and pair records are kept for dry packings
    // only, which have no bridges: they stay packed (a write-back here
    // would cost the next step a new pack)
    const bool packed_dry = c->pair_packed && c->S.forceCap != 1;
    if (!packed_dry && (rc = dem_pair_unpack(c))) return rc;
    unsigned long long h[8] = {0};
    HIPCHK(hipMemcpyAsync(c->stats, h, sizeof h, hipMemcpyHostToDevice, c->stream));
    StepArgs A = dem_step_args(c, true);
    if (packed_dry) A.h = nullptr;  // no bridge flags to read
    launch_stats(c->P, A, c->stats, c->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(h, c->stats, sizeof h, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    out->n_particles = c->total;
    out->n_nebor = (long long)h[0];
    out->n_contact = (long long)h[1];
    out->n_wall_slots = (long long)h[2];
    out->n_wall_contact = (long long)h[3];
    out->n_bridge = (long long)h[4];
    return DEM_OK;
}
