// dem_batch.cpp -- batched stepping of many instances and packing analysis.
//
// The reference runs one packing per process: its constants are process
// globals (kernel.cu:15-19, heydem.h:196-204) and main.cpp:425-623 steps
// that one packing.  A batch holds n ordinary contexts on one stream and
// launches one step kernel per step for all dry instances (per-wave
// compaction) and one for all others (one thread per particle); the blocks
// of each launch are tasks (instance, block) listed largest instance first.
// Rebuild requests stay per instance, and each instance takes exactly the
// steps and rebuilds dem_step would give it.
//
// Two schedules.  The device-driven one (the default, for instances of at
// most 65,536 particles) keeps every instance's state in device tables: the
// host enqueues blocks of slot_every steps and, after each block, a rebuild
// slot on a second stream that rebuilds the instances that asked while the
// next block steps the others; the rebuilt instances resume after that
// block.  The host reads nothing back until the run ends (apart from a
// look every few dozen blocks to see whether every instance is through).
// The host-chunked one runs chunks of steps, reads the rebuild requests
// back, rebuilds on the host's decision and runs the next chunk.
#include <limits.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "../../include/dem_batch.h"
#include "dem_host.h"
#include "dem_launch.h"

static const int kBatchChunk = 64;
// instances up to this size are rebuilt by the batched rebuild (one block
// sorts an instance); larger ones by the multi-block single-instance path
static const int kBatchRebuildMax = 1 << 16;

struct dem_batch {
    std::vector<dem_ctx *> inst;
    hipStream_t stream = nullptr;
    int precision = DEM_PREC_FP64;
    DevInst *d_inst = nullptr;
    BatchInst *d_tab = nullptr, *h_tab = nullptr;  // h_tab pinned
    AnaArgs *d_ana = nullptr, *h_ana = nullptr;    // h_ana pinned
    double *d_out = nullptr, *h_out = nullptr;     // analysis results, h_out pinned
    int *d_stop = nullptr, *h_stop = nullptr;      // h_stop pinned
    int *h_ovf = nullptr;                          // pinned
    int *d_ovf = nullptr;                          // overflow flags of a batched rebuild's entries
    int2 *d_tasks = nullptr;
    int ntask_wave = 0, ntask_loop = 0, task_cap = 0;
    std::vector<int> task_totals;                  // totals the task table was built for
    std::vector<int> task_kernel;
    RbInst *d_rb = nullptr, *h_rb = nullptr;       // h_rb pinned
    int2 *d_rtask = nullptr;                       // particle then cell tasks of a rebuild
    long rtask_cap = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr;
    double step_ms = 0.0, rebuild_ms = 0.0;
    long long launches = 0, n_rebuild = 0;
    // device-driven schedule (on by default): steps between rebuild slots
    bool device_schedule = true;
    int slot_every = 4;
    DBatchInst *d_dins = nullptr, *h_dins = nullptr;  // h_dins pinned
    DBatchCtl *d_ctl = nullptr, *h_ctl = nullptr;     // h_ctl pinned
    DBatchCtl *d_nxt = nullptr;                       // control records a slot hands to kb_commit
    hipStream_t rstream = nullptr;                    // rebuild slots, beside the steps
    hipEvent_t evS[4] = {}, evR[4] = {};              // step-side and slot-side marks, a ring
    int2 *d_stask = nullptr;                          // static rebuild tasks: particles, then cells
    int nstask_p = 0, nstask_c = 0;
    long long slots = 0, host_syncs = 0;
};

static void batch_free(dem_batch *b) {
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->rstream) (void)hipStreamSynchronize(b->rstream);
    for (dem_ctx *c : b->inst) dem_destroy(c);
    (void)hipFree(b->d_inst);
    (void)hipFree(b->d_tab);
    (void)hipFree(b->d_ana);
    (void)hipFree(b->d_out);
    (void)hipFree(b->d_stop);
    (void)hipFree(b->d_ovf);
    (void)hipFree(b->d_tasks);
    (void)hipFree(b->d_rb);
    (void)hipFree(b->d_rtask);
    (void)hipFree(b->d_dins);
    (void)hipFree(b->d_ctl);
    (void)hipFree(b->d_nxt);
    (void)hipFree(b->d_stask);
    (void)hipHostFree(b->h_dins);
    (void)hipHostFree(b->h_ctl);
    (void)hipHostFree(b->h_rb);
    (void)hipHostFree(b->h_tab);
    (void)hipHostFree(b->h_ana);
    (void)hipHostFree(b->h_out);
    (void)hipHostFree(b->h_stop);
    (void)hipHostFree(b->h_ovf);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->ev2) (void)hipEventDestroy(b->ev2);
    for (int q = 0; q < 4; ++q) {
        if (b->evS[q]) (void)hipEventDestroy(b->evS[q]);
        if (b->evR[q]) (void)hipEventDestroy(b->evR[q]);
    }
    if (b->rstream) (void)hipStreamDestroy(b->rstream);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

extern "C" int dem_batch_create(int n, const dem_input *in, const double *const *dia, int precision,
                                dem_batch **out) {
    if (n < 1 || !in || !out) return dem_fail(DEM_ERR_ARG, "bad argument");
    *out = nullptr;
    dem_batch *b = new dem_batch();
    b->precision = precision;
    int rc = DEM_OK;
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete b;
        return dem_fail(DEM_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(e));
    }
    for (int k = 0; k < n && rc == DEM_OK; ++k) {
        dem_ctx *c = nullptr;
        rc = dem_create_on(&in[k], dia ? dia[k] : nullptr, precision, b->stream, &c);
        if (rc == DEM_OK) b->inst.push_back(c);
    }
    const size_t un = (size_t)n;
    if (rc == DEM_OK) {
        e = hipMalloc((void **)&b->d_inst, un * sizeof(DevInst));
        if (e == hipSuccess) e = hipMalloc((void **)&b->d_tab, un * sizeof(BatchInst));
        if (e == hipSuccess) e = hipMalloc((void **)&b->d_ana, un * sizeof(AnaArgs));
        if (e == hipSuccess) e = hipMalloc((void **)&b->d_out, un * 3 * sizeof(double));
        if (e == hipSuccess) e = hipMalloc((void **)&b->d_stop, un * sizeof(int));
        if (e == hipSuccess) e = hipMalloc((void **)&b->d_ovf, un * sizeof(int));
        if (e == hipSuccess) e = hipHostMalloc((void **)&b->h_tab, un * sizeof(BatchInst), 0);
        if (e == hipSuccess) e = hipHostMalloc((void **)&b->h_ana, un * sizeof(AnaArgs), 0);
        if (e == hipSuccess) e = hipHostMalloc((void **)&b->h_out, un * 3 * sizeof(double), 0);
        if (e == hipSuccess) e = hipHostMalloc((void **)&b->h_stop, un * sizeof(int), 0);
        if (e == hipSuccess) e = hipHostMalloc((void **)&b->h_ovf, un * sizeof(int), 0);
        if (e == hipSuccess) e = hipMalloc((void **)&b->d_rb, un * sizeof(RbInst));
        if (e == hipSuccess) e = hipHostMalloc((void **)&b->h_rb, un * sizeof(RbInst), 0);
        if (e == hipSuccess) e = hipMalloc((void **)&b->d_dins, un * sizeof(DBatchInst));
        if (e == hipSuccess) e = hipMalloc((void **)&b->d_ctl, un * sizeof(DBatchCtl));
        if (e == hipSuccess) e = hipHostMalloc((void **)&b->h_dins, un * sizeof(DBatchInst), 0);
        if (e == hipSuccess) e = hipHostMalloc((void **)&b->h_ctl, un * sizeof(DBatchCtl), 0);
        if (e == hipSuccess) e = hipEventCreate(&b->ev0);
        if (e == hipSuccess) e = hipEventCreate(&b->ev1);
        if (e == hipSuccess) e = hipEventCreate(&b->ev2);
        if (e == hipSuccess) e = hipMalloc((void **)&b->d_nxt, un * sizeof(DBatchCtl));
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->rstream, hipStreamNonBlocking);
        for (int q = 0; q < 4 && e == hipSuccess; ++q) {
            e = hipEventCreateWithFlags(&b->evS[q], hipEventDisableTiming);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&b->evR[q], hipEventDisableTiming);
        }
        if (e != hipSuccess) rc = dem_fail(DEM_ERR_HIP, "batch allocation: %s", hipGetErrorString(e));
    }
    if (rc != DEM_OK) {
        batch_free(b);
        return rc;
    }
    *out = b;
    return DEM_OK;
}

extern "C" int dem_batch_destroy(dem_batch *b) {
    if (b) batch_free(b);
    return DEM_OK;
}

extern "C" int dem_batch_size(dem_batch *b) { return b ? (int)b->inst.size() : 0; }

extern "C" dem_ctx *dem_batch_instance(dem_batch *b, int k) {
    if (!b || k < 0 || k >= (int)b->inst.size()) return nullptr;
    return b->inst[k];
}

// constants of every instance (they change with dem_set_material & co.)
static int upload_insts(dem_batch *b) {
    std::vector<DevInst> h(b->inst.size());
    for (size_t k = 0; k < b->inst.size(); ++k) h[k] = b->inst[k]->P;
    HIPCHK(hipMemcpyAsync(b->d_inst, h.data(), h.size() * sizeof(DevInst), hipMemcpyHostToDevice,
                          b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));  // h goes out of scope
    return DEM_OK;
}

// task table: (instance, block) for the wave-kernel instances, then for the
// loop-kernel instances; largest instance first within each group
static int build_tasks(dem_batch *b) {
    const int n = (int)b->inst.size();
    std::vector<int> tot(n), kern(n);
    for (int k = 0; k < n; ++k) {
        tot[k] = b->inst[k]->total;
        kern[k] = dem_step_kernel(b->inst[k]) == DEM_KERNEL_LOOP ? DEM_KERNEL_LOOP : DEM_KERNEL_WAVE;
    }
    if (tot == b->task_totals && kern == b->task_kernel && b->d_tasks) return DEM_OK;
    std::vector<int> order(n);
    for (int k = 0; k < n; ++k) order[k] = k;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return tot[x] > tot[y]; });
    std::vector<int2> t;
    int nw = 0;
    for (int pass = 0; pass < 2; ++pass) {
        for (int k : order) {
            const bool wave = kern[k] == DEM_KERNEL_WAVE;
            if (wave != (pass == 0)) continue;
            const int bs = wave ? WVB_BS : LPB_BS;
            const int nblk = (tot[k] + bs - 1) / bs;
            for (int q = 0; q < nblk; ++q) t.push_back(make_int2(k, q));
        }
        if (pass == 0) nw = (int)t.size();
    }
    if ((int)t.size() > b->task_cap) {
        (void)hipFree(b->d_tasks);
        b->d_tasks = nullptr;
        b->task_cap = 0;
        HIPCHK(hipMalloc((void **)&b->d_tasks, (t.size() ? t.size() : 1) * sizeof(int2)));
        b->task_cap = (int)t.size();
    }
    if (!t.empty()) {
        HIPCHK(hipMemcpyAsync(b->d_tasks, t.data(), t.size() * sizeof(int2), hipMemcpyHostToDevice,
                              b->stream));
        HIPCHK(hipStreamSynchronize(b->stream));
    }
    b->ntask_wave = nw;
    b->ntask_loop = (int)t.size() - nw;
    b->task_totals = tot;
    b->task_kernel = kern;
    return DEM_OK;
}

// table entry of a batched rebuild of c; performs the host side of the
// rebuild (the buffer flips and bookkeeping of dem_do_rebuild)
static void rebuild_entry(dem_ctx *c, int k, RbInst &R) {
    dem_cnt_hstry_save(c);
    c->old_total = c->saved ? c->old_total : 0;
    memset(&R, 0, sizeof(R));
    R.pa = c->P4[c->cur]; R.va = c->VW[c->cur]; R.lqa = c->lq[c->Lidx];
    R.pb = c->P4[1 - c->cur]; R.vb = c->VW[1 - c->cur]; R.lqb = c->lq[1 - c->Lidx];
    R.lvla = c->level[c->pcur]; R.ida = c->id[c->pcur];
    R.lvlb = c->level[1 - c->pcur]; R.idb = c->id[1 - c->pcur];
    R.keys = c->keys; R.vals = c->vals; R.keys2 = c->keys2; R.vals2 = c->vals2;
    R.passes = (c->key_bits + 7) / 8 < 1 ? 1 : (c->key_bits + 7) / 8;
    R.cstart = c->cstart; R.cend = c->cend; R.ncells = c->P.ncellTab;
    const int o = 1 - c->hcur;
    R.nbr = c->nbr[c->hcur]; R.nnb = c->nnb[c->hcur]; R.nbr_old = c->nbr[o]; R.nnb_old = c->nnb[o];
    const int ol = 1 - c->hl;
    R.hp = c->hp[c->hl]; R.hp_old = c->hp[ol];
    R.amask = c->amask[c->hl]; R.amask_old = c->amask[ol];
    R.actoff = c->actoff[c->hl]; R.actoff_old = c->actoff[ol];
    R.nmask = c->nmask;
    R.K = c->K;
    R.wl = c->wl[c->hcur]; R.nwl = c->nwl[c->hcur]; R.wl_old = c->wl[o]; R.nwl_old = c->nwl[o];
    R.hw = c->hw[c->hcur]; R.hw_old = c->hw[o];
    R.overflow = c->overflow;
    R.cap = c->cap; R.total = c->total; R.old_total = c->saved ? c->old_total : 0;
    R.inst = k;
    R.Tw = c->Tw;
    c->cur = 1 - c->cur;
    c->Lidx = 1 - c->Lidx;
    c->pcur = 1 - c->pcur;
    c->saved = false;
    c->has_lists = true;
    c->old_total = c->total;
    c->rebuild_pending = false;
    c->n_rebuild++;
    c->forces_valid = false;
    c->layout_age = 0;
}

// rebuilds every instance in need among those with steps left (the small
// ones in one batched rebuild, the large ones one by one), then checks
// their overflow flags with one synchronisation.  The entries of the
// batched rebuild flag overflow in d_ovf[entry], read back by one copy
// (not one per instance); the others use their own flag.
static int batch_rebuilds(dem_batch *b, const std::vector<int> &left) {
    const int n = (int)b->inst.size();
    std::vector<int> done, single;
    std::vector<int2> pt, ct, all;  // all: the task list copied to the device
    int nr = 0;
    HIPCHK(hipEventRecord(b->ev0, b->stream));
    for (int k = 0; k < n; ++k) {
        dem_ctx *c = b->inst[k];
        if (left[k] <= 0 || !c->rebuild_pending) continue;
        if (c->total > kBatchRebuildMax || c->total < 1) {
            int rc = dem_do_rebuild(c);
            if (rc) return rc;
            single.push_back(k);
            continue;
        }
        done.push_back(k);
        rebuild_entry(c, k, b->h_rb[nr]);
        b->h_rb[nr].overflow = b->d_ovf + nr;
        for (int q = 0; q < (c->total + 255) / 256; ++q) pt.push_back(make_int2(nr, q));
        for (int q = 0; q < (c->P.ncellTab + 255) / 256; ++q) ct.push_back(make_int2(nr, q));
        ++nr;
    }
    if (done.empty() && single.empty()) return DEM_OK;
    if (nr > 0) {
        const long need = (long)(pt.size() + ct.size());
        if (need > b->rtask_cap) {
            (void)hipFree(b->d_rtask);
            b->d_rtask = nullptr;
            b->rtask_cap = 0;
            HIPCHK(hipMalloc((void **)&b->d_rtask, (size_t)need * 2 * sizeof(int2)));
            b->rtask_cap = need * 2;
        }
        all = pt;
        all.insert(all.end(), ct.begin(), ct.end());
        HIPCHK(hipMemcpyAsync(b->d_rb, b->h_rb, (size_t)nr * sizeof(RbInst), hipMemcpyHostToDevice,
                              b->stream));
        HIPCHK(hipMemcpyAsync(b->d_rtask, all.data(), all.size() * sizeof(int2), hipMemcpyHostToDevice,
                              b->stream));
        launch_rebuild_batch(b->d_inst, b->d_rb, nr, b->d_rtask, (int)pt.size(),
                             b->d_rtask + pt.size(), (int)ct.size(), b->stream);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipEventRecord(b->ev1, b->stream));
    // flags: h_ovf[0, nr) the batched entries (done[q] is entry q), then one
    // per instance rebuilt on its own
    if (nr > 0)
        HIPCHK(hipMemcpyAsync(b->h_ovf, b->d_ovf, (size_t)nr * sizeof(int), hipMemcpyDeviceToHost,
                              b->stream));
    for (size_t q = 0; q < single.size(); ++q)
        HIPCHK(hipMemcpyAsync(&b->h_ovf[nr + q], b->inst[single[q]]->overflow, sizeof(int),
                              hipMemcpyDeviceToHost, b->stream));
    // the host vectors copied above must outlive the copies: synchronise
    HIPCHK(hipStreamSynchronize(b->stream));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    b->rebuild_ms += ms;
    b->n_rebuild += (long long)(done.size() + single.size());
    int first = -1;
    for (size_t q = 0; q < done.size() + single.size(); ++q) {
        const int k = q < done.size() ? done[q] : single[q - done.size()];
        b->inst[k]->lists_overflow = b->h_ovf[q] > 0 ? b->h_ovf[q] : 0;
        if (b->h_ovf[q] > 0 && first < 0) first = k;
    }
    if (first >= 0)
        return dem_fail(DEM_ERR_OVERFLOW,
                        "instance %d: a neighbour list needed %d slots, more than its size level "
                        "allows", first, b->inst[first]->lists_overflow);
    return DEM_OK;
}

// ---------------------------------------------------------------------
// Device-driven schedule: after the rebuilds due at the start (host path,
// one synchronisation), the host enqueues slot_every step launches and one
// rebuild slot at a time and never waits for them; the kernels keep every
// instance's segment (buffer parities, steps left, rebuild request).  An
// instance that asks for a rebuild idles until the next slot, at most
// slot_every - 1 launches.  Every 64 slots the host reads the instances'
// state back to see whether all are done.  Results are those of dem_step
// on each instance: same steps, same rebuild points.
// ---------------------------------------------------------------------
static bool device_schedule_applies(dem_batch *b) {
    if (!b->device_schedule) return false;
    for (dem_ctx *c : b->inst)
        if (c->total < 1 || c->total > kBatchRebuildMax) return false;
    return true;
}

static int static_rebuild_tasks(dem_batch *b) {
    std::vector<int2> t;
    for (int k = 0; k < (int)b->inst.size(); ++k)
        for (int q = 0; q < (b->inst[k]->total + 255) / 256; ++q) t.push_back(make_int2(k, q));
    const int np = (int)t.size();
    for (int k = 0; k < (int)b->inst.size(); ++k)
        for (int q = 0; q < (b->inst[k]->P.ncellTab + 255) / 256; ++q) t.push_back(make_int2(k, q));
    (void)hipFree(b->d_stask);
    b->d_stask = nullptr;
    HIPCHK(hipMalloc((void **)&b->d_stask, t.size() * sizeof(int2)));
    HIPCHK(hipMemcpyAsync(b->d_stask, t.data(), t.size() * sizeof(int2), hipMemcpyHostToDevice,
                          b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));  // t goes out of scope
    b->nstask_p = np;
    b->nstask_c = (int)t.size() - np;
    return DEM_OK;
}

static int batch_step_device(dem_batch *b, int nsteps) {
    const int n = (int)b->inst.size();
    const bool mixed = b->precision == DEM_PREC_MIXED;
    int rc;
    std::vector<int> left(n, nsteps);
    if ((rc = batch_rebuilds(b, left)) || (rc = build_tasks(b)) || (rc = static_rebuild_tasks(b)))
        return rc;
    for (int k = 0; k < n; ++k) {
        dem_ctx *c = b->inst[k];
        DBatchInst &D = b->h_dins[k];
        memset(&D, 0, sizeof(D));
        for (int q = 0; q < 2; ++q) {
            D.P4[q] = c->P4[q]; D.VW[q] = c->VW[q]; D.lq[q] = c->lq[q];
            D.level[q] = c->level[q]; D.id[q] = c->id[q];
            D.nbr[q] = c->nbr[q]; D.nnb[q] = c->nnb[q]; D.hp[q] = c->hp[q];
            D.amask[q] = c->amask[q]; D.actoff[q] = c->actoff[q];
            D.wl[q] = c->wl[q]; D.nwl[q] = c->nwl[q]; D.hw[q] = c->hw[q];
        }
        D.keys = c->keys; D.vals = c->vals; D.keys2 = c->keys2; D.vals2 = c->vals2;
        D.cstart = c->cstart; D.cend = c->cend; D.nmask = c->nmask;
        D.overflow = c->overflow;
        D.stop_at = b->d_stop + k;
        D.cap = c->cap; D.total = c->total; D.K = c->K; D.ncells = c->P.ncellTab;
        D.passes = (c->key_bits + 7) / 8 < 1 ? 1 : (c->key_bits + 7) / 8;
        D.dTw = c->S.topv * c->S.dt;
        DBatchCtl &C = b->h_ctl[k];
        memset(&C, 0, sizeof(C));
        C.cur = c->cur; C.Lidx = c->Lidx; C.pcur = c->pcur; C.hcur = c->hcur; C.hl = c->hl;
        C.base = 0;
        C.left = nsteps;
        C.Tw = c->Tw;
    }
    HIPCHK(hipMemcpyAsync(b->d_dins, b->h_dins, (size_t)n * sizeof(DBatchInst), hipMemcpyHostToDevice,
                          b->stream));
    HIPCHK(hipMemcpyAsync(b->d_ctl, b->h_ctl, (size_t)n * sizeof(DBatchCtl), hipMemcpyHostToDevice,
                          b->stream));
    launch_fill_i32(b->d_stop, INT_MAX, n, b->stream);
    HIPCHK(hipEventRecord(b->ev0, b->stream));
    // Blocks of R global steps on the step stream.  After block j the
    // rebuild slot j runs on the rebuild stream while block j + 1 steps the
    // other instances; then the step stream waits for the slot, kb_commit
    // hands the rebuilt instances their new control records, and they
    // resume with block j + 2.  Nothing returns to the host in between; it
    // looks at the control records only to see whether every instance is
    // through: first once nsteps global steps have run, then no earlier
    // than the instance furthest from its end could finish (it takes at
    // most one step per global step), and every 256 blocks.
    const int R = b->slot_every < 1 ? 1 : b->slot_every;
    // an instance that rebuilds after every step takes 2R global steps per step
    const long long gmax = ((long long)nsteps + 1) * 2 * R;
    long long G = 0, slots = 0, blocks = 0, last_check = 0, next_check = nsteps;
    int pend = -1;
    bool done = nsteps == 0;
    while (!done) {
        for (int r = 0; r < R; ++r, ++G) {
            launch_step_dbatch(b->d_inst, b->d_dins, b->d_ctl, b->d_tasks, b->ntask_wave, true, mixed,
                               (int)G, b->stream);
            launch_step_dbatch(b->d_inst, b->d_dins, b->d_ctl, b->d_tasks + b->ntask_wave,
                               b->ntask_loop, false, mixed, (int)G, b->stream);
        }
        if (pend >= 0) {
            HIPCHK(hipStreamWaitEvent(b->stream, b->evR[pend], 0));
            launch_commit_slot(b->d_dins, b->d_ctl, b->d_nxt, n, b->stream);
            pend = -1;
        }
        HIPCHK(hipGetLastError());
        ++blocks;
        if (G >= next_check || blocks - last_check >= 256 || G >= gmax) {
            last_check = blocks;
            HIPCHK(hipMemcpyAsync(b->h_ctl, b->d_ctl, (size_t)n * sizeof(DBatchCtl),
                                  hipMemcpyDeviceToHost, b->stream));
            HIPCHK(hipMemcpyAsync(b->h_stop, b->d_stop, (size_t)n * sizeof(int), hipMemcpyDeviceToHost,
                                  b->stream));
            HIPCHK(hipStreamSynchronize(b->stream));
            ++b->host_syncs;
            done = true;
            int most = 0;  // steps left of the instance furthest from its end
            for (int k = 0; k < n; ++k) {
                const DBatchCtl &C = b->h_ctl[k];
                const int seen = (int)G - C.base, st = b->h_stop[k];
                int seg = st < seen ? st : seen;
                seg = seg < C.left ? seg : C.left;
                // still running, or rebuilding at the next slot
                if (!C.failed && seg < C.left) {
                    done = false;
                    most = C.left - seg > most ? C.left - seg : most;
                }
            }
            if (done) break;
            next_check = G + (most > R ? most : R);
            if (G >= gmax) return dem_fail(DEM_ERR_STATE, "batched run made no progress");
        }
        const int q = (int)(slots & 3);
        HIPCHK(hipEventRecord(b->evS[q], b->stream));
        HIPCHK(hipStreamWaitEvent(b->rstream, b->evS[q], 0));
        launch_rebuild_slot(b->d_inst, b->d_dins, b->d_ctl, b->d_nxt, b->d_rb, n, (int)G - 1,
                            (int)G + R, b->d_stask, b->nstask_p, b->d_stask + b->nstask_p, b->nstask_c,
                            b->rstream);
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(b->evR[q], b->rstream));
        pend = q;
        ++slots;
    }
    HIPCHK(hipEventRecord(b->ev1, b->stream));
    HIPCHK(hipMemcpyAsync(b->h_ctl, b->d_ctl, (size_t)n * sizeof(DBatchCtl), hipMemcpyDeviceToHost,
                          b->stream));
    HIPCHK(hipMemcpyAsync(b->h_stop, b->d_stop, (size_t)n * sizeof(int), hipMemcpyDeviceToHost,
                          b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    b->step_ms += ms;
    b->launches += 2LL * G;
    b->slots += slots;
    int first_bad = -1;
    for (int k = 0; k < n; ++k) {
        dem_ctx *c = b->inst[k];
        const DBatchCtl &C = b->h_ctl[k];
        int seen = (int)G - C.base;
        seen = seen < 0 ? 0 : seen;
        const int st = b->h_stop[k];
        int seg = st < seen ? st : seen;
        seg = seg < C.left ? seg : C.left;
        if (C.failed) seg = 0;  // halted at the rebuild that overflowed
        const int ex = C.executed + seg;
        c->cur = (C.cur + seg) & 1;
        c->Lidx = (C.Lidx + seg) & 1;
        c->pcur = C.pcur;
        c->hcur = C.hcur;
        c->hl = C.hl;
        double tw = C.Tw;
        for (int q = 0; q < seg; ++q) tw = tw + c->S.topv * c->S.dt;
        c->Tw = tw;
        for (int q = 0; q < ex; ++q) c->ctime = c->ctime + c->S.dt;
        c->iter += ex;
        c->fused_steps += ex;
        if (C.nreb > 0) {
            c->n_rebuild += C.nreb;
            b->n_rebuild += C.nreb;
            c->rebuild_interval = C.seg_steps;
            c->iter_at_rebuild = c->iter - seg;
            c->saved = false;
            c->has_lists = true;
            c->old_total = c->total;
            c->layout_age = seg;
        } else {
            c->layout_age += ex;
        }
        c->forces_valid = false;
        c->rebuild_pending = st <= seen;  // asked in its last segment, no step left to rebuild for
        if (C.failed > 0) {
            c->lists_overflow = C.failed;
            if (first_bad < 0) first_bad = k;
        }
    }
    if (first_bad >= 0)
        return dem_fail(DEM_ERR_OVERFLOW,
                        "instance %d: a neighbour list needed %d slots, more than its size level "
                        "allows", first_bad, b->inst[first_bad]->lists_overflow);
    return DEM_OK;
}

extern "C" int dem_batch_flush_rebuilds(dem_batch *b) {
    if (!b) return dem_fail(DEM_ERR_ARG, "null batch");
    const int n = (int)b->inst.size();
    for (int k = 0; k < n; ++k) {
        dem_ctx *c = b->inst[k];
        if (c->rebuild_pending) {
            int rc = dem_pair_unpack(c);
            if (rc) return rc;
        }
    }
    int rc = upload_insts(b);
    if (rc) return rc;
    return batch_rebuilds(b, std::vector<int>(n, 1));
}

extern "C" int dem_batch_set_schedule(dem_batch *b, int device, int slot_every) {
    if (!b || slot_every < 1) return dem_fail(DEM_ERR_ARG, "bad argument");
    b->device_schedule = device != 0;
    b->slot_every = slot_every;
    return DEM_OK;
}

extern "C" int dem_batch_schedule_stats(dem_batch *b, long long *slots, long long *host_syncs) {
    if (!b) return dem_fail(DEM_ERR_ARG, "null batch");
    if (slots) *slots = b->slots;
    if (host_syncs) *host_syncs = b->host_syncs;
    return DEM_OK;
}

extern "C" int dem_batch_step(dem_batch *b, int nsteps) {
    if (!b || nsteps < 0) return dem_fail(DEM_ERR_ARG, "bad argument");
    const int n = (int)b->inst.size();
    int rc;
    for (int k = 0; k < n; ++k) {
        dem_ctx *c = b->inst[k];
        if (c->precision != b->precision) return dem_fail(DEM_ERR_STATE, "mixed precisions in batch");
        if (!c->rebuild_pending && c->lists_overflow > 0)
            return dem_fail(DEM_ERR_OVERFLOW, "instance %d: a neighbour list needed %d slots, more "
                            "than its size level allows", k, c->lists_overflow);
        // the batched kernels step the directed records
        if ((rc = dem_pair_unpack(c))) return rc;
    }
    if ((rc = upload_insts(b))) return rc;
    if (device_schedule_applies(b)) return batch_step_device(b, nsteps);
    std::vector<int> left(n, nsteps);
    const bool mixed = b->precision == DEM_PREC_MIXED;
    // chunk length: an instance that requested a rebuild idles for the rest
    // of its chunk, so the chunk follows the run lengths seen last time
    // (the 1st quartile over the running instances, between 8 and 64: on
    // 128 wet pour-and-settle packings 5 % faster than the 3rd quartile,
    // the idle launches saved outweighing the extra host round trips);
    // chunking does not change any instance's sequence of steps
    std::vector<int> last_run(n, kBatchChunk);
    for (;;) {
        bool any = false;
        for (int k = 0; k < n; ++k) any = any || left[k] > 0;
        if (!any) break;
        if ((rc = batch_rebuilds(b, left))) return rc;
        if ((rc = build_tasks(b))) return rc;
        std::vector<int> runs;
        for (int k = 0; k < n; ++k)
            if (left[k] > 0) runs.push_back(last_run[k]);
        std::sort(runs.begin(), runs.end());
        int chunk = runs.empty() ? kBatchChunk : runs[runs.size() / 4];
        chunk = chunk < 8 ? 8 : (chunk > kBatchChunk ? kBatchChunk : chunk);
        int S = 0;
        for (int k = 0; k < n; ++k) {
            dem_ctx *c = b->inst[k];
            BatchInst &T = b->h_tab[k];
            memset(&T, 0, sizeof(T));
            T.A = dem_step_args(c, true);
            T.A.stop_at = b->d_stop + k;
            for (int q = 0; q < 2; ++q) {
                T.P4[q] = c->P4[q];
                T.VW[q] = c->VW[q];
                T.lq[q] = c->lq[q];
            }
            T.cur = c->cur;
            T.Lidx = c->Lidx;
            T.limit = left[k] < chunk ? (left[k] > 0 ? left[k] : 0) : chunk;
            T.Tw0 = c->Tw;
            T.dTw = c->S.topv * c->S.dt;
            S = T.limit > S ? T.limit : S;
        }
        HIPCHK(hipMemcpyAsync(b->d_tab, b->h_tab, (size_t)n * sizeof(BatchInst), hipMemcpyHostToDevice,
                              b->stream));
        launch_fill_i32(b->d_stop, INT_MAX, n, b->stream);
        HIPCHK(hipEventRecord(b->ev0, b->stream));
        for (int s = 0; s < S; ++s) {
            launch_step_batch(b->d_inst, b->d_tab, b->d_tasks, b->ntask_wave, true, mixed, s, b->stream);
            launch_step_batch(b->d_inst, b->d_tab, b->d_tasks + b->ntask_wave, b->ntask_loop, false, mixed,
                              s, b->stream);
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(b->ev1, b->stream));
        HIPCHK(hipMemcpyAsync(b->h_stop, b->d_stop, (size_t)n * sizeof(int), hipMemcpyDeviceToHost,
                              b->stream));
        HIPCHK(hipStreamSynchronize(b->stream));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
        b->step_ms += ms;
        b->launches += S;
        for (int k = 0; k < n; ++k) {
            const int lim = b->h_tab[k].limit;
            if (lim <= 0) continue;
            dem_ctx *c = b->inst[k];
            const int ex = dem_finish_chunk(c, lim, b->h_stop[k]);
            c->forces_valid = c->store_forces != 0 && ex > 0;
            left[k] -= ex;
            last_run[k] = c->rebuild_pending ? ex + 1 : 2 * lim;
        }
    }
    return DEM_OK;
}

extern "C" int dem_batch_timing(dem_batch *b, double *step_ms, long long *launches, double *rebuild_ms,
                                long long *n_rebuild) {
    if (!b) return dem_fail(DEM_ERR_ARG, "null batch");
    if (step_ms) *step_ms = b->step_ms;
    if (launches) *launches = b->launches;
    if (rebuild_ms) *rebuild_ms = b->rebuild_ms;
    if (n_rebuild) *n_rebuild = b->n_rebuild;
    return DEM_OK;
}

// ---------------------------------------------------------------------
// analysis
// ---------------------------------------------------------------------
static int analyze(const std::vector<dem_ctx *> &cs, hipStream_t st, DevInst *d_inst, AnaArgs *d_ana,
                   AnaArgs *h_ana, double *d_out, double *h_out, double *phi, double *Z, double *top) {
    const int n = (int)cs.size();
    std::vector<DevInst> hi(n);
    std::vector<int> rebuilt;
    for (int k = 0; k < n; ++k) {
        dem_ctx *c = cs[k];
        if (c->rebuild_pending || !c->has_lists) {
            int rc = dem_do_rebuild(c);  // lists of the current positions
            if (rc) return rc;
            rebuilt.push_back(k);
        }
    }
    for (int k : rebuilt) {  // a truncated list would undercount Z
        int ov = 0;
        HIPCHK(hipMemcpyAsync(&ov, cs[k]->overflow, sizeof(int), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        cs[k]->lists_overflow = ov > 0 ? ov : 0;
        if (ov > 0)
            return dem_fail(DEM_ERR_OVERFLOW, "instance %d: a neighbour list needed %d slots", k, ov);
    }
    for (int k = 0; k < n; ++k) {
        dem_ctx *c = cs[k];
        hi[k] = c->P;
        AnaArgs &G = h_ana[k];
        G.p = c->P4[c->cur];
        G.nbr = c->nbr[c->hcur];
        G.nnb = c->nnb[c->hcur];
        G.wl = c->wl[c->hcur];
        G.nwl = c->nwl[c->hcur];
        G.cap = c->cap;
        G.total = c->total;
        G.Tw = c->Tw;
    }
    HIPCHK(hipMemcpyAsync(d_inst, hi.data(), (size_t)n * sizeof(DevInst), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d_ana, h_ana, (size_t)n * sizeof(AnaArgs), hipMemcpyHostToDevice, st));
    launch_analyze(d_inst, d_ana, d_out, n, st);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(h_out, d_out, (size_t)n * 3 * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    for (int k = 0; k < n; ++k) {
        if (phi) phi[k] = h_out[3 * k];
        if (Z) Z[k] = h_out[3 * k + 1];
        if (top) top[k] = h_out[3 * k + 2];
    }
    return DEM_OK;
}

extern "C" int dem_batch_analyze(dem_batch *b, double *phi, double *Z, double *top) {
    if (!b) return dem_fail(DEM_ERR_ARG, "null batch");
    return analyze(b->inst, b->stream, b->d_inst, b->d_ana, b->h_ana, b->d_out, b->h_out, phi, Z, top);
}

extern "C" int dem_analyze(dem_ctx *c, double *phi, double *Z, double *top) {
    if (!c) return dem_fail(DEM_ERR_ARG, "null context");
    DevInst *d_inst = nullptr;
    AnaArgs *d_ana = nullptr;
    double *d_out = nullptr;
    AnaArgs h_ana;
    double h_out[3];
    int rc = DEM_OK;
    hipError_t e = hipMalloc((void **)&d_inst, sizeof(DevInst));
    if (e == hipSuccess) e = hipMalloc((void **)&d_ana, sizeof(AnaArgs));
    if (e == hipSuccess) e = hipMalloc((void **)&d_out, 3 * sizeof(double));
    if (e != hipSuccess) rc = dem_fail(DEM_ERR_HIP, "analysis allocation: %s", hipGetErrorString(e));
    if (rc == DEM_OK)
        rc = analyze(std::vector<dem_ctx *>{c}, c->stream, d_inst, d_ana, &h_ana, d_out, h_out, phi, Z,
                     top);
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(d_inst);
    (void)hipFree(d_ana);
    (void)hipFree(d_out);
    return rc;
}
