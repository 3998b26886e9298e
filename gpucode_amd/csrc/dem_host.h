// dem_host.h -- host-side context of the DEM engine.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/dem_capi.h"
#include "dem_internal.h"

int dem_fail(int code, const char *fmt, ...);
void dem_calpara_from_material(dem_setup *S);
extern "C" int dem_read_deck(const char *path, dem_input *in);

#define HIPCHK(call)                                                                    \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess)                                                           \
            return dem_fail(DEM_ERR_HIP, "%s:%d %s: %s", __FILE__, __LINE__, #call,      \
                            hipGetErrorString(e_));                                     \
    } while (0)

struct dem_ctx {
    dem_input in;
    dem_setup S;
    int precision = DEM_PREC_FP64;
    int cap = 0, total = 0, K = 0;
    hipStream_t stream = nullptr;
    bool own_stream = true;      // false when a batch lends its stream
    DevInst P;
    double Tw = 0.0;
    std::vector<double> rad_h;   // reduced radius by original id
    std::vector<int> level_h;    // size level by original id

    // particle records, cell-sorted, two buffers
    double4 *P4[2] = {};         // (x, y, z, r)
    double4 *VW[2] = {};         // (vx, vy, vz, wx), (wy, wz, pmove, -)
    int cur = 0;
    int *lq[2] = {};             // bridge counts: lq[Lidx] = lqdCnt, other = lqdCnt_Old
    int Lidx = 0;
    int *level[2] = {}, *id[2] = {};
    int pcur = 0;

    // lists and wall history, current = [hcur], old = [1-hcur]; the
    // particle-pair history and its layout, current = [hl], old = [1-hl]
    // (both flip at a rebuild, hl alone at a layout refresh)
    int *nbr[2] = {}, *nnb[2] = {};
    int *wl[2] = {}, *nwl[2] = {};
    HistRec *hp[2] = {}, *hw[2] = {};   // hp: compact region K*cap, then slot-major K*cap
    unsigned long long *amask[2] = {};  // compact-record slots of each particle
    int *actoff[2] = {};                // start of each particle's compact records
    unsigned long long *nmask = nullptr;  // slots that can touch before the next rebuild
    int hcur = 0, hl = 0;
    // shared pair records of the wave kernel (PAIR_MIRROR in dem_internal.h),
    // made from the directed records before a chunk and written back before
    // anything else reads the history
    bool pair_on = true;         // use them when dem_step runs the wave kernel
    bool pair_bad = false;       // a mirror record did not match: directed records only
    bool pair_packed = false;    // the pair records hold the current history
    int hqc = 0;                 // the pair buffer the next step reads
    PairHot *hq[2] = {};         // hot part of the pair records, two buffers
    double2 *hc[2] = {};         // their plstRad/plstDfm where not implied
    long hq_cap = 0;             // records per pair buffer
    int2 *nl = nullptr;          // near lists: (partner, pair record | PAIR_MIRROR)
    long nl_cap = 0;
    unsigned *pcnt = nullptr, *poff = nullptr;  // 3 x cap counts and their scans (+1: near total)
    int *pair_err = nullptr;
    int n_pack = 0;
    unsigned xb_work[9] = {};    // wave kernel: XCD ranges of chunks cut for equal work at the pack
    bool xb_work_ok = false;
    int relayout_every = 2048;   // steps between layout refreshes (0: never)
    long long layout_age = 0;    // steps since the layout was made
    int n_relayout = 0;
    bool has_lists = false, saved = false;
    int old_total = 0;           // particles covered by the old lists

    // search cells and sort scratch
    int *cstart = nullptr, *cend = nullptr;
    unsigned *keys = nullptr, *vals = nullptr, *keys2 = nullptr, *vals2 = nullptr, *sort_tmp = nullptr;
    int key_bits = 1;

    // force arrays (separate-kernel path and optional output of the fused path)
    double *F[7] = {};           // Fx Fy Fz Mx My Mz Mr
    int store_forces = 0;
    int variant = DEM_KERNEL_AUTO;  // fused step kernel choice
    int wave_split = 0;             // wave kernel sub-waves per chunk (0: one, or 1, 2, 4)
    bool want_morton = true;        // Morton storage order when the grid allows

    // device flags and counters
    int *stop_at = nullptr, *overflow = nullptr, *upd_count = nullptr;
    unsigned long long *stats = nullptr;

    // host bookkeeping (main.cpp:425-623)
    int iter = 0, n_rebuild = 0;
    int iter_at_rebuild = -1;     // iter of the last rebuild
    int rebuild_interval = -1;    // steps between the last two rebuilds (-1: not known)
    double ctime = 0.0;
    bool rebuild_pending = true;
    // > 0 after a rebuild whose lists overflowed (the slots a list needed):
    // the lists are truncated, so stepping is refused until the state is
    // replaced or a rebuild fits
    int lists_overflow = 0;
    bool forces_valid = false;

    // timing of the fused path
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // rebuild stage profile (dem_profile_rebuild): when on, an event after
    // each stage of a rebuild
    static const int kRbStages = 9;
    hipEvent_t rb_ev[kRbStages + 1] = {};
    bool rb_prof = false;
    // neighbour search form: nf_mode DEM_NSEARCH_*; under AUTO the first two
    // rebuilds time the tiled and the plain form (nf_ms) and nf_choice keeps
    // the faster (-1 while undecided), for 256 rebuilds
    int nf_mode = 0, nf_choice = -1, nf_trials = 0, nf_timed = 0, nf_since = 0;
    bool nf_pending = false;
    float nf_ms[2] = {0.f, 0.f};
    hipEvent_t nf_ev[2] = {};
    double fused_ms = 0.0, rebuild_ms = 0.0;
    long long fused_steps = 0, fused_launches = 0;
    long long bytes = 0;

    StateBuf buf(int b, int lqd_idx) const {
        StateBuf s;
        s.p = P4[b];
        s.v = VW[b];
        s.lqd = lq[lqd_idx];
        return s;
    }
};

// DevInst from the derived setup (one instance)
void dem_fill_inst(const dem_setup &S, DevInst &P);
// dem_fill_inst plus the context's storage order and sort-key width
void dem_refresh_inst(dem_ctx *c);
// Internal pieces of the rebuild and step, shared by the C ABI functions.
int dem_do_rebuild(dem_ctx *c);
int dem_lists_ok(dem_ctx *c);
// shared pair records: make them from the directed records, or write them
// back (a no-op when they are not in use)
int dem_pair_pack(dem_ctx *c);
int dem_pair_unpack(dem_ctx *c);
// frees the pair buffers
void dem_pair_free(dem_ctx *c);
// StepArgs of the context's current buffers (fused: read cur, write 1-cur)
StepArgs dem_step_args(dem_ctx *c, bool fused);
// DEM_KERNEL_LOOP, _COMPACT or _WAVE: the kernel dem_step launches
int dem_step_kernel(const dem_ctx *c);
int dem_wave_split(const dem_ctx *c);
// host bookkeeping after a chunk of nsteps launches whose device stop
// index read back as stop; returns the steps executed
int dem_finish_chunk(dem_ctx *c, int nsteps, int stop);
// dem_create on a stream owned by the caller
int dem_create_on(const dem_input *in, const double *dia_m, int precision, hipStream_t stream,
                  dem_ctx **out);
