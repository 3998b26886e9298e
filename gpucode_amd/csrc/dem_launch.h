// dem_launch.h -- host-side launchers of the kernels in dem_kernels.hip.
#pragma once
#include <hip/hip_runtime.h>

#include "dem_internal.h"

void launch_fill_i32(int *p, int v, long n, hipStream_t s);
void launch_fill_f64(double *p, double v, long n, hipStream_t s);

long scan_tmp_elems(long n);
void launch_excl_scan(const unsigned *in, unsigned *out, long n, unsigned *tmp, hipStream_t s);
long radix_tmp_elems(long n);
void launch_radix_sort(unsigned *k, unsigned *v, unsigned *k2, unsigned *v2, long n, int bits,
                       unsigned *tmp, hipStream_t s);

void launch_calc_hash(const StateBuf &s, const int *level, const int *id, const DevInst &P,
                      unsigned *keys, unsigned *vals, int total, hipStream_t st);
void launch_permute(const unsigned *perm, const StateBuf &a, const StateBuf &b, const int *lvl_a,
                    int *lvl_b, const int *id_a, int *id_b, int total, hipStream_t s);
void launch_cell_range(const unsigned *keys, const StateBuf &s, const int *level, const DevInst &P,
                       int *cstart, int *cend, int ncells, int total, hipStream_t st);
bool neighbor_tiles_apply(const DevInst &P);
void launch_neighbor_find(const StateBuf &s, const int *level, const DevInst &P, const unsigned *keys,
                          const int *cstart, const int *cend, int *nbr, int *nnb, int *overflow,
                          bool tiled, int cap, int total, hipStream_t st);
void launch_wall_list(const StateBuf &s, const DevInst &P, double Tw, int *wl, int *nwl, int cap,
                      int total, hipStream_t st);
void launch_hist_remap(const unsigned *perm, const int *id_new, const int *id_old, const int *nbr,
                       const int *nnb, const unsigned long long *am_new, const int *off_new,
                       const int *nbr_old, const int *nnb_old, const unsigned long long *am_old,
                       const int *off_old, HistRec *h, const HistRec *h_old, const int *wl,
                       const int *nwl, const int *wl_old, const int *nwl_old, HistRec *hw,
                       const HistRec *hw_old, int cap, long Kcap, int total, int old_total,
                       hipStream_t s);
void launch_active_slots(const double4 *p, const int *nbr, const int *nnb, unsigned long long *amask,
                         int *actoff, unsigned long long *nmask, double near_gap, unsigned *cnt,
                         unsigned *scan_tmp, int cap, int total, hipStream_t s);
void launch_hist_relayout(const int *nnb, const unsigned long long *am_old, const int *off_old,
                          const unsigned long long *am_new, const int *off_new,
                          const HistRec *h_old, HistRec *h, int cap, long Kcap, int total,
                          hipStream_t s);
void launch_hist_slots(HistRec *h, const int *nnb, const unsigned long long *amask, const int *actoff,
                       HistRec *slots, int cap, long Kcap, int total, bool to_slots, hipStream_t s);

void launch_step_fused(const DevInst &P, const StepArgs &A, bool wet, bool mixed, hipStream_t s);
// dry step with contact compaction; needs K <= 64
void launch_step_compact(const DevInst &P, const StepArgs &A, int K, bool mixed, hipStream_t s);
// pair: shared pair records (A.noff, A.nrec, A.hq_in, A.hq_out; see PAIR_MIRROR)
void launch_step_wave(const DevInst &P, const StepArgs &A, bool mixed, bool pair, bool wet,
                      hipStream_t s);
// shared pair records: counts (c: 4 arrays of cap), fill from the directed
// records (o: exclusive scans of the counts), write back
// work of each chunk of 64 particles (contacts: per-particle contact counts)
void launch_chunk_work(const unsigned *contacts, unsigned *work, int total, hipStream_t s);
void launch_pair_count(const double4 *p, const int *nbr, const int *nnb,
                       const unsigned long long *nmask, unsigned *c, int cap, int total,
                       hipStream_t s);
void launch_pair_fill(const double4 *p, double4 *vw, const int *nbr, const int *nnb,
                      const unsigned long long *nmask, const unsigned *o, unsigned base2,
                      const HistRec *hd, const unsigned long long *amask, const int *actoff,
                      int2 *nl, PairHot *q0, PairHot *q1, double2 *c0, double2 *c1, int *bad,
                      int cap, long Kcap, int total, hipStream_t s);
void launch_pair_unpack(const double4 *p, const int *nnb, const unsigned long long *nmask,
                        const unsigned *noff, const int2 *nl, const PairHot *q, const double2 *c,
                        HistRec *hd, const unsigned long long *amask, const int *actoff, int cap,
                        long Kcap, int total, hipStream_t s);
void launch_analyze(const DevInst *insts, const AnaArgs *args, double *out, int n, hipStream_t s);
void launch_rebuild_batch(const DevInst *insts, const RbInst *tab, int ninst, const int2 *ptasks,
                          int nptasks, const int2 *ctasks, int nctasks, hipStream_t s);
// particles per task of the batched step kernels: a wave per task for the
// wave kernel, 256 threads for the loop kernel
#define WVB_BS 64
#define LPB_BS 256
void launch_step_batch(const DevInst *insts, const BatchInst *tab, const int2 *tasks, int ntasks,
                       bool wave, bool mixed, int step, hipStream_t s);
void launch_pp_force(const DevInst &P, const StepArgs &A, bool wet, bool mixed, hipStream_t s);
void launch_pw_force(const DevInst &P, const StepArgs &A, bool mixed, hipStream_t s);
void launch_update(const DevInst &P, const StepArgs &A, int *count, hipStream_t s);

// what: 0 position, 1 velocity, 2 spin (AoS [id][3]), 3 radius, 4 path length
void launch_gather_state(const int *id, const StateBuf &s, int what, double *out, int total,
                         hipStream_t st);
void launch_gather3(const int *id, const double *a, const double *b, const double *c, double *out,
                    int total, hipStream_t s);
void launch_gather1(const int *id, const double *a, double *out, int total, hipStream_t s);
void launch_gather1i(const int *id, const int *a, int *out, int total, hipStream_t s);
void launch_scatter_state(const double *pos, const double *vel, const double *angv,
                          const double *rad, const StateBuf &s, int n, int off, hipStream_t st);
void launch_stats(const DevInst &P, const StepArgs &A, unsigned long long *out, hipStream_t s);
void launch_step_dbatch(const DevInst *insts, const DBatchInst *dins, const DBatchCtl *ctl,
                        const int2 *tasks, int ntasks, bool wave, bool mixed, int g, hipStream_t s);
void launch_rebuild_slot(const DevInst *insts, const DBatchInst *dins, const DBatchCtl *ctl,
                         DBatchCtl *nxt, RbInst *tab, int n, int g_end, int resume, const int2 *ptasks,
                         int nptasks, const int2 *ctasks, int nctasks, hipStream_t s);
void launch_commit_slot(const DBatchInst *dins, DBatchCtl *ctl, const DBatchCtl *nxt, int n,
                        hipStream_t s);
void launch_scale_velocities(const StateBuf &s, double f, int total, hipStream_t st);
void launch_stream_copy(const void *a, void *b, long nbytes, int blocks_per_cu, hipStream_t s);
