/*
 * dem_tune.h -- autotuning of the neighbour-list skin, step kernel and
 * storage order for a packing.
 *
 * No counterpart in the reference, whose skin (CUTGAPFORNEIGHBORLIST,
 * ReadData.cpp:363) and cell size (ReadData.cpp:369-375) are fixed user
 * inputs.  A larger skin means longer lists (more work per step) and fewer
 * rebuilds (the trigger is a path length of 0.3 SearchGap, kernel.cu:582);
 * the best trade depends on how fast the packing moves.  The contact laws do
 * not depend on the skin, but the cell size and the list contents do, so a
 * tuned run sums contacts in another order than an untuned one: results
 * agree to rounding, not bit for bit.
 */
#ifndef DEM_TUNE_H
#define DEM_TUNE_H

#include "dem_capi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dem_tune_result {
    double cutGap;       /* chosen CUTGAPFORNEIGHBORLIST [d_max] */
    int kernel;          /* chosen DEM_KERNEL_* */
    int morton;          /* chosen storage order (dem_set_order) */
    double ms_per_step;  /* its time per step, rebuilds included */
    double ms_default;   /* the same for the input's own cutGap, AUTO kernel, Morton order */
    int n_tried;         /* candidates timed */
} dem_tune_result;

/* Times `steps` steps of the packing (in, dia; n particles at pos[n][3],
 * velocities vel or NULL) for every candidate: skins cutGaps[ncut] (NULL:
 * 0.05, 0.1, 0.2, 0.3, 0.5), each applicable step kernel (WAVE and COMPACT
 * for dry packings with at most 64 slots, LOOP always) and both storage
 * orders.  Skins whose search gap is below the largest capillary rupture
 * distance of a wet packing are skipped (bridges would be lost), and so
 * are skins whose lists would overflow.  table, if not NULL, receives
 * ncut x 3 x 2 times [ms per step], -1 where skipped, indexed
 * [skin][kernel LOOP/COMPACT/WAVE][order row-major/Morton]. */
int dem_autotune(const dem_input *in, const double *dia, int n, const double *pos, const double *vel,
                 int steps, const double *cutGaps, int ncut, int precision, dem_tune_result *best,
                 double *table);

#ifdef __cplusplus
}
#endif
#endif
