/*
 * dem_batch.h -- batched multi-instance stepping and packing analysis.
 *
 * No counterpart in the reference, which runs one packing per process with
 * its constants in process-global __constant__ symbols (kernel.cu:15-19)
 * and PARAMETER (heydem.h:196-204).  Here each instance is an ordinary
 * dem_ctx (every dem_* call of dem_capi.h works on it) whose constants and
 * list pointers also sit in per-instance device tables, so one kernel
 * launch steps all instances.  Each instance keeps its own rebuild trigger
 * (path length > 0.3 SearchGap, kernel.cu:582, 601) and takes exactly the
 * steps and rebuilds it would take alone: dem_batch_step gives results
 * bit-identical to dem_step on each instance separately.
 */
#ifndef DEM_BATCH_H
#define DEM_BATCH_H

#include "dem_capi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dem_batch dem_batch;

/* n instances; in[k] and dia[k] (may be NULL, or dia[k] NULL) as for
 * dem_create.  All instances share one HIP stream and one precision. */
int dem_batch_create(int n, const dem_input *in, const double *const *dia, int precision,
                     dem_batch **out);
int dem_batch_destroy(dem_batch *b);
int dem_batch_size(dem_batch *b);
/* The k-th instance, for dem_set_particles, dem_get_* and the rest of the
 * single-instance API.  Owned by the batch. */
dem_ctx *dem_batch_instance(dem_batch *b, int k);
/* Advances every instance by nsteps fused steps, rebuilding each instance
 * when its own trigger fires; one launch per step covers all instances. */
int dem_batch_step(dem_batch *b, int nsteps);
/* Schedule of dem_batch_step (no counterpart in the reference, which runs
 * one packing per process).  device != 0, the default: per-instance state
 * in device tables; after every slot_every global steps (default 4) a
 * rebuild slot on a second stream rebuilds the instances that asked, while
 * the next slot_every steps run for the others, and they resume after
 * that; the host reads nothing back per rebuild (applies when every
 * instance has 1 to 65,536 particles).  device = 0: chunks of steps with
 * the rebuild requests read back by the host.  Either way every instance
 * takes exactly the steps and rebuilds it takes alone. */
int dem_batch_set_schedule(dem_batch *b, int device, int slot_every);
/* dem_flush_rebuild for every instance, as one batched rebuild. */
int dem_batch_flush_rebuilds(dem_batch *b);
/* Rebuild slots run and host synchronisations of the device-driven
 * schedule so far (no counterpart in the reference). */
int dem_batch_schedule_stats(dem_batch *b, long long *slots, long long *host_syncs);
/* Event timings of dem_batch_step: step-kernel time, launches, rebuild
 * time, instance rebuilds. */
int dem_batch_timing(dem_batch *b, double *step_ms, long long *launches, double *rebuild_ms,
                     long long *n_rebuild);

/* Packing analysis of the current state, one value per instance.
 * phi: solid fraction of the core cylinder of radius R - 2 d_max around the
 * container axis, between 2 d_max above the floor and 2 d_max below the bed
 * top; the bed top is the highest z + r of the particles that touch a
 * particle or a wall.  A sphere inside the cylinder counts with its exact
 * intersection with the slab; a sphere crossing the cylinder wall counts
 * with the area of its z-slices inside the cylinder (disc-disc lens area)
 * integrated over its z extent in the slab by 32-point Gauss-Legendre
 * quadrature.
 * Z: mean number of particle-particle contacts (gap < 0) per particle.
 * top: the bed top [reduced length]. */
int dem_batch_analyze(dem_batch *b, double *phi, double *Z, double *top);
int dem_analyze(dem_ctx *ctx, double *phi, double *Z, double *top);

#ifdef __cplusplus
}
#endif
#endif
