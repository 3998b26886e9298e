/*
 * dem_capi.h -- C ABI of the MI355X-native DEM time-step engine (gfx950).
 *
 * The engine replaces the reference's device layer: the extern "C" launch
 * wrappers of ParticleSystem.cu:18-382 (declared heydem.h:547-613), the
 * constant setters of kernel.cu:22-50 (heydem.h:453-457) and the allocation
 * helpers of Allocation.cpp (heydem.h:366-402).  Differences of form, kept
 * deliberately:
 *   - all state lives in an opaque dem_ctx (no process globals, no
 *     __constant__ symbols), so several contexts and batched instances can
 *     coexist;
 *   - particle state is held as packed per-particle records (position and
 *     radius; velocity, spin and path length) permuted into cell order at
 *     every rebuild; double3 arrays in original particle order appear only
 *     at the explicit import/export calls below;
 *   - every call returns a status code (DEM_OK = 0) instead of exit();
 *     dem_last_error() gives the message;
 *   - each context owns one HIP stream; calls on one context must come from
 *     one host thread at a time.
 * All lengths, times and forces at this boundary are in the reference's
 * reduced units (UnitReduce.cpp:14-20: length = largest diameter,
 * time = sqrt(d/g), force = m g) unless a name ends in _si.
 */
#ifndef DEM_CAPI_H
#define DEM_CAPI_H

#ifdef __cplusplus
extern "C" {
#endif

#define DEM_MAX_LEVELS 8
#define DEM_WALLS 3

enum {
    DEM_OK = 0,
    DEM_ERR_ARG = -1,      /* bad argument or size */
    DEM_ERR_HIP = -2,      /* a HIP runtime call failed */
    DEM_ERR_OVERFLOW = -3, /* a neighbour list exceeded its level's nNebors */
    DEM_ERR_LEVEL = -4,    /* a radius fell outside every size level */
    DEM_ERR_NOMEM = -5,
    DEM_ERR_STATE = -6     /* call out of order */
};

/* Pair-force arithmetic.  State (x, v, omega, contact history) is FP64 in
 * both modes. */
enum { DEM_PREC_FP64 = 0, DEM_PREC_MIXED = 1 };

/* Fused step kernel: AUTO picks the per-wave contact-compacting kernel for
 * dry packings whose lists have at most 64 slots, the one-thread-per-
 * particle loop otherwise.  All variants give identical results. */
enum { DEM_KERNEL_AUTO = 0, DEM_KERNEL_LOOP = 1, DEM_KERNEL_COMPACT = 2, DEM_KERNEL_WAVE = 3 };

/* The heyi.in keywords (ReadData.cpp:75-188), SI units, plus overrides. */
typedef struct dem_input {
    int NP;                 /* NUMPARTICLE (capacity) */
    int nLevels;            /* SIZELEVEL */
    int loadStage;          /* LOADMODE (1 = packing) */
    int forceCap;           /* FoceCapillary: 1 = capillary bridges on */
    double tPacking;        /* packing_TIME [s] */
    double realSize;        /* REALSIZE: largest diameter [m] */
    double cutGap;          /* CUTGAPFORNEIGHBORLIST [d_max] */
    double expandSize;      /* EXPANDSIZE [r_max] */
    double cirX, cirY, Bw, R, H; /* CylinderBC [m] */
    double feedLmin, feedLmax;   /* FeedZone [m] */
    double mat[2][7];       /* MATERIAL rows (particle, wall):
                               density ymod pois sfrc rfrc dmpn yldp */
    double stension, vol, gapmn; /* LIQUID */
    double dtFactor;        /* SIMULATION */
    int nOutPacking;        /* SIMULATION */
    int sizeLevelInput;     /* SIZELEVEL_INPUT: 1 = levelDia given */
    double levelDia[DEM_MAX_LEVELS];
    double gravity;         /* [m/s^2]; 0 selects the reference's 9.8f */
    double dtOverride;      /* [s]; 0 = ReadData.cpp:260-263 formula */
    double theta;           /* liquid contact angle [rad]; 0 = reference law */
    int ruptureMode;        /* 0: 0.99 V^0.33 (kernel.cu:522);
                               1: (1 + theta/2) V^(1/3) */
    int gridClampQuirk;     /* must be 0: dem_create refuses 1 with
                               DEM_ERR_ARG.  kernel.cu:627-629 clamps cell
                               coordinates to gridSize, which sends a particle
                               outside the box past the cell table; here cells
                               clamp to gridSize-1 (identical for every
                               particle inside the box) */
} dem_input;

/* Derived quantities in reduced units (ReadData.cpp:243-375,
 * UnitReduce.cpp:14-104, main.cpp:328-348). */
typedef struct dem_calpara {
    double emod, pois, sfrc, ystress, ydsp_t, yforce_t, ystiff_t;
} dem_calpara;

typedef struct dem_setup {
    double rlunit, rmunit, rfunit, rsunit, rtunit, rvunit, reunit;
    double dt, searchGap;
    double origin[3], endPoint[3];
    double cirX, cirY, R, Tw, Bw, topv, btmv;
    double feedLmin, feedLmax, tPacking;
    double density[2], emod[2], ymod[2], pois[2], sfrc[2], rfrc[2], dmpn[2], yldp[2];
    double stension, vol, gapmn, layer, brknmx;
    double theta;
    dem_calpara pp, pw;
    int nLevels, forceCap, loadStage, ruptureMode, gridClampQuirk, nPPCnt;
    double levelSize[DEM_MAX_LEVELS];
    int levelNNebors[DEM_MAX_LEVELS];
    int levelNPars[DEM_MAX_LEVELS];
    int levelNebSId[DEM_MAX_LEVELS];
    double cellSize[DEM_MAX_LEVELS];
    int gridSize[DEM_MAX_LEVELS][3];
    int numCells[DEM_MAX_LEVELS];
    int ncellIdx[DEM_MAX_LEVELS];
    int sysNCells;
    double radMin, radMax;
} dem_setup;

/* Contact counts of the current state (for the roofline bytes). */
typedef struct dem_stats {
    long long n_particles, n_nebor, n_contact, n_wall_slots, n_wall_contact;
    long long n_bridge;
} dem_stats;

typedef struct dem_ctx dem_ctx;

const char *dem_last_error(void);

/* ---- setup ---------------------------------------------------------- */
/* Host-side derivation only (ReadData + UnitReduce + contact constants).
 * dia_m: NP diameters [m] or NULL (all = realSize).  rad, pLevel, neborIdx
 * (NP each) receive reduced radii, size levels and in-level indices. */
int dem_derive(const dem_input *in, const double *dia_m, dem_setup *out,
               double *rad, int *pLevel, int *neborIdx);

/* Allocates a context for one packing: replaces ParticleInfoAlloc
 * (Allocation.cpp:22-53), PPCntHstry_Alloc (115-174), WallCntHstry_Alloc
 * (57-110), NeborListAlloc (181-235), ParamAlloc (251-287),
 * SearchCellAlloc (291-303) and the constant uploads setMaterial/setLiquid/
 * setParameters/setCalpara (kernel.cu:22-50, main.cpp:356-359). */
int dem_create(const dem_input *in, const double *dia_m, int precision, dem_ctx **out);
int dem_destroy(dem_ctx *ctx);
int dem_get_setup(dem_ctx *ctx, dem_setup *out);

/* Constant setters, per context (kernel.cu:22-50).  mat: 2 rows of
 * density ymod pois sfrc rfrc dmpn yldp, reduced units (after UnitReduce);
 * the contact constants are re-derived as in main.cpp:328-348. */
int dem_set_material(dem_ctx *ctx, const double mat[2][7]);
/* stension, vol (volume fraction * pi/6, i.e. after UnitReduce), gapmn,
 * layer; theta [rad]; ruptureMode */
int dem_set_liquid(dem_ctx *ctx, double stension, double vol, double gapmn,
                   double layer, double theta, int ruptureMode, int forceCap);
/* dt and SearchGap (PARAMETER, heydem.h:196-204) */
int dem_set_parameters(dem_ctx *ctx, double dt, double searchGap);

/* ---- particle state (import / export; AoS [n][3], original ids) ------ */
/* Sets the first `total` particles (ids 0..total-1); clears lists and
 * history; the next step rebuilds.  vel/angv may be NULL (zero). */
int dem_set_particles(dem_ctx *ctx, int total, const double *pos, const double *vel,
                      const double *angv);
/* Appends n particles (ids total..total+n-1) at rest, as ParticleFeed does
 * (main.cpp:443-448); the next step rebuilds (FeedFlag, main.cpp:473). */
int dem_feed(dem_ctx *ctx, int n, const double *pos);
int dem_total(dem_ctx *ctx);
int dem_get_particles(dem_ctx *ctx, double *pos, double *vel, double *angv);
/* F, M, Mr of the last step (dForce/dMom/dRMom, after UpdateSystemG has
 * scaled Mr by dt/I as kernel.cu:591 does).  Needs dem_set_store_forces(1)
 * for the fused dem_step path. */
int dem_get_forces(dem_ctx *ctx, double *F, double *M, double *Mr);
int dem_set_store_forces(dem_ctx *ctx, int on);
int dem_get_pmove(dem_ctx *ctx, double *pmove);
int dem_get_lqdcnt(dem_ctx *ctx, int *lqdCnt);

/* NeborToCPU (ParticleSystem.cu:296-334): lists as original ids in slot
 * order, num[total], list[total][maxK] (-1 past num). */
int dem_get_nebor(dem_ctx *ctx, int *num, int *list, int maxK);
int dem_get_hist_pp(dem_ctx *ctx, int maxK, double *dispt, double *nDsp,
                    double *nDsp_mx, double *plstRad, double *plstDfm, int *stage,
                    int *lqdBrg);
int dem_get_walls(dem_ctx *ctx, int *num, int *list /* [total][3] */);
int dem_get_hist_pw(dem_ctx *ctx, double *dispt, double *nDsp, double *nDsp_mx,
                    double *plstRad, double *plstDfm, int *stage, int *lqdBrg);
/* Restores lists + history saved by the getters above (restart, DumpLoad);
 * the next step rebuilds and remaps it (RESTART, main.cpp:473-519). */
int dem_set_contacts(dem_ctx *ctx, int maxK, const int *num, const int *list,
                     const double *dispt, const double *nDsp, const double *nDsp_mx,
                     const double *plstRad, const double *plstDfm, const int *stage,
                     const int *lqdBrg, const int *wnum, const int *wlist,
                     const double *wdispt, const double *wnDsp, const double *wnDsp_mx,
                     const double *wplstRad, const double *wplstDfm, const int *wstage,
                     const int *wlqdBrg);
/* Sorted (hash, original index) pairs and CellStart/CellEnd of the last
 * build, in the reference's form (calcHashG + sort_by_key +
 * reorderDataAndFindCellStartG). */
int dem_get_hash(dem_ctx *ctx, int *hash, int *index);
int dem_get_cells(dem_ctx *ctx, int *cellStart, int *cellEnd);

/* ---- the reference's step wrappers ---------------------------------- */
/* PPCntHstryCopy + WallCntHstryCopy (ParticleSystem.cu:336-375) and the
 * list copies of main.cpp:479-485: the current lists and history become
 * "old" (a buffer swap, no copy). */
int dem_cnt_hstry_save(dem_ctx *ctx);
/* BldNeborList (ParticleSystem.cu:141-163): hash, radix sort, cell range,
 * neighbour search.  Also permutes the particle state into cell order and
 * zeroes pMoveforReBld (main.cpp:492). */
int dem_bld_nebor_list(dem_ctx *ctx);
/* BldCirWall (ParticleSystem.cu:193-204) */
int dem_bld_cir_wall(dem_ctx *ctx);
/* HstryCopy + HstryCopyPW (ParticleSystem.cu:171-226): history of every
 * surviving pair moves to its new slot; other slots are zero. */
int dem_hstry_copy(dem_ctx *ctx);
/* VarInit (Allocation.cpp:308-317) */
int dem_var_init(dem_ctx *ctx);
/* PpInterCalc (ParticleSystem.cu:231-250) */
int dem_pp_inter_calc(dem_ctx *ctx);
/* CirPWCalc (ParticleSystem.cu:256-271) */
int dem_cir_pw_calc(dem_ctx *ctx);
/* UpdateSystem (ParticleSystem.cu:279-289); *reNebld receives the count
 * of particles over the rebuild threshold (d_ReNebld). */
int dem_update_system(dem_ctx *ctx, int *reNebld);
/* main.cpp:473-527 as one call (save, build, walls, history remap). */
int dem_rebuild(dem_ctx *ctx);
/* One reference time step through the separate kernels above
 * (main.cpp:470-564). */
int dem_step_unfused(dem_ctx *ctx);

/* ---- the fused path --------------------------------------------------- */
/* nsteps time steps, one fused force+wall+integrate kernel each, the
 * rebuild decided on the device (no per-step host sync; the host checks
 * once per chunk and re-runs skipped steps after a rebuild). */
int dem_step(dem_ctx *ctx, int nsteps);
/* Runs nsteps fused steps with no rebuild check between them and returns
 * the average kernel time [ms] measured with HIP events on the context's
 * stream.  Raises DEM_ERR_STATE if a rebuild would have been needed. */
int dem_time_steps(dem_ctx *ctx, int nsteps, double *ms_per_step, int *rebuild_needed);
int dem_sync(dem_ctx *ctx);
/* DEM_KERNEL_* selection for dem_step.  LOOP: one thread per particle
 * walks its slots.  WAVE: per-wave compaction of the slots in contact (wet
 * packings: in contact, bridged or close enough for a bridge to form), any
 * list length.  COMPACT (block-wide compaction) is refused for wet
 * packings and lists longer than 64 slots.  AUTO takes WAVE.  Every
 * choice gives the same results. */
int dem_set_kernel(dem_ctx *ctx, int variant);
/* Neighbour search form.  TILED stages the positions of each aligned tile
 * of 4 x 4 x 4 cells and its one-cell rim in LDS and scans them from there
 * (one size level in Morton order only); PLAIN has one thread per particle
 * scan its 27 cells in memory; AUTO (default) times both at the context's
 * first rebuilds and keeps the faster, again every 256 rebuilds.  The lists
 * are identical. */
enum { DEM_NSEARCH_AUTO = 0, DEM_NSEARCH_PLAIN = 1, DEM_NSEARCH_TILED = 2 };
int dem_set_neighbor_search(dem_ctx *ctx, int mode);
/* The form the next rebuild uses (AUTO while still timing). */
int dem_neighbor_search_used(dem_ctx *ctx);
/* Storage order of the particles between rebuilds: 1 (default) = Morton
 * order of the cells, 0 = the reference's row-major cell hash.  Lists,
 * forces and the exported hash/cell tables are identical either way. */
int dem_set_order(dem_ctx *ctx, int morton);
/* History layout refresh in dem_step: every `every` steps (default 2048;
 * 0 turns it off) without a rebuild, the contact records are regathered
 * into the compact per-wave region from the contacts of that moment.
 * Records only move, so results are identical with any setting.  No
 * reference counterpart (the reference keeps one slot-major layout). */
int dem_set_relayout(dem_ctx *ctx, int every);
/* Block shape of the wave kernel: every chunk of 64 particles is stepped by
 * `parts` waves (1, 2 or 4) of 64 / parts owners each; 0 (default) is one
 * wave per chunk.  Results are the same for every choice.  No reference
 * counterpart (the reference's launch shape is fixed,
 * ParticleSystem.cu:231-250). */
int dem_set_wave_split(dem_ctx *ctx, int parts);
/* Layout refreshes done so far. */
int dem_n_relayout(dem_ctx *ctx);
/* Shared pair records for the WAVE kernel (on by default): a pair of
 * equal radii keeps one history record instead of the reference's two
 * directed ones (kernel.cu:413-552 evaluates (i, j) and (j, i) with
 * records that stay exact mirrors: equal fields, tangential spring of
 * opposite sign), updated by the lower storage index and read, spring
 * negated, by the other.  Results are identical bit for bit; the history
 * writes of a step halve.  The directed records are restored before
 * anything else reads them (export, rebuild, other kernels).  A packing
 * whose records are found not to be mirrors keeps the directed records,
 * and while rebuilds come less than 64 steps apart the directed records
 * are used (a pack costs about what 30 steps gain).  No reference
 * counterpart. */
int dem_set_pair_records(dem_ctx *ctx, int on);
/* 1 when dem_step would use shared pair records for this context. */
int dem_pair_records_used(dem_ctx *ctx);
int dem_iter(dem_ctx *ctx);
double dem_ctime(dem_ctx *ctx);
int dem_n_rebuild(dem_ctx *ctx);
int dem_contact_stats(dem_ctx *ctx, dem_stats *out);
/* Accumulated event timings of dem_step: fused-kernel time, steps
 * executed, kernels launched, rebuild time, rebuild count. */
int dem_get_timing(dem_ctx *ctx, double *fused_ms, long long *fused_steps,
                   long long *fused_launches, double *rebuild_ms, int *n_rebuild);
/* Measurement aid with no reference counterpart: reps rebuilds of the
 * current state (main.cpp:473-527 with the lists that result unchanged), each
 * stage timed with HIP events on the context's stream; stage_ms[9] receives
 * the mean time of: 0 history save (pair-record write-back), 1 hash,
 * 2 radix sort, 3 permutation into cell order, 4 cell ranges, 5 neighbour
 * search, 6 compact-history layout, 7 wall lists, 8 history remap. */
int dem_profile_rebuild(dem_ctx *ctx, int reps, double *stage_ms);
/* Multiplies every velocity and spin by factor >= 0, keeping positions,
 * lists and contact history: a settling aid (factor 0 quenches the packing
 * to rest) with no reference counterpart. */
int dem_scale_velocities(dem_ctx *ctx, double factor);
/* Bits of the cell key the radix sort orders (8 per pass); a measurement
 * aid with no reference counterpart. */
int dem_key_bits(dem_ctx *ctx);
/* The rebuild the next step would start with, done now (nothing when none
 * is due): lets a caller end a timed run with the rebuild its last step
 * asked for.  No counterpart in the reference, whose loop rebuilds at the
 * top of the next step (main.cpp:470-527). */
int dem_flush_rebuild(dem_ctx *ctx);
/* Bytes of device memory held by the context. */
long long dem_device_bytes(dem_ctx *ctx);
/* Name of the step kernel dem_step launches for this context's packing
 * ("k_step_wave", "k_step_wave_mixed", "k_step_compact" or "k_step_fused");
 * a measurement aid with no reference counterpart. */
const char *dem_step_kernel_name(dem_ctx *ctx);
/* Measurement utility with no reference counterpart: a plain device copy
 * of nbytes, timed over iters launches with HIP events; *gbps = read +
 * write bytes per second / 1e9 (the attainable HBM bandwidth, and the
 * known-size kernel that calibrates rocprofv3's FETCH_SIZE/WRITE_SIZE). */
int dem_stream_copy(long long nbytes, int iters, double *gbps);

/* ---- host interfaces (reference names kept) -------------------------- */
/* heyi.in keyword deck (ReadData.cpp:7-188, main.cpp:65-68); optional
 * extra keys GRAVITY, TIMESTEP, CONTACTANGLE, RUPTUREMODE. */
int dem_read_deck(const char *path, dem_input *in);
/* ParticleFeed (ParticleFeed.cpp:8-72): appends positions for particles
 * numParticle.. into pos[NP][3] (reduced units) with the reference's
 * sampling law, seeded by *seed (updated); returns the new count or -1. */
int dem_particle_feed(const dem_setup *S, const double *rad, double *pos, int numParticle,
                      int NP, unsigned long long *seed);
/* WrtTec (WriteTec.cpp:9-38): same Tecplot text format. */
int dem_write_tec(const char *path, int outs, double time_s, int total, const double *pos,
                  const double *rad, const double *vel, const double *angv,
                  const double *force);
/* LoadPara (DumpLoad.cpp:7-18) and a full restart dump in the same style. */
int dem_load_para(const char *path, int *NP, int *nBatch, int *TotalParticle);
int dem_dump_write(dem_ctx *ctx, const char *path, int nBatch);
/* The same restart dump with the records after the text header written as
 * raw little-endian arrays (only the live list slots): a tenth of the size
 * and written at disk speed, for multi-million-particle restarts. */
int dem_dump_write_binary(dem_ctx *ctx, const char *path, int nBatch);
/* Reads either form (DumpLoad.cpp:7-18 reads the same SYSTEM Info header);
 * the next step rebuilds and remaps the history (RESTART, main.cpp:473). */
int dem_dump_read(dem_ctx *ctx, const char *path, int *nBatch);
int dem_set_lqdcnt(dem_ctx *ctx, const int *lqdCnt);
int dem_set_clock(dem_ctx *ctx, int iter, double ctime, double Tw);
/* The run log <genfile>_Log.dat: its header (main.cpp:93-139; restart 0 =
 * a new run), one record per output (main.cpp:591-618; the particle count
 * is written while total < NP) and the run's total (main.cpp:631-634). */
int dem_log_header(const char *path, const char *genfile, int restart, int NP, int nBatch,
                   int total);
int dem_log_output(const char *path, int outs, double runtime_s, double elapsed_ms, int nrebuild,
                   int total, int NP, int iter);
int dem_log_total(const char *path, double seconds);

#ifdef __cplusplus
}
#endif
#endif
