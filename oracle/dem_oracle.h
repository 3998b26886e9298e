/*
 * dem_oracle.h -- FP64 CPU restatement of the reference DEM time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This oracle is the checker for the HIP engine in
 * gpucode_amd/.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only to check or to time a CPU baseline.
 * The product never links, loads or calls it.
 *
 * It restates, in plain C99 (+ optional OpenMP over particles), the algorithm
 * of the reference's device and host code:
 *   kernel.cu:65-81      fCap                      -> orc_fcap
 *   kernel.cu:89-228     ContactForceD             -> orc_contact_pp
 *   kernel.cu:236-359    PWcontactD                -> orc_contact_pw
 *   kernel.cu:370-403    MultiPWForceD             -> (inside orc_pw_force)
 *   kernel.cu:413-552    PPForceCalcG              -> orc_pp_force
 *   kernel.cu:563-606    UpdateSystemG             -> orc_update
 *   kernel.cu:617-673    calcGridPos/Hash, calcHashG -> orc_calc_hash
 *   ParticleSystem.cu:61-69  thrust::sort_by_key   -> orc_sort (stable)
 *   kernel.cu:682-724    reorderDataAndFindCellStartG -> orc_cell_range
 *   kernel.cu:729-855    NeighborFindG             -> orc_neighbor_find
 *   kernel.cu:869-911    HstryCopyG                -> orc_hstry_copy
 *   kernel.cu:918-950    HstryCopyPWG              -> orc_hstry_copy_pw
 *   kernel.cu:958-1023   CirPWInterG               -> orc_pw_force
 *   kernel.cu:1034-1073  CirWallListG              -> orc_wall_list
 *   ParticleSystem.cu:336-375 Wall/PPCntHstryCopy  -> orc_hist_save_and_clear
 *   Allocation.cpp:308-317 VarInit                 -> orc_var_init
 *   ReadData.cpp:243-375 derived quantities        -> orc_derive
 *   UnitReduce.cpp:14-104 reduced units            -> orc_derive
 *   main.cpp:328-348     contact constants         -> orc_derive
 *   main.cpp:433-568     one time step             -> orc_step
 *
 * Arithmetic follows double_math.h (dot = x*x'+y*y'+z*z' left to right,
 * length = sqrt(dot), cross as written there) and is compiled with
 * -ffp-contract=off so that no FMA is formed.
 *
 * Parity pins (what each function is checked against) are listed in the
 * header comment of dem_oracle.c.
 */
#ifndef DEM_ORACLE_H
#define DEM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_LEVELS 8
#define ORC_WALLS 3 /* StlMaxWallPerBall, main.cpp:264 */

/* Input deck: the heyi.in keywords (ReadData.cpp:75-188), SI units, plus the
 * overrides the workloads need (fixed dt, gravity, contact angle). */
typedef struct orc_input {
    int NP;                 /* NUMPARTICLE */
    int nLevels;            /* SIZELEVEL */
    int loadStage;          /* LOADMODE (1 = packing) */
    int forceCap;           /* FoceCapillary Yes=1 */
    double tPacking;        /* packing_TIME [s] */
    double realSize;        /* REALSIZE: largest diameter [m] */
    double cutGap;          /* CUTGAPFORNEIGHBORLIST [d_max] */
    double expandSize;      /* EXPANDSIZE [r_max] */
    double cirX, cirY, Bw, R, H; /* CylinderBC [m] */
    double feedLmin, feedLmax;   /* FeedZone [m] */
    double mat[2][7];       /* MATERIAL: density ymod pois sfrc rfrc dmpn yldp */
    double stension, vol, gapmn; /* LIQUID */
    double dtFactor;        /* SIMULATION */
    int nOutPacking;        /* SIMULATION */
    int sizeLevelInput;     /* SIZELEVEL_INPUT Yes=1 */
    double levelDia[ORC_MAX_LEVELS]; /* level diameters [m] when sizeLevelInput */
    double gravity;         /* [m/s^2]; the reference uses 9.8f */
    double dtOverride;      /* [s]; 0 = the reference formula (ReadData.cpp:260-263) */
    double theta;           /* contact angle [rad]; 0 = the reference law */
    int ruptureMode;        /* 0: 0.99 V^0.33 (kernel.cu:522); 1: (1+theta/2) V^(1/3) */
    int gridClampQuirk;     /* 1: clamp cell coords to gridSize (kernel.cu:627-629) */
} orc_input;

/* p-p / p-w contact constants (heydem.h:220-242, main.cpp:328-348). */
typedef struct orc_calpara {
    double emod, pois, sfrc, ystress, ydsp_t, yforce_t, ystiff_t;
} orc_calpara;

/* Everything derived from the deck, in reduced units (UnitReduce.cpp). */
typedef struct orc_setup {
    double rlunit, rmunit, rfunit, rsunit, rtunit, rvunit, reunit;
    double dt, searchGap;
    double origin[3], endPoint[3];
    double cirX, cirY, R, Tw, Bw, topv, btmv;
    double feedLmin, feedLmax, tPacking;
    double density[2], emod[2], ymod[2], pois[2], sfrc[2], rfrc[2], dmpn[2], yldp[2];
    double stension, vol, gapmn, layer, brknmx;
    double theta;
    orc_calpara pp, pw;
    int nLevels, forceCap, loadStage, ruptureMode, gridClampQuirk, nPPCnt;
    double levelSize[ORC_MAX_LEVELS];  /* level max radius, reduced */
    int levelNNebors[ORC_MAX_LEVELS];
    int levelNPars[ORC_MAX_LEVELS];
    int levelNebSId[ORC_MAX_LEVELS];   /* sum_{l<L} nNebors[l]*nPars[l] */
    double cellSize[ORC_MAX_LEVELS];
    int gridSize[ORC_MAX_LEVELS][3];
    int numCells[ORC_MAX_LEVELS];
    int ncellIdx[ORC_MAX_LEVELS];      /* inclusive scan of numCells */
    int sysNCells;
    double radMin, radMax;             /* reduced */
} orc_setup;

/* Seven history fields of one directed contact slot (heydem.h:153-190). */
typedef struct orc_hist1 {
    double dispt[3];
    double nDsp, nDsp_mx, plstRad, plstDfm;
    int stage, lqdBrg;
} orc_hist1;

typedef struct orc_state orc_state;

/* --- derivation ------------------------------------------------------- */
/* dia: NP diameters [m] (the heyi_dia.dat list already tiled over NP; NULL =
 * all equal to realSize).  Fills setup and per-particle rad/level/neborIdx. */
int orc_derive(const orc_input *in, const double *dia, orc_setup *S,
               double *rad, int *pLevel, int *neborIdx);

/* --- single-contact laws (for closed-form tests) ------------------------ */
void orc_contact_pp(const orc_setup *S, orc_hist1 *h, double nrmDsp,
                    const double unitVec[3], double iRad, const double iVel[3],
                    const double iAngv[3], double jRad, const double jVel[3],
                    const double jAngv[3], int stage, double F[3], double M[3],
                    double *Mr);
void orc_contact_pw(const orc_setup *S, orc_hist1 *h, double nrmDsp,
                    const double unitVec[3], double iRad, const double iVel[3],
                    const double iAngv[3], const double jVel[3], int stage,
                    double F[3], double M[3], double *Mr);
double orc_fcap(const orc_setup *S, double Splus, double ijVol, double effectR);
/* Willett's coefficients, [f1 0..3, f2 0..2, f3 0..3, f4 0..2][theta^0..2] */
extern const double ORC_WILLETT[14][3];
double orc_fcap_coef(const double coef[14][3], const orc_setup *S, double Splus,
                     double ijVol, double effectR);
/* rounding-reproducible exp / log / x^y (x > 0) used for the reference's
 * libm calls; see dem_oracle.c */
double orc_exp(double x);
double orc_log(double x);
double orc_pow(double x, double y);
double orc_rupture(const orc_setup *S, double ijVol);

/* --- whole state -------------------------------------------------------- */
orc_state *orc_create(const orc_input *in, const double *dia);
void orc_free(orc_state *st);
const orc_setup *orc_get_setup(const orc_state *st);
int orc_num_threads(int n); /* sets OpenMP threads when >0; returns active count */

/* positions/velocities/spins in reduced units, AoS [total][3] */
void orc_set_particles(orc_state *st, int total, const double *pos,
                       const double *vel, const double *angv);
void orc_feed(orc_state *st, int n, const double *pos);
void orc_get_particles(const orc_state *st, double *pos, double *vel, double *angv);
void orc_get_forces(const orc_state *st, double *F, double *M, double *Mr);
void orc_get_pmove(const orc_state *st, double *pmove);
void orc_get_lqdcnt(const orc_state *st, int *lqdCnt);
int orc_total(const orc_state *st);
int orc_iter(const orc_state *st);
double orc_ctime(const orc_state *st);
double orc_tw(const orc_state *st);
int orc_n_rebuild(const orc_state *st);
int orc_overflow(const orc_state *st);

/* neighbour lists in slot order: num[total], list[total][maxK] (orig ids) */
int orc_get_nebor(const orc_state *st, int *num, int *list, int maxK);
/* histories in slot order: [total][maxK] each field; dispt [total][maxK][3] */
int orc_get_hist_pp(const orc_state *st, int maxK, double *dispt, double *nDsp,
                    double *nDsp_mx, double *plstRad, double *plstDfm,
                    int *stage, int *lqdBrg);
void orc_get_walls(const orc_state *st, int *num, int *list /* [total][3] */);
int orc_set_contacts(orc_state *st, int maxK, const int *num, const int *list,
                     const double *dispt, const double *nDsp, const double *nDsp_mx,
                     const double *plstRad, const double *plstDfm, const int *stage,
                     const int *lqdBrg, const int *wnum, const int *wlist,
                     const double *wdispt, const double *wnDsp, const double *wnDsp_mx,
                     const double *wplstRad, const double *wplstDfm, const int *wstage,
                     const int *wlqdBrg);
void orc_set_lqdcnt(orc_state *st, const int *lqdCnt);
void orc_get_hist_pw(const orc_state *st, double *dispt, double *nDsp,
                     double *nDsp_mx, double *plstRad, double *plstDfm,
                     int *stage, int *lqdBrg); /* each [total][3] */
/* sorted hash / index and cell range after the last build */
void orc_get_hash(const orc_state *st, int *hash, int *index);
void orc_get_cells(const orc_state *st, int *cellStart, int *cellEnd);

/* --- kernels as loops (the step is built from these) -------------------- */
void orc_calc_hash(orc_state *st);
void orc_sort(orc_state *st);
void orc_cell_range(orc_state *st);
void orc_neighbor_find(orc_state *st);
void orc_wall_list(orc_state *st);
void orc_hist_save_and_clear(orc_state *st);
void orc_hstry_copy(orc_state *st);
void orc_hstry_copy_pw(orc_state *st);
void orc_var_init(orc_state *st);
void orc_pp_force(orc_state *st);
void orc_pw_force(orc_state *st);
int orc_update(orc_state *st);

/* main.cpp:473-527 with the flag set; returns 0 or -1 on list overflow */
int orc_rebuild(orc_state *st);
/* main.cpp:439-568 without the feed; returns 0 or -1 on overflow */
int orc_step(orc_state *st);
int orc_run(orc_state *st, int nsteps);
/* forces a rebuild at the next step (FeedFlag/RESTART, main.cpp:473) */
void orc_force_rebuild(orc_state *st);
/* time one force+walls+integrate step (no rebuild) over the current lists */
double orc_time_force_steps(orc_state *st, int nsteps);

#ifdef __cplusplus
}
#endif
#endif
