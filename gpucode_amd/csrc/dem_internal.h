// dem_internal.h -- device-side data layout of the DEM engine (gfx950).
//
// Layout in HBM (cap = particle capacity, K = max neighbour slots), all in
// cell-sorted particle order:
//   P  [cap] double4    (x, y, z, r)                       32 B, two buffers
//   VW [2*cap] double4  (vx, vy, vz, wx), (wy, wz, pmove, -) 64 B, two buffers
//   lqd [cap] int        bridge count, two buffers (lqdCnt / lqdCnt_Old)
//   nbr [K*cap] int      slot-major: nbr[k*cap + i] = sorted index of the
//                        k-th neighbour of i; nnb[cap] counts
//   H  [2*K*cap] HistRec 64-byte contact history, one record per directed
//                        slot: a compact region (the slots in contact when
//                        the layout was made, per particle from actoff[i],
//                        bit k of amask[i] marking them), then a slot-major
//                        region for every other slot (see hist_idx)
//   nmask [cap] u64      slots that can touch before the next rebuild
//   wl [3*cap] int, nwl [cap], HW [3*cap] HistRec  wall slots
// Buffers A (read) and B (written) of P/VW swap every fused step.  One
// record per particle or slot means a gather touches one 32/64-byte sector
// instead of one line per field; slot-major order keeps slot k of 64
// consecutive particles contiguous for the wavefront that walks them.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define DEM_PI 3.1415926  // heydem.h:15
#define DEM_EMPTY (-1)    // 0xffffffff cell sentinel (ParticleSystem.cu:92-93)
#define DEM_MAXL 8

// Per-instance constants (replace the __constant__ symbols of
// kernel.cu:15-19).  Every precomputed product keeps the association order
// of the reference expression it stands for.
struct DevLaw {
    double c43E;       // 4.0/3.0 * emod                       (kernel.cu:123)
    double ydsp_t, yforce_t, ystiff_t;
    double dsmxA;      // 3.0 * sfrc                            (kernel.cu:188)
    double dsmxB;      // 16.0*(1-pois)*emod/(4.0-2.0*pois)     (kernel.cu:188)
    double sfrc;
    double damp;       // -2.0*sqrt(dmpn*dmpn')*emod            (kernel.cu:164, 301)
    int damp_on;       // load stage selects damping            (kernel.cu:162, 299)
    int pad_;
};

struct DevInst {
    DevLaw pp, pw;
    double rolling;            // 0.5*(rfrc0+rfrc0)             (kernel.cu:226)
    double dt, searchGap, rebuildThr;  // rebuildThr = 0.3*SearchGap (kernel.cu:601)
    // liquid (kernel.cu:65-81, 429-550)
    double vol, gapmn, layer2, twoPI, stension;
    double wc[14];             // Willett coefficients at theta
    float wcf[14];             // the same rounded to FP32 (the mixed-precision bridge force)
    double theta;
    int forceCap, ruptureMode;
    // cylinder (heydem.h:39-46)
    double cirX, cirY, R, Bw, topv;
    // grids (one per size level)
    double origin[3];
    double levelSize[DEM_MAXL];
    double cellSize[DEM_MAXL];
    int gx[DEM_MAXL], gy[DEM_MAXL], gz[DEM_MAXL];
    int cellOff[DEM_MAXL];     // hash offset of each level grid
    int levelK[DEM_MAXL];      // nNebors of each level
    int nLevels;
    // storage order of the particles: 0 = row-major cell hash (the
    // reference's sort key), 1 = Morton (Z-order) code of the cell, which
    // keeps 3-D neighbours close in memory, each size level in a range of
    // its own; 2 (several levels) = the levels interleaved in one Morton
    // order: the key of a cell of level L is the Morton code of the
    // level-0 cell holding its centre, then L, so the particles of every
    // level at one place are stored together (a sphere's neighbours of
    // other levels then sit next to it, not in another level's range).
    // Either way the particles of one cell are contiguous and stored by
    // original index.
    int morton;
    unsigned mortonOff[DEM_MAXL];  // key offset of each level's Morton range (morton 1)
    int levelBits;             // bits of the level in the key (morton 2)
    // The CellStart/CellEnd tables are stored in blocks of 4 x 4 x 4 cells:
    // block (x/4, y/4, z/4) of level L in row-major block order from
    // blkOff[L], the 64 cells of a block row-major inside it.  The 27 cells
    // a particle scans then lie in one to eight 256-byte blocks, which the
    // particles stored next to it (Morton order keeps a block's particles
    // together) read as well.  bbx, bby: blocks per row and per layer.
    int bbx[DEM_MAXL], bby[DEM_MAXL], blkOff[DEM_MAXL];
    int ncellTab;              // entries of each table (>= the reference's cell count)
};

// table entry of cell (x, y, z) of level L (see DevInst::blkOff)
__host__ __device__ inline int cell_slot(const DevInst &P, int L, int x, int y, int z) {
    return P.blkOff[L] + (((z >> 2) * P.bby[L] + (y >> 2)) * P.bbx[L] + (x >> 2)) * 64 +
           (((z & 3) << 4) | ((y & 3) << 2) | (x & 3));
}
// first entry of the cells (x, y, z) of one row, x = 0, 4, 8, ...; cell x
// is at row_slot + (x >> 2) * 64 + (x & 3)
__host__ __device__ inline int row_slot(const DevInst &P, int L, int y, int z) {
    return P.blkOff[L] + ((z >> 2) * P.bby[L] + (y >> 2)) * P.bbx[L] * 64 +
           (((z & 3) << 4) | ((y & 3) << 2));
}

// Gap at a rebuild above which a pair cannot touch before the next rebuild:
// until then each centre stays within rebuildThr of where it was, so a gap
// shrinks by at most 2 rebuildThr; the factor 2.2 adds a margin far above
// rounding.
__host__ __device__ inline double near_gap_of(const DevInst &P) { return 2.2 * P.rebuildThr; }

// Seven history fields of one directed contact (heydem.h:153-190) plus the
// flags word: stage in bits 0-7, lqdBrg in bit 8.
struct __attribute__((aligned(64))) HistRec {
    double dx, dy, dz, nd, ndm, prad, pdfm;
    int flags, pad;
};

// Shared pair records.  For a pair of equal radii the reference's two
// directed evaluations, (i, j) and (j, i), compute the same overlap, the
// same load/unload verdict and a tangential spring of opposite sign, bit
// for bit: every operation either repeats on identical operands or acts on
// exactly negated ones (d = (|dx| - r) - r, n and the contact-point
// velocity flip sign, round-to-nearest is symmetric).  Their history
// records therefore stay exact mirrors, and one record per pair serves
// both: the owner (the lower storage index) updates it, the mirror side
// reads it with the spring negated and writes nothing.  Pairs of unequal
// radii ((|dx| - ri) - rj and (|dx| - rj) - ri can round apart) keep one
// record per direction, each side the owner of its own.  The records are
// double-buffered (read hq_in, write hq_out), so the mirror side never sees
// the owner's update of the same step; a record that is not updated in a
// step is identical in both buffers, except one updated in the previous
// step, which the owner copies forward once when the contact ends (bit k of
// the contact mask kept in the spare word of the particle's velocity
// record).
#define PAIR_MIRROR 0x80000000u

// A shared pair record as the wave kernel keeps it: the fields every
// contact step reads and writes (spring, overlap, largest overlap, flags)
// in 48 bytes, and plstRad/plstDfm in a separate 16-byte array that is
// touched only when the flags do not imply them.  Implied: plstDfm = +0 and
// plstRad = the pair's R* (PR_RSTAR) or 0 (neither flag), the values of
// every contact that has not yielded and of every fresh record; otherwise
// PR_COLD is set and the pair (plstRad, plstDfm) is stored.  Lossless.
struct __attribute__((aligned(16))) PairHot {
    double dx, dy, dz, nd, ndm;
    int flags, pad;
};
#define PR_COLD 0x200
#define PR_RSTAR 0x400

typedef double4 dvec4;

struct StateBuf {
    dvec4 *p;    // (x, y, z, r)
    dvec4 *v;    // 2 records per particle
    int *lqd;
};

// Everything one step kernel touches.
struct StepArgs {
    StateBuf in, out;          // read A, write B (in == out for in-place)
    const int *nbr, *nnb;
    HistRec *h;                // history (see hist_idx)
    const unsigned long long *amask;  // per particle: slots whose records are in the compact region
    const int *actoff;         // per particle: start of its compact records
    const unsigned long long *nmask;  // per particle: slots that can touch before the next rebuild
    // shared pair records (the wave kernel's pair mode, see PAIR_MIRROR):
    // the near list of particle i, nl[noff[i] .. noff[i+1]), holds its near
    // slots (the set bits of nmask[i]) in slot order as (partner, pair
    // record index, bit 31 set on the mirror side); records are read from
    // hq_in and written to hq_out
    const int *noff;
    const int2 *nl;
    const PairHot *hq_in;
    PairHot *hq_out;
    const double2 *hc_in;      // plstRad, plstDfm of records with PR_COLD
    double2 *hc_out;
    const int *wl, *nwl;
    HistRec *hw;
    double *Fx, *Fy, *Fz, *Mx, *My, *Mz, *Mr;  // optional force output
    int *stop_at;              // device rebuild request (step index)
    int cap, total, step, store_forces;
    int K;                     // neighbour slots per particle (list stride)
    unsigned xb[9];            // wave kernel: chunks [xb[x], xb[x+1]) go to XCD x
    int wsplit;                // wave kernel: sub-waves per chunk of 64 (1, 2 or 4)
    double Tw;                 // lid height for this step
};

// One instance of a batched launch: its step arguments apart from the
// per-step buffers, both record buffers, the buffer parities at the start
// of the chunk, the number of steps it may take in this chunk and its lid
// height (Tw after step s of the chunk is Tw0 + (s+1) increments of dTw,
// added one at a time as the single-instance driver does).
struct BatchInst {
    StepArgs A;
    double4 *P4[2], *VW[2];
    int *lq[2];
    int cur, Lidx, limit, pad;
    double Tw0, dTw;
};

// Device-driven batched stepping (dem_batch_step with every instance small
// enough for the batched rebuild): the host enqueues steps and rebuild slots
// without reading anything back, and the kernels keep each instance's
// state.  DBatchInst holds both copies of every buffer of an instance;
// DBatchCtl is its state at the start of its current segment (the steps
// since its last rebuild), which the rebuild slot kernels update.
struct DBatchInst {
    double4 *P4[2], *VW[2];
    int *lq[2], *level[2], *id[2];
    int *nbr[2], *nnb[2];
    HistRec *hp[2];
    unsigned long long *amask[2];
    int *actoff[2];
    int *wl[2], *nwl[2];
    HistRec *hw[2];
    unsigned *keys, *vals, *keys2, *vals2;
    int *cstart, *cend;
    unsigned long long *nmask;
    int *overflow;
    int *stop_at;              // the step kernels' rebuild request, local step + 1
    int cap, total, K, ncells, passes, pad;
    double dTw;                // lid increment per step
};

struct DBatchCtl {
    int cur, Lidx, pcur, hcur, hl;  // buffer parities at the segment's start
    int base;                  // global step index at which the segment began
    int left;                  // steps the instance may still take from base
    int armed;                 // rebuilt in the current slot
    int failed;                // > 0: a list overflowed (slots needed); halted
    int nreb;                  // rebuilds in this run
    int executed;              // steps before the current segment, this run
    int seg_steps;             // steps of the segment the last rebuild ended
    double Tw;                 // lid height at the segment's start
};

// Inputs of the packing analysis of one instance (current records).
struct AnaArgs {
    const double4 *p;
    const int *nbr, *nnb, *wl, *nwl;
    int cap, total;
    double Tw;
};

// One instance of a batched rebuild: the buffers before (a) and after (b)
// the permutation into cell order, the sort arrays, the cell tables, and
// the new and old lists and histories.
struct RbInst {
    const double4 *pa, *va;
    const int *lqa, *lvla, *ida;
    double4 *pb, *vb;
    int *lqb, *lvlb, *idb;
    unsigned *keys, *vals, *keys2, *vals2;
    int *cstart, *cend;
    int *nbr, *nnb;
    const int *nbr_old, *nnb_old;
    HistRec *hp;
    const HistRec *hp_old;
    unsigned long long *amask;
    const unsigned long long *amask_old;
    int *actoff;
    const int *actoff_old;
    unsigned long long *nmask;
    int *wl, *nwl;
    const int *wl_old, *nwl_old;
    HistRec *hw;
    const HistRec *hw_old;
    int *overflow;
    int cap, total, old_total, ncells, passes, inst, K, pad;
    double Tw;
};
