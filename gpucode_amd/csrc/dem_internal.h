// dem_internal.h -- device-side data layout of the DEM engine (gfx950).
//
// Layout in HBM (cap = particle capacity, K = max neighbour slots):
//   particle state  SoA FP64, in cell-sorted order, two buffers (A read,
//                   B written by the fused step, swapped every step):
//                   x y z | vx vy vz | wx wy wz, plus lqdCnt (int) per buffer
//   rad, pmove      FP64 [cap];  level, inst, id [cap] int
//   neighbour list  int32 nbr[k*cap + i] (sorted index of j), nnb[cap];
//                   a current and an old copy (swapped at each rebuild)
//   p-p history     per slot k*cap+i: dispt x,y,z, nDsp, nDsp_mx, plstRad,
//                   plstDfm (FP64 SoA) + flags (int: stage | lqdBrg << 8);
//                   current and old copies
//   wall list       wl[k*cap + i] (k < 3), nwl[cap]; wall history as p-p
//                   with 3 slots; current and old copies
// Slot-major SoA keeps the accesses of one wavefront to slot k of 64
// consecutive particles contiguous.
#pragma once
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
    int pad2_;
};

struct HistPtr {
    double *dx, *dy, *dz, *nd, *ndm, *prad, *pdfm;
    int *flags;
};

struct StateBuf {
    double *x, *y, *z, *vx, *vy, *vz, *wx, *wy, *wz;
    int *lqd;
};

// Everything one step kernel touches.
struct StepArgs {
    StateBuf in, out;          // read A, write B (in == out for in-place)
    const double *rad;
    double *pmove;
    const int *nbr, *nnb;
    HistPtr h;
    const int *wl, *nwl;
    HistPtr hw;
    double *Fx, *Fy, *Fz, *Mx, *My, *Mz, *Mr;  // optional force output
    int *stop_at;              // device rebuild request (step index)
    int cap, total, step, store_forces;
    double Tw;                 // lid height for this step
};
