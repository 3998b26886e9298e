// dem_kernels.hip -- hand-written gfx950 kernels of the DEM time step.
//
// One thread owns one particle i (cell-sorted order) and gathers every
// directed pair (i, j) for itself, so forces accumulate in registers with
// no atomics.  Each directed pair has its own history slot, as in the
// reference (kernel.cu:413-552), except in the wave kernel's pair mode,
// where a pair of equal radii shares one record (see PAIR_MIRROR in
// dem_internal.h).  Work is HBM-bound integer/FP64 streaming; nothing here
// is a dense contraction, so there is no MFMA.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <type_traits>

#include "dem_internal.h"
#include "dem_launch.h"

#define DEV __device__ __forceinline__

// ---------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------
static inline unsigned blocks_for(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// Blocks are dispatched round-robin over the 8 XCDs; remap so that each
// XCD walks one contiguous range of particles (neighbouring particles share
// that XCD's L2).
DEV unsigned xcd_block(unsigned nb) {
    unsigned b = blockIdx.x;
    unsigned per = nb / 8, rem = nb % 8, x = b % 8, k = b / 8;
    if (nb < 64) return b;
    // XCD x owns [start_x, start_x + per + (x < rem))
    unsigned start = x * per + (x < rem ? x : rem);
    return start + k;
}

// the same map, each XCD's range walked backwards when rev (odd steps of
// the batched wave kernel: a launch starts on the tasks the one before
// finished last, whose data are still in L2)
DEV unsigned xcd_block_dir(unsigned nb, bool rev) {
    const unsigned b = blockIdx.x;
    if (nb < 64) return rev ? nb - 1u - b : b;
    const unsigned per = nb / 8, rem = nb % 8, x = b % 8, k = b / 8;
    const unsigned start = x * per + (x < rem ? x : rem), cnt = per + (x < rem ? 1u : 0u);
    return rev ? start + cnt - 1u - k : start + k;
}

__global__ void k_fill_i32(int *p, int v, long n) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_fill_f64(double *p, double v, long n) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
void launch_fill_i32(int *p, int v, long n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_fill_i32, dim3(blocks_for(n, 256)), dim3(256), 0, s, p, v, n);
}
void launch_fill_f64(double *p, double v, long n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_fill_f64, dim3(blocks_for(n, 256)), dim3(256), 0, s, p, v, n);
}

// ---------------------------------------------------------------------
// exclusive scan of uint32 (radix offsets)
// ---------------------------------------------------------------------
#define SCAN_BS 256
#define SCAN_ITEMS 4
#define SCAN_TILE (SCAN_BS * SCAN_ITEMS)

DEV unsigned wave_incl_scan(unsigned v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        unsigned t = __shfl_up(v, d, 64);
        if (lane >= d) v += t;
    }
    return v;
}

// exclusive scan over a 256-thread block; returns the exclusive prefix and
// the block total
DEV unsigned block_excl_scan(unsigned v, unsigned *total, unsigned *lds /*[4]*/) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    unsigned inc = wave_incl_scan(v);
    if (lane == 63) lds[w] = inc;
    __syncthreads();
    unsigned base = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < SCAN_BS / 64; ++k) {
        unsigned t = lds[k];
        if (k < w) base += t;
        tot += t;
    }
    __syncthreads();
    *total = tot;
    return base + inc - v;
}

__global__ void __launch_bounds__(SCAN_BS) k_scan_local(const unsigned *in, unsigned *out, long n,
                                                        unsigned *bsum) {
    __shared__ unsigned lds[4];
    long base = (long)blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    unsigned v[SCAN_ITEMS], s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        v[k] = (base + k < n) ? in[base + k] : 0u;
        s += v[k];
    }
    unsigned tot;
    unsigned p = block_excl_scan(s, &tot, lds);
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        if (base + k < n) out[base + k] = p;
        p += v[k];
    }
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

// single block: exclusive scan of nb block sums in place
__global__ void __launch_bounds__(SCAN_BS) k_scan_top(unsigned *bsum, long nb) {
    __shared__ unsigned lds[4];
    unsigned carry = 0;
    for (long c = 0; c < nb; c += SCAN_TILE) {
        long base = c + threadIdx.x * SCAN_ITEMS;
        unsigned v[SCAN_ITEMS], s = 0;
#pragma unroll
        for (int k = 0; k < SCAN_ITEMS; ++k) {
            v[k] = (base + k < nb) ? bsum[base + k] : 0u;
            s += v[k];
        }
        unsigned tot;
        unsigned p = block_excl_scan(s, &tot, lds) + carry;
#pragma unroll
        for (int k = 0; k < SCAN_ITEMS; ++k) {
            if (base + k < nb) bsum[base + k] = p;
            p += v[k];
        }
        carry += tot;
    }
}

__global__ void k_scan_add(unsigned *out, long n, const unsigned *bsum) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += bsum[i / SCAN_TILE];
}

long scan_tmp_elems(long n) { return (n + SCAN_TILE - 1) / SCAN_TILE; }

void launch_excl_scan(const unsigned *in, unsigned *out, long n, unsigned *tmp, hipStream_t s) {
    if (n <= 0) return;
    long nb = scan_tmp_elems(n);
    hipLaunchKernelGGL(k_scan_local, dim3((unsigned)nb), dim3(SCAN_BS), 0, s, in, out, n, tmp);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(SCAN_BS), 0, s, tmp, nb);
    hipLaunchKernelGGL(k_scan_add, dim3(blocks_for(n, 256)), dim3(256), 0, s, out, n, tmp);
}

// ---------------------------------------------------------------------
// stable LSD radix sort of (key, value) uint32 pairs, 8-bit digits
// ---------------------------------------------------------------------
#define RS_BS 256
#define RS_ROUNDS 8
#define RS_TILE (RS_BS * RS_ROUNDS)

__global__ void __launch_bounds__(RS_BS) k_rs_hist(const unsigned *keys, long n, int shift,
                                                   unsigned *hist, long nblk) {
    __shared__ unsigned cnt[256];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    long base = (long)blockIdx.x * RS_TILE;
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; ++r) {
        long e = base + (long)r * RS_BS + threadIdx.x;
        if (e < n) atomicAdd(&cnt[(keys[e] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(long)threadIdx.x * nblk + blockIdx.x] = cnt[threadIdx.x];
}

// lanes of this wavefront holding the same 9-bit value d
DEV unsigned long long wave_match9(unsigned d) {
    unsigned long long m = ~0ull;
#pragma unroll
    for (int b = 0; b < 9; ++b) {
        int bit = (d >> b) & 1u;
        unsigned long long bb = __ballot(bit);
        m &= bit ? bb : ~bb;
    }
    return m;
}

__global__ void __launch_bounds__(RS_BS) k_rs_scatter(const unsigned *kin, const unsigned *vin,
                                                      unsigned *kout, unsigned *vout, long n,
                                                      int shift, const unsigned *offs, long nblk) {
    __shared__ unsigned base[256];
    __shared__ unsigned wcnt[RS_BS / 64][256];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    base[tid] = offs[(long)tid * nblk + blockIdx.x];
    const unsigned long long lt = (1ull << lane) - 1ull;
    long tile = (long)blockIdx.x * RS_TILE;
    for (int r = 0; r < RS_ROUNDS; ++r) {
#pragma unroll
        for (int q = 0; q < RS_BS / 64; ++q) wcnt[q][tid] = 0;
        __syncthreads();
        long e = tile + (long)r * RS_BS + tid;
        bool valid = e < n;
        unsigned key = 0, val = 0, d = 256u;
        if (valid) {
            key = kin[e];
            val = vin[e];
            d = (key >> shift) & 255u;
        }
        unsigned long long peers = wave_match9(d);
        unsigned rank = (unsigned)__popcll(peers & lt);
        bool leader = (peers & lt) == 0ull;
        if (valid && leader) wcnt[w][d] = (unsigned)__popcll(peers);
        __syncthreads();
        if (valid) {
            unsigned pos = base[d] + rank;
            for (int q = 0; q < w; ++q) pos += wcnt[q][d];
            kout[pos] = key;
            vout[pos] = val;
        }
        __syncthreads();
        unsigned add = 0;
#pragma unroll
        for (int q = 0; q < RS_BS / 64; ++q) add += wcnt[q][tid];
        base[tid] += add;
        __syncthreads();
    }
}

long radix_tmp_elems(long n) {
    long nblk = (n + RS_TILE - 1) / RS_TILE;
    long h = 256 * nblk;
    return 2 * h + scan_tmp_elems(h) + 16;
}

// Sorts (k, v) of length n by the low `bits` bits of k.  k2/v2 are
// ping-pong buffers; the result ends in (k, v).
void launch_radix_sort(unsigned *k, unsigned *v, unsigned *k2, unsigned *v2, long n, int bits,
                       unsigned *tmp, hipStream_t s) {
    if (n <= 1) return;
    long nblk = (n + RS_TILE - 1) / RS_TILE;
    long h = 256 * nblk;
    unsigned *hist = tmp, *offs = tmp + h, *stmp = tmp + 2 * h;
    int passes = (bits + 7) / 8;
    if (passes < 1) passes = 1;
    unsigned *ka = k, *va = v, *kb = k2, *vb = v2;
    for (int p = 0; p < passes; ++p) {
        int shift = 8 * p;
        hipLaunchKernelGGL(k_rs_hist, dim3((unsigned)nblk), dim3(RS_BS), 0, s, ka, n, shift, hist,
                           nblk);
        launch_excl_scan(hist, offs, h, stmp, s);
        hipLaunchKernelGGL(k_rs_scatter, dim3((unsigned)nblk), dim3(RS_BS), 0, s, ka, va, kb, vb,
                           n, shift, offs, nblk);
        unsigned *t = ka; ka = kb; kb = t;
        t = va; va = vb; vb = t;
    }
    if (ka != k) {
        (void)hipMemcpyAsync(k, ka, n * sizeof(unsigned), hipMemcpyDeviceToDevice, s);
        (void)hipMemcpyAsync(v, va, n * sizeof(unsigned), hipMemcpyDeviceToDevice, s);
    }
}

// ---------------------------------------------------------------------
// cell hash (kernel.cu:617-673)
// ---------------------------------------------------------------------
DEV int cell_coord(double p, double o, double cs, int g) {
    int c = (int)floor((p - o) / cs);
    if (c > g - 1) c = g - 1;
    if (c < 0) c = 0;
    return c;
}

// keys[id] = hash, vals[id] = current index: sorting the id-ordered pairs
// stably by hash yields (hash, original index) order, the order
// thrust::sort_by_key produces from calcHashG's output.
// spreads the low 10 bits of v to every third bit
DEV unsigned spread3(unsigned v) {
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

// row-major cell hash of the reference (kernel.cu:627-640) and, in Morton
// mode, the Z-order key used as the storage order
DEV void cell_of(const DevInst &P, const double4 q, int L, int &hash, unsigned &key) {
    const double cs = P.cellSize[L];
    const int cx = cell_coord(q.x, P.origin[0], cs, P.gx[L]);
    const int cy = cell_coord(q.y, P.origin[1], cs, P.gy[L]);
    const int cz = cell_coord(q.z, P.origin[2], cs, P.gz[L]);
    hash = (cz * P.gy[L] * P.gx[L]) + (cy * P.gx[L]) + cx + P.cellOff[L];
    if (P.morton == 2) {
        // the level-0 cell holding the centre of this cell: every level's
        // cells are at least as large as level 0's, so two cells of one
        // level never share it
        const double c0 = P.cellSize[0];
        const unsigned xc = (unsigned)floor(((double)cx + 0.5) * cs / c0);
        const unsigned yc = (unsigned)floor(((double)cy + 0.5) * cs / c0);
        const unsigned zc = (unsigned)floor(((double)cz + 0.5) * cs / c0);
        key = ((spread3(xc) | (spread3(yc) << 1) | (spread3(zc) << 2)) << P.levelBits) | (unsigned)L;
    } else if (P.morton == 1) {
        key = P.mortonOff[L] + (spread3(cx) | (spread3(cy) << 1) | (spread3(cz) << 2));
    } else {
        key = (unsigned)hash;
    }
}

DEV void calc_hash_one(const DevInst &P, const double4 *p, const int *level, const int *id,
                       unsigned *keys, unsigned *vals, int i) {
    int h;
    unsigned key;
    cell_of(P, p[i], level[i], h, key);
    keys[id[i]] = key;
    vals[id[i]] = (unsigned)i;
}

__global__ void k_calc_hash(const double4 *p, const int *level, const int *id, DevInst P,
                            unsigned *keys, unsigned *vals, int total) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    calc_hash_one(P, p, level, id, keys, vals, i);
}

void launch_calc_hash(const StateBuf &s, const int *level, const int *id, const DevInst &P,
                      unsigned *keys, unsigned *vals, int total, hipStream_t st) {
    if (total <= 0) return;
    hipLaunchKernelGGL(k_calc_hash, dim3(blocks_for(total, 256)), dim3(256), 0, st, s.p, level,
                       id, P, keys, vals, total);
}

// ---------------------------------------------------------------------
// permutation into cell order; zeroes the path length (main.cpp:492)
// ---------------------------------------------------------------------
DEV void permute_one(const unsigned *perm, const StateBuf &a, const StateBuf &b, const int *lvl_a,
                     int *lvl_b, const int *id_a, int *id_b, int i) {
    int o = (int)perm[i];
    b.p[i] = a.p[o];
    b.v[2 * i] = a.v[2 * o];
    double4 v1 = a.v[2 * o + 1];
    v1.z = 0.0;
    b.v[2 * i + 1] = v1;
    b.lqd[i] = a.lqd[o];
    lvl_b[i] = lvl_a[o];
    id_b[i] = id_a[o];
}

__global__ void k_permute(const unsigned *perm, StateBuf a, StateBuf b, const int *lvl_a,
                          int *lvl_b, const int *id_a, int *id_b, int total) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    permute_one(perm, a, b, lvl_a, lvl_b, id_a, id_b, i);
}

void launch_permute(const unsigned *perm, const StateBuf &a, const StateBuf &b, const int *lvl_a,
                    int *lvl_b, const int *id_a, int *id_b, int total, hipStream_t s) {
    if (total <= 0) return;
    hipLaunchKernelGGL(k_permute, dim3(blocks_for(total, 256)), dim3(256), 0, s, perm, a, b, lvl_a,
                       lvl_b, id_a, id_b, total);
}

// ---------------------------------------------------------------------
// cell ranges (kernel.cu:682-724)
// ---------------------------------------------------------------------
// CellStart/CellEnd in the blocked cell layout (cell_slot); a cell's
// particles are contiguous in storage in either order (equal sort keys =
// same cell)
DEV void cell_range_one(const unsigned *keys, const double4 *p, const int *level, const DevInst &P,
                        int *cstart, int *cend, int total, int k) {
    const unsigned key = keys[k];
    const bool first = k == 0 || key != keys[k - 1];
    const bool last = k == total - 1 || key != keys[k + 1];
    if (!first && !last) return;
    const double4 q = p[k];
    const int L = level[k];
    const double cs = P.cellSize[L];
    const int h = cell_slot(P, L, cell_coord(q.x, P.origin[0], cs, P.gx[L]),
                            cell_coord(q.y, P.origin[1], cs, P.gy[L]),
                            cell_coord(q.z, P.origin[2], cs, P.gz[L]));
    if (first) cstart[h] = k;
    if (last) cend[h] = k + 1;
}

__global__ void k_cell_range(const unsigned *keys, const double4 *p, const int *level, DevInst P,
                             int *cstart, int *cend, int total) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= total) return;
    cell_range_one(keys, p, level, P, cstart, cend, total, k);
}

void launch_cell_range(const unsigned *keys, const StateBuf &s, const int *level, const DevInst &P,
                       int *cstart, int *cend, int ncells, int total, hipStream_t st) {
    launch_fill_i32(cstart, DEM_EMPTY, ncells, st);
    launch_fill_i32(cend, DEM_EMPTY, ncells, st);
    if (total <= 0) return;
    hipLaunchKernelGGL(k_cell_range, dim3(blocks_for(total, 256)), dim3(256), 0, st, keys, s.p,
                       level, P, cstart, cend, total);
}

// ---------------------------------------------------------------------
// neighbour search (kernel.cu:729-855), in the reference's scan order:
// level, z row, y row, then the trimmed contiguous sorted range of the row.
// The file is compiled with -ffp-contract=off, so the distance test rounds
// like the host restatement.
// ---------------------------------------------------------------------
#define NFG_KEEP 8

// The part of one scan row (level L, cells (sx..ex, yy, zz)) that can hold
// a neighbour of the sphere at o: a partner of level L has radius at most
// levelSize[L], so it lies within reach = r_i + levelSize[L] + searchGap of
// o.  A row whose yz box is farther than that is skipped (returns false);
// otherwise [sx, ex] is narrowed to the cells whose x interval is within
// reach.  Edge cells hold the particles clamped into them, so their box is
// open on the outer side.  The margin (1e-6 of a cell) is far above the
// rounding of the cell assignment and of the distance test; skipped cells
// hold no hit, and the remaining ones are scanned in the same order, so
// the list is the one of the full scan.  Used on packings of several size
// levels, where a sphere scans up to (2 off + 1)^2 rows of cells smaller
// than its own reach (the three-level 200k bed's search 0.45 -> 0.24 ms);
// on one level nearly every row is within reach and the test costs more
// than it saves (the 2M bed's search 0.44 -> 0.48 ms).  inv = 1 / cellSize[L]
// (the margin covers its rounding).
DEV bool row_reach(const DevInst &P, int L, double inv, const double4 o, double ri, int yy, int zz,
                   int &sx, int &ex) {
    const double cs = P.cellSize[L];
    const double m = 1.0e-6 * cs;
    const double reach = ri + P.levelSize[L] + P.searchGap + m;
    const double ylo = P.origin[1] + yy * cs, zlo = P.origin[2] + zz * cs;
    double gy = 0.0, gz = 0.0;
    if (yy > 0 && o.y < ylo) gy = ylo - o.y;
    else if (yy < P.gy[L] - 1 && o.y > ylo + cs) gy = o.y - (ylo + cs);
    if (zz > 0 && o.z < zlo) gz = zlo - o.z;
    else if (zz < P.gz[L] - 1 && o.z > zlo + cs) gz = o.z - (zlo + cs);
    const double g2 = gy * gy + gz * gz;
    if (!(g2 < reach * reach)) return false;
    const double w = sqrt(reach * reach - g2) + m;
    const double xl = floor((o.x - w - P.origin[0]) * inv), xh = floor((o.x + w - P.origin[0]) * inv);
    if (xl > (double)sx) sx = xl > (double)ex ? ex + 1 : (int)xl;
    if (xh < (double)ex) ex = xh < (double)sx ? sx - 1 : (int)xh;
    return sx <= ex;
}

DEV void nfg_hit(int j, int i, const double4 q, const double4 o, double ri, double gap, int *keep,
                 bool write, int *nbr, int cap, int num0, int K, int &c) {
    if (j == i) return;
    double dx = o.x - q.x, dy2 = o.y - q.y, dz2 = o.z - q.z;
    double len = sqrt(dx * dx + dy2 * dy2 + dz2 * dz2);
    if (len < q.w + ri + gap) {
        if (write) {
            if (num0 + c < K) nbr[(long)(num0 + c) * cap + i] = j;
        } else {
#pragma unroll
            for (int t = 0; t < NFG_KEEP; ++t)
                if (t == c) keep[t] = j;
        }
        ++c;
    }
}

// one row: the cell ranges of four cells at a time are loaded together,
// candidates two at a time (the order of the hits is the scan order)
DEV int nfg_row(const double4 *p, const DevInst &P, const int *cstart, const int *cend, int i,
                const double4 o, double ri, int rowb, int sx, int ex, int *keep, bool write, int *nbr,
                int cap, int num0, int K) {
    int c = 0;
    const double gap = P.searchGap;
    for (int x0 = sx; x0 <= ex; x0 += 4) {
        int s4[4], e4[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const bool in = x0 + u <= ex;
            const int x = x0 + u, h = rowb + ((x >> 2) << 6) + (x & 3);
            s4[u] = in ? cstart[h] : DEM_EMPTY;
            e4[u] = in ? cend[h] : 0;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (s4[u] == DEM_EMPTY) continue;
            const int e0 = e4[u];
            int j = s4[u];
            for (; j + 1 < e0; j += 2) {
                const double4 q0 = p[j], q1 = p[j + 1];
                nfg_hit(j, i, q0, o, ri, gap, keep, write, nbr, cap, num0, K, c);
                nfg_hit(j + 1, i, q1, o, ri, gap, keep, write, nbr, cap, num0, K, c);
            }
            if (j < e0) nfg_hit(j, i, p[j], o, ri, gap, keep, write, nbr, cap, num0, K, c);
        }
    }
    return c;
}

// TRIM: skip the parts of rows out of reach (row_reach; packings of several
// size levels)
template <bool TRIM = false>
DEV void neighbor_find_one(const double4 *p, const int *level, const DevInst &P, const int *cstart,
                           const int *cend, int *nbr, int *nnb, int *overflow, int cap, int i) {
    const double4 o = p[i];
    const double ri = o.w;
    const int iL = level[i];
    const int K = P.levelK[iL];
    int num = 0;
    int keep[NFG_KEEP];  // unused: the rows write their hits straight into the list
    for (int L = 0; L < P.nLevels; ++L) {
        const double cs = P.cellSize[L];
        const int gx = P.gx[L], gy = P.gy[L], gz = P.gz[L];
        const int cx = cell_coord(o.x, P.origin[0], cs, gx);
        const int cy = cell_coord(o.y, P.origin[1], cs, gy);
        const int cz = cell_coord(o.z, P.origin[2], cs, gz);
        const int off = (int)ceil(P.levelSize[iL] / P.levelSize[L]);
        const double inv = TRIM ? 1.0 / cs : 0.0;
        for (int dz = -off; dz <= off; ++dz) {
            const int zz = cz + dz;
            if (zz < 0 || zz > gz - 1) continue;
            for (int dy = -off; dy <= off; ++dy) {
                const int yy = cy + dy;
                if (yy < 0 || yy > gy - 1) continue;
                int sx = cx - off, ex = cx + off;
                if (sx < 0) sx = 0;
                if (ex > gx - 1) ex = gx - 1;
                // the reference scans the sorted range from the first to the
                // last non-empty cell of the row; cell by cell in x order is
                // the same sequence, and works for either storage order.  The
                // part of the row out of reach holds no hit and is skipped.
                // The cell ranges of four cells load together, candidates two
                // at a time, so a thread keeps several loads in flight.
                if (TRIM && !row_reach(P, L, inv, o, ri, yy, zz, sx, ex)) continue;
                num += nfg_row(p, P, cstart, cend, i, o, ri, row_slot(P, L, yy, zz), sx, ex, keep,
                               true, nbr, cap, num, K);
            }
        }
    }
    if (num > K) {
        atomicMax(overflow, num);
        num = K;
    }
    nnb[i] = num;
}

// The same search by a group of 16 lanes per particle, for the batched
// rebuild of small instances, where one thread per particle leaves the chip
// nearly empty and the search time is the latency of one particle's walk
// over its rows.  The rows of the scan (level, z row, y row in the order
// above) are dealt out 16 at a time, one per lane; a lane walks its row's
// cells in x order and candidates in storage order, so the list is the
// concatenation of the rows' hits in row order: a 16-lane prefix sum of the
// hit counts places each lane's hits.  Up to NFG_KEEP hits of a row stay
// in registers from the counting pass; a row with more is walked again.
// g = lane within the particle's group of 16; every lane of the group calls
// this with the same i (valid = false for a group past the end)
DEV void neighbor_find_group16(const double4 *p, const int *level, const DevInst &P,
                               const int *cstart, const int *cend, int *nbr, int *nnb, int *overflow,
                               int cap, int i, int g, bool valid) {
    double4 o = make_double4(0.0, 0.0, 0.0, 0.0);
    int iL = 0;
    if (valid) {
        o = p[i];
        iL = level[i];
    }
    const double ri = o.w;
    const int K = P.levelK[iL];
    // rows per level and their total
    int nrows = 0;
    for (int L = 0; L < P.nLevels; ++L) {
        const int off = (int)ceil(P.levelSize[iL] / P.levelSize[L]);
        nrows += (2 * off + 1) * (2 * off + 1);
    }
    if (!valid) nrows = 0;
    int num = 0;
    const int gbase = (threadIdx.x & 63) & ~15;
    for (int r0 = 0; r0 < nrows; r0 += 16) {
        // this lane's row: level, then (dz, dy)
        int r = r0 + g, L = 0, rowb = 0, sx = 0, ex = -1;
        bool have = r < nrows;
        if (have) {
            int off = 0, w = 0;
            for (L = 0; L < P.nLevels; ++L) {
                off = (int)ceil(P.levelSize[iL] / P.levelSize[L]);
                w = 2 * off + 1;
                if (r < w * w) break;
                r -= w * w;
            }
            const double cs = P.cellSize[L];
            const int gx = P.gx[L], gy = P.gy[L], gz = P.gz[L];
            const int cx = cell_coord(o.x, P.origin[0], cs, gx);
            const int cy = cell_coord(o.y, P.origin[1], cs, gy);
            const int cz = cell_coord(o.z, P.origin[2], cs, gz);
            const int zz = cz + (r / w - off), yy = cy + (r % w - off);
            if (zz < 0 || zz > gz - 1 || yy < 0 || yy > gy - 1) {
                have = false;
            } else {
                sx = cx - off;
                ex = cx + off;
                if (sx < 0) sx = 0;
                if (ex > gx - 1) ex = gx - 1;
                if (P.nLevels > 1) have = row_reach(P, L, 1.0 / cs, o, ri, yy, zz, sx, ex);
                rowb = row_slot(P, L, yy, zz);
            }
        }
        int keep[NFG_KEEP];
        const int c = have ? nfg_row(p, P, cstart, cend, i, o, ri, rowb, sx, ex, keep, false, nbr, cap,
                                     0, K)
                           : 0;
        // inclusive scan of c over the 16 lanes of the group
        int inc = c;
#pragma unroll
        for (int d = 1; d < 16; d <<= 1) {
            const int t = __shfl_up(inc, d, 16);
            if (g >= d) inc += t;
        }
        const int tot = __shfl(inc, gbase + 15, 64);
        const int at = num + inc - c;
        if (c > NFG_KEEP) {
            (void)nfg_row(p, P, cstart, cend, i, o, ri, rowb, sx, ex, keep, true, nbr, cap, at, K);
        } else {
#pragma unroll
            for (int t = 0; t < NFG_KEEP; ++t)
                if (t < c && at + t < K) nbr[(long)(at + t) * cap + i] = keep[t];
        }
        num += tot;
    }
    if (valid && g == 0) {
        if (num > K) {
            atomicMax(overflow, num);
            num = K;
        }
        nnb[i] = num;
    }
}

template <bool TRIM>
__global__ void __launch_bounds__(128) k_neighbor_find(const double4 *p, const int *level,
                                                       DevInst P, const int *cstart,
                                                       const int *cend, int *nbr, int *nnb,
                                                       int *overflow, int cap, int total) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    neighbor_find_one<TRIM>(p, level, P, cstart, cend, nbr, nnb, overflow, cap, i);
}

// ---------------------------------------------------------------------
// The same search with LDS-staged cell tiles, for one size level in Morton
// storage order.  The particles of an aligned tile of 4 x 4 x 4 cells (the
// cells whose Morton codes share all but their low 6 bits) are contiguous in
// storage.  A block takes one tile at a time, copies the positions of the
// 6 x 6 x 6 cells around it (the tile and its one-cell rim, the cells its
// particles scan) into LDS once, and each particle scans its 27 cells from
// LDS in the reference's order: every candidate position is read from HBM
// or L2 once per tile instead of once per particle that scans it.  A tile
// whose rim holds more candidates than LDS takes scans from global memory.
// ---------------------------------------------------------------------
#define NF_BS 256
#define NF_EXT 6                         // tile edge + rim, in cells
#define NF_CELLS (NF_EXT * NF_EXT * NF_EXT)
#define NF_MAXP 768                      // staged candidate positions

// every third bit of v from bit 0, packed (the inverse of spread3)
DEV unsigned compact3(unsigned v) {
    v &= 0x09249249u;
    v = (v | (v >> 2)) & 0x030c30c3u;
    v = (v | (v >> 4)) & 0x0300f00fu;
    v = (v | (v >> 8)) & 0x030000ffu;
    v = (v | (v >> 16)) & 0x000003ffu;
    return v;
}

// Block b takes the tiles that start among storage positions
// [256 b, 256 b + 256), one after another (a tile may run past the range).
__global__ void __launch_bounds__(NF_BS) k_neighbor_tiles(const double4 *p, const int *level,
                                                          DevInst P, const unsigned *keys,
                                                          const int *cstart, const int *cend,
                                                          int *nbr, int *nnb, int *overflow, int cap,
                                                          int total) {
    __shared__ double4 sp[NF_MAXP];
    __shared__ int soff[NF_CELLS + 1];   // LDS offset of each rim cell's candidates, and the total
    __shared__ int sglob[NF_CELLS];      // storage index of each rim cell's first particle
    __shared__ int sstart[NF_BS];        // the tiles starting in this block's range
    __shared__ int swave[NF_BS / 64];    // per-wave totals of a scan
    __shared__ int sval[2];              // tile count; tile end
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int gx = P.gx[0], gy = P.gy[0], gz = P.gz[0];
    const int K = P.levelK[0];
    const double cs = P.cellSize[0];
    {
        const int k = blockIdx.x * NF_BS + tid;
        const bool st = k < total && (k == 0 || (keys[k] >> 6) != (keys[k - 1] >> 6));
        const unsigned long long m = __ballot(st);
        if (lane == 0) swave[wv] = __popcll(m);
        __syncthreads();
        int before = 0, all = 0;
        for (int w = 0; w < NF_BS / 64; ++w) {
            before += w < wv ? swave[w] : 0;
            all += swave[w];
        }
        if (st) sstart[before + __popcll(m & ((1ull << lane) - 1))] = k;
        if (tid == 0) sval[0] = all;
        __syncthreads();
    }
    const int nst = sval[0];
    for (int t = 0; t < nst; ++t) {
        const int k0 = sstart[t];
        const unsigned key0 = keys[k0];
        const unsigned pre = key0 >> 6;
        // the tile's end: the first later position with another prefix
        if (tid == 0) sval[1] = total;
        __syncthreads();
        int k1 = total;
        for (int c0 = k0 + 1; c0 < total; c0 += NF_BS) {
            const int j = c0 + tid;
            if (j < total && (keys[j] >> 6) != pre) atomicMin(&sval[1], j);
            __syncthreads();
            k1 = sval[1];
            __syncthreads();
            if (k1 < total) break;
        }
        // rim cells and their candidate counts (cells outside the grid hold none)
        const unsigned tp = (key0 - P.mortonOff[0]) >> 6;
        const int x0 = 4 * (int)compact3(tp) - 1, y0 = 4 * (int)compact3(tp >> 1) - 1,
                  z0 = 4 * (int)compact3(tp >> 2) - 1;
        int cnt = 0;
        if (tid < NF_CELLS) {
            const int x = x0 + tid % NF_EXT, y = y0 + (tid / NF_EXT) % NF_EXT,
                      z = z0 + tid / (NF_EXT * NF_EXT);
            int g0 = 0;
            if (x >= 0 && x < gx && y >= 0 && y < gy && z >= 0 && z < gz) {
                const int h = cell_slot(P, 0, x, y, z);
                const int s0 = cstart[h], e0 = cend[h];
                if (s0 != DEM_EMPTY) {
                    g0 = s0;
                    cnt = e0 - s0;
                }
            }
            sglob[tid] = g0;
        }
        // exclusive prefix sum over the block: within each wave by shuffles,
        // then the totals of the waves before
        int incl = cnt;
        for (int d = 1; d < 64; d <<= 1) {
            const int u = __shfl_up(incl, d);
            if (lane >= d) incl += u;
        }
        if (lane == 63) swave[wv] = incl;
        __syncthreads();
        int woff = 0;
        for (int w = 0; w < wv; ++w) woff += swave[w];
        if (tid <= NF_CELLS) soff[tid] = woff + incl - cnt;
        __syncthreads();
        const int ncand = soff[NF_CELLS];
        const bool staged = ncand <= NF_MAXP;
        if (staged) {
            // candidate q belongs to the last cell whose offset is <= q
            for (int q = tid; q < ncand; q += NF_BS) {
                int lo = 0, hi = NF_CELLS - 1;
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (soff[mid] <= q) lo = mid;
                    else hi = mid - 1;
                }
                sp[q] = p[sglob[lo] + (q - soff[lo])];
            }
        }
        __syncthreads();
        for (int i = k0 + tid; i < k1; i += NF_BS) {
            if (!staged) {
                neighbor_find_one(p, level, P, cstart, cend, nbr, nnb, overflow, cap, i);
                continue;
            }
            const double4 o = p[i];
            const double ri = o.w;
            const int cx = cell_coord(o.x, P.origin[0], cs, gx);
            const int cy = cell_coord(o.y, P.origin[1], cs, gy);
            const int cz = cell_coord(o.z, P.origin[2], cs, gz);
            int num = 0;
            for (int dz = -1; dz <= 1; ++dz) {
                const int zz = cz + dz;
                if (zz < 0 || zz > gz - 1) continue;
                for (int dy = -1; dy <= 1; ++dy) {
                    const int yy = cy + dy;
                    if (yy < 0 || yy > gy - 1) continue;
                    int sx = cx - 1, ex = cx + 1;
                    if (sx < 0) sx = 0;
                    if (ex > gx - 1) ex = gx - 1;
                    const int rowc = (zz - z0) * NF_EXT * NF_EXT + (yy - y0) * NF_EXT - x0;
                    for (int x = sx; x <= ex; ++x) {
                        const int c = rowc + x;
                        const int q0 = soff[c], q1 = soff[c + 1];
                        const int g0 = sglob[c] - q0;
                        for (int q = q0; q < q1; ++q) {
                            const int j = g0 + q;
                            if (j == i) continue;
                            const double4 b = sp[q];
                            double dx = o.x - b.x, dy2 = o.y - b.y, dz2 = o.z - b.z;
                            double len = sqrt(dx * dx + dy2 * dy2 + dz2 * dz2);
                            if (len < b.w + ri + P.searchGap) {
                                if (num < K) nbr[(long)num * cap + i] = j;
                                ++num;
                            }
                        }
                    }
                }
            }
            if (num > K) {
                atomicMax(overflow, num);
                num = K;
            }
            nnb[i] = num;
        }
        __syncthreads();  // LDS is reused by the next tile
    }
}

bool neighbor_tiles_apply(const DevInst &P) {
    return P.morton == 1 && P.nLevels == 1 && (P.mortonOff[0] & 63u) == 0 && P.gx[0] <= 1024 &&
           P.gy[0] <= 1024 && P.gz[0] <= 1024;
}

// tiled: the LDS-staged form (only where neighbor_tiles_apply holds)
void launch_neighbor_find(const StateBuf &s, const int *level, const DevInst &P, const unsigned *keys,
                          const int *cstart, const int *cend, int *nbr, int *nnb, int *overflow,
                          bool tiled, int cap, int total, hipStream_t st) {
    if (total <= 0) return;
    if (tiled && neighbor_tiles_apply(P))
        hipLaunchKernelGGL(k_neighbor_tiles, dim3(blocks_for(total, NF_BS)), dim3(NF_BS), 0, st, s.p,
                           level, P, keys, cstart, cend, nbr, nnb, overflow, cap, total);
    else if (P.nLevels > 1)
        hipLaunchKernelGGL(k_neighbor_find<true>, dim3(blocks_for(total, 128)), dim3(128), 0, st, s.p,
                           level, P, cstart, cend, nbr, nnb, overflow, cap, total);
    else
        hipLaunchKernelGGL(k_neighbor_find<false>, dim3(blocks_for(total, 128)), dim3(128), 0, st,
                           s.p, level, P, cstart, cend, nbr, nnb, overflow, cap, total);
}

// ---------------------------------------------------------------------
// cylinder wall candidates (kernel.cu:1034-1073)
// ---------------------------------------------------------------------
DEV void wall_list_one(const double4 *p, const DevInst &P, double Tw, int *wl, int *nwl, int cap,
                       int i) {
    const double4 q = p[i];
    const double r = q.w;
    int n = 0;
    if (q.z >= (Tw - (2.0 * r + P.searchGap))) wl[(long)(n++) * cap + i] = 0;
    if (q.z <= ((2.0 * r + P.searchGap) + P.Bw)) wl[(long)(n++) * cap + i] = 1;
    double ax = P.cirX - q.x, ay = P.cirY - q.y;
    double dc = sqrt(ax * ax + ay * ay);
    if (dc >= P.R - (2.0 * r + P.searchGap)) wl[(long)(n++) * cap + i] = 2;
    nwl[i] = n;
}

__global__ void k_wall_list(const double4 *p, DevInst P, double Tw, int *wl, int *nwl, int cap,
                            int total) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    wall_list_one(p, P, Tw, wl, nwl, cap, i);
}

void launch_wall_list(const StateBuf &s, const DevInst &P, double Tw, int *wl, int *nwl, int cap,
                      int total, hipStream_t st) {
    if (total <= 0) return;
    hipLaunchKernelGGL(k_wall_list, dim3(blocks_for(total, 256)), dim3(256), 0, st, s.p, P, Tw, wl,
                       nwl, cap, total);
}

// ---------------------------------------------------------------------
// history remap across a rebuild (kernel.cu:869-950 with
// ParticleSystem.cu:336-375): each new slot takes the history of the same
// partner in the old list, or zero.  Partners are compared by original id
// because the particle order changes at every rebuild.
// ---------------------------------------------------------------------
DEV HistRec hist_zero() {
    HistRec r;
    r.dx = 0.0; r.dy = 0.0; r.dz = 0.0; r.nd = 0.0; r.ndm = 0.0; r.prad = 0.0; r.pdfm = 0.0;
    r.flags = 0;
    r.pad = 0;
    return r;
}

// History layout.  The history array holds two regions of K*cap records:
// a compact region, then a slot-major region.  At a rebuild the slots of
// particle i that are in or near contact get their records in the compact region,
// in slot order, from actoff[i] on (a prefix sum over the particles in
// storage order), so that the contact records of a wave's particles are
// contiguous; bit k of amask[i] marks those slots.  Every other slot k
// keeps its record at K*cap + k*cap + i.  Where a record lives changes
// nothing else: each directed slot has exactly one record, as in the
// reference, and contacts are still summed in slot order.
DEV long hist_idx(unsigned long long am, int off, int k, int i, long cap, long Kcap) {
    if (k < 64 && ((am >> k) & 1ull)) return (long)off + __popcll(am & ((1ull << k) - 1ull));
    return Kcap + (long)k * cap + i;
}

// slots of particle i (the first 64) in contact now: amask[i] and their
// count.  Only the layout depends on this.  Contacts that form later sit in
// the slot-major region until the next rebuild or layout refresh; keeping
// near-contacts in the compact region as well leaves holes between the
// records that are read, and costs more than it saves.
//
// At a rebuild (nmask given) also the slots that can touch before the next
// one: a rebuild comes as soon as a particle's path length passes
// rebuildThr, so until then no centre moves farther than that from where
// it was at the rebuild, and a pair's gap shrinks by at most 2 rebuildThr.
// Slots with a larger gap now (near_gap_of, with a 10 % margin) cannot
// touch, and the wave step kernel does not test them.  Bits cover the
// first 64 slots; the kernels that use nmask need K <= 64.
DEV void active_slots_one(const double4 *p, const int *nbr, const int *nnb, unsigned long long *amask,
                          unsigned *cnt, unsigned long long *nmask, double near_gap, int cap,
                          int i) {
    const int n = nnb[i];
    const double4 pi = p[i];
    unsigned long long m = 0ull, nm = 0ull;
    for (int a = 0; a < n && a < 64; ++a) {
        const double4 q = p[nbr[(long)a * cap + i]];
        const double dx = pi.x - q.x, dy = pi.y - q.y, dz = pi.z - q.z;
        const double rs = pi.w + q.w;
        const double gap = sqrt(dx * dx + dy * dy + dz * dz) - rs;
        if (gap < 0.0) m |= 1ull << a;
        if (gap < near_gap) nm |= 1ull << a;
    }
    amask[i] = m;
    cnt[i] = (unsigned)__popcll(m);
    if (nmask) nmask[i] = nm;
}

__global__ void k_active_slots(const double4 *p, const int *nbr, const int *nnb,
                               unsigned long long *amask, unsigned *cnt,
                               unsigned long long *nmask, double near_gap, int cap, int total) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < total) active_slots_one(p, nbr, nnb, amask, cnt, nmask, near_gap, cap, i);
}

// amask and actoff of the lists (cnt is scratch of total entries), and at a
// rebuild nmask (null at a layout refresh)
void launch_active_slots(const double4 *p, const int *nbr, const int *nnb, unsigned long long *amask,
                         int *actoff, unsigned long long *nmask, double near_gap, unsigned *cnt,
                         unsigned *scan_tmp, int cap, int total, hipStream_t s) {
    if (total <= 0) return;
    hipLaunchKernelGGL(k_active_slots, dim3(blocks_for(total, 256)), dim3(256), 0, s, p, nbr, nnb, amask,
                       cnt, nmask, near_gap, cap, total);
    launch_excl_scan(cnt, (unsigned *)actoff, total, scan_tmp, s);
}

// layout refresh: the same lists, records moved from one layout (am_old,
// off_old in h_old) to another (am_new, off_new in h); every listed slot's
// record is copied unchanged
__global__ void k_hist_relayout(const int *nnb, const unsigned long long *am_old,
                                const int *off_old, const unsigned long long *am_new,
                                const int *off_new, const HistRec *h_old, HistRec *h, int cap,
                                long Kcap, int total) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const unsigned long long amo = am_old[i], amn = am_new[i];
    const int ofo = off_old[i], ofn = off_new[i];
    const int n = nnb[i];
    for (int k = 0; k < n; ++k)
        h[hist_idx(amn, ofn, k, i, cap, Kcap)] = h_old[hist_idx(amo, ofo, k, i, cap, Kcap)];
}

void launch_hist_relayout(const int *nnb, const unsigned long long *am_old, const int *off_old,
                          const unsigned long long *am_new, const int *off_new,
                          const HistRec *h_old, HistRec *h, int cap, long Kcap, int total,
                          hipStream_t s) {
    if (total <= 0) return;
    hipLaunchKernelGGL(k_hist_relayout, dim3(blocks_for(total, 256)), dim3(256), 0, s, nnb, am_old,
                       off_old, am_new, off_new, h_old, h, cap, Kcap, total);
}

DEV void hist_remap_one(const unsigned *perm, const int *id_new, const int *id_old, const int *nbr,
                        const int *nnb, const unsigned long long *am_new, const int *off_new,
                        const int *nbr_old, const int *nnb_old, const unsigned long long *am_old,
                        const int *off_old, HistRec *h, const HistRec *h_old, int cap, long Kcap,
                        int old_total, int i) {
    int p = (int)perm[i];
    const bool had = p < old_total;
    int nold = had ? nnb_old[p] : 0;
    const unsigned long long amo = had ? am_old[p] : 0ull;
    const int ofo = had ? off_old[p] : 0;
    const unsigned long long amn = am_new[i];
    const int ofn = off_new[i];
    int n = nnb[i];
    // the old partners' ids, gathered once (a short old list stays in
    // registers); the last old slot with the partner's id wins, as in
    // kernel.cu:895-905 (partners are unique, so there is one)
    constexpr int kReg = 32;
    int oid[kReg];
#pragma unroll
    for (int b = 0; b < kReg; ++b) oid[b] = b < nold ? id_old[nbr_old[(long)b * cap + p]] : -1;
    for (int a = 0; a < n; ++a) {
        long sa = (long)a * cap + i;
        int jid = id_new[nbr[sa]];
        int hit = -1;
#pragma unroll
        for (int b = 0; b < kReg; ++b)
            if (oid[b] == jid) hit = b;
        for (int b = kReg; b < nold; ++b)
            if (id_old[nbr_old[(long)b * cap + p]] == jid) hit = b;
        h[hist_idx(amn, ofn, a, i, cap, Kcap)] =
            hit >= 0 ? h_old[hist_idx(amo, ofo, hit, p, cap, Kcap)] : hist_zero();
    }
}

__global__ void k_hist_remap(const unsigned *perm, const int *id_new, const int *id_old,
                             const int *nbr, const int *nnb, const unsigned long long *am_new,
                             const int *off_new, const int *nbr_old, const int *nnb_old,
                             const unsigned long long *am_old, const int *off_old, HistRec *h,
                             const HistRec *h_old, int cap, long Kcap, int total, int old_total) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    hist_remap_one(perm, id_new, id_old, nbr, nnb, am_new, off_new, nbr_old, nnb_old, am_old, off_old, h,
                   h_old, cap, Kcap, old_total, i);
}

// the directed slot-major view (record k*cap+i = slot k of particle i) of
// the history, and back
__global__ void k_hist_to_slots(const HistRec *h, const int *nnb, const unsigned long long *amask,
                                const int *actoff, HistRec *out, int cap, long Kcap, int total) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const unsigned long long am = amask[i];
    const int off = actoff[i];
    for (int a = 0; a < nnb[i]; ++a) out[(long)a * cap + i] = h[hist_idx(am, off, a, i, cap, Kcap)];
}

__global__ void k_hist_from_slots(HistRec *h, const int *nnb, const unsigned long long *amask,
                                  const int *actoff, const HistRec *in, int cap, long Kcap, int total) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const unsigned long long am = amask[i];
    const int off = actoff[i];
    for (int a = 0; a < nnb[i]; ++a) h[hist_idx(am, off, a, i, cap, Kcap)] = in[(long)a * cap + i];
}

void launch_hist_slots(HistRec *h, const int *nnb, const unsigned long long *amask, const int *actoff,
                       HistRec *slots, int cap, long Kcap, int total, bool to_slots, hipStream_t s) {
    if (total <= 0) return;
    if (to_slots)
        hipLaunchKernelGGL(k_hist_to_slots, dim3(blocks_for(total, 256)), dim3(256), 0, s, h, nnb, amask,
                           actoff, slots, cap, Kcap, total);
    else
        hipLaunchKernelGGL(k_hist_from_slots, dim3(blocks_for(total, 256)), dim3(256), 0, s, h, nnb,
                           amask, actoff, slots, cap, Kcap, total);
}

DEV void hist_remap_pw_one(const unsigned *perm, const int *wl, const int *nwl, const int *wl_old,
                           const int *nwl_old, HistRec *h, const HistRec *h_old, int cap,
                           int old_total, int i) {
    int p = (int)perm[i];
    int nold = p < old_total ? nwl_old[p] : 0;
    for (int a = 0; a < nwl[i]; ++a) {
        long sa = (long)a * cap + i;
        int w = wl[sa], hit = -1;
        for (int b = 0; b < nold; ++b)
            if (wl_old[(long)b * cap + p] == w) hit = b;
        h[sa] = hit >= 0 ? h_old[(long)hit * cap + p] : hist_zero();
    }
}

__global__ void k_hist_remap_pw(const unsigned *perm, const int *wl, const int *nwl,
                                const int *wl_old, const int *nwl_old, HistRec *h,
                                const HistRec *h_old, int cap, int total, int old_total) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    hist_remap_pw_one(perm, wl, nwl, wl_old, nwl_old, h, h_old, cap, old_total, i);
}

void launch_hist_remap(const unsigned *perm, const int *id_new, const int *id_old, const int *nbr,
                       const int *nnb, const unsigned long long *am_new, const int *off_new,
                       const int *nbr_old, const int *nnb_old, const unsigned long long *am_old,
                       const int *off_old, HistRec *h, const HistRec *h_old, const int *wl,
                       const int *nwl, const int *wl_old, const int *nwl_old, HistRec *hw,
                       const HistRec *hw_old, int cap, long Kcap, int total, int old_total,
                       hipStream_t s) {
    if (total <= 0) return;
    hipLaunchKernelGGL(k_hist_remap, dim3(blocks_for(total, 128)), dim3(128), 0, s, perm, id_new,
                       id_old, nbr, nnb, am_new, off_new, nbr_old, nnb_old, am_old, off_old, h, h_old,
                       cap, Kcap, total, old_total);
    hipLaunchKernelGGL(k_hist_remap_pw, dim3(blocks_for(total, 256)), dim3(256), 0, s, perm, wl,
                       nwl, wl_old, nwl_old, hw, hw_old, cap, total, old_total);
}

// ---------------------------------------------------------------------
// Shared pair records (see PAIR_MIRROR in dem_internal.h): made from the
// directed records of the near slots (nmask), and written back to them,
// with the lists unchanged.  A slot owns its record unless its partner has
// the same radius and a lower storage index; owned records of pairs in
// contact now come first, in (particle, slot) order, then the other owned
// records, so that the records a step reads are contiguous per wave.
// ---------------------------------------------------------------------
DEV unsigned long long near_slots(const unsigned long long *nmask, const int *nnb, int i) {
    const int n = nnb[i];
    unsigned long long nm = nmask[i];
    if (n < 64) nm &= (1ull << n) - 1ull;
    return nm;
}

DEV bool owns_record(int i, int j, double ri, double rj) { return j > i || rj != ri; }

DEV double pair_rstar(double ri, double rj) { return ri * rj / (ri + rj); }  // as pair_compute

// the full record from its stored form (cold: its plstRad, plstDfm, read
// only when PR_COLD is set; R: the pair's R*)
DEV HistRec pair_expand(const PairHot &p, double2 cold, double R) {
    HistRec h;
    h.dx = p.dx; h.dy = p.dy; h.dz = p.dz; h.nd = p.nd; h.ndm = p.ndm;
    if (p.flags & PR_COLD) {
        h.prad = cold.x;
        h.pdfm = cold.y;
    } else {
        h.prad = (p.flags & PR_RSTAR) ? R : 0.0;
        h.pdfm = 0.0;
    }
    h.flags = p.flags & ~(PR_COLD | PR_RSTAR);
    h.pad = 0;
    return h;
}

// the stored form of h; need_cold: cold must be written as well
DEV PairHot pair_compress(const HistRec &h, double R, double2 &cold, bool &need_cold) {
    PairHot p;
    p.dx = h.dx; p.dy = h.dy; p.dz = h.dz; p.nd = h.nd; p.ndm = h.ndm;
    int f = h.flags & ~(PR_COLD | PR_RSTAR);
    const long long bp = __double_as_longlong(h.prad), bd = __double_as_longlong(h.pdfm);
    need_cold = false;
    if (bd == 0 && bp == __double_as_longlong(R)) {
        f |= PR_RSTAR;
    } else if (!(bd == 0 && bp == 0)) {
        f |= PR_COLD;
        cold = make_double2(h.prad, h.pdfm);
        need_cold = true;
    }
    p.flags = f;
    p.pad = 0;
    return p;
}

// per particle: owned near slots in contact now (c[0]), other owned near
// slots (c[1]), near slots (c[2]), near slots in contact now (c[3]); c[q]
// is c + q*cap
__global__ void k_pair_count(const double4 *p, const int *nbr, const int *nnb,
                             const unsigned long long *nmask, unsigned *c, int cap, int total) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const double4 pi = p[i];
    unsigned long long nm = near_slots(nmask, nnb, i);
    unsigned a = 0, b = 0, t = 0;
    const unsigned n = (unsigned)__popcll(nm);
    while (nm) {
        const int k = __ffsll((long long)nm) - 1;
        nm &= nm - 1ull;
        const int j = nbr[(long)k * cap + i];
        const double4 q = p[j];
        const double dx = pi.x - q.x, dy = pi.y - q.y, dz = pi.z - q.z;
        const bool touch = sqrt(dx * dx + dy * dy + dz * dz) - pi.w - q.w < 0.0;
        t += touch ? 1u : 0u;
        if (!owns_record(i, j, pi.w, q.w)) continue;
        if (touch) a++;
        else b++;
    }
    c[i] = a;
    c[(long)cap + i] = b;
    c[2L * cap + i] = n;
    c[3L * cap + i] = t;
}

// work of each chunk of 64 particles for the wave kernel: 72 + its contacts
// (a wave's time, fitted on the 2M bed, is about 6.8 us + 0.094 us per
// contact entry)
__global__ void k_chunk_work(const unsigned *contacts, unsigned *work, int total) {
    const int ch = blockIdx.x * blockDim.x + threadIdx.x;
    const int lo = ch * 64;
    if (lo >= total) return;
    const int hi = lo + 64 < total ? lo + 64 : total;
    unsigned w = 72;
    for (int i = lo; i < hi; ++i) w += contacts[i];
    work[ch] = w;
}

void launch_chunk_work(const unsigned *contacts, unsigned *work, int total, hipStream_t s) {
    if (total <= 0) return;
    hipLaunchKernelGGL(k_chunk_work, dim3(blocks_for(blocks_for(total, 64), 256)), dim3(256), 0, s,
                       contacts, work, total);
}

// owned slots: assign records, store the directed records in both buffers,
// and write the near list (partner, record) of particle i in slot order;
// mirror slots get record -1 until k_pair_fill_mirror.  Clears the contact
// mask in the spare word of the velocity records the next step reads.
__global__ void k_pair_fill_owner(const double4 *p, double4 *vw, const int *nbr, const int *nnb,
                                  const unsigned long long *nmask, const unsigned *o, unsigned base2,
                                  const HistRec *hd, const unsigned long long *amask,
                                  const int *actoff, int2 *nl, PairHot *q0, PairHot *q1,
                                  double2 *c0, double2 *c1, int cap, long Kcap, int total) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    vw[2L * i + 1].w = 0.0;
    const double4 pi = p[i];
    unsigned long long nm = near_slots(nmask, nnb, i);
    const unsigned long long am = amask[i];
    const int ao = actoff[i];
    unsigned r1 = o[i], r2 = base2 + o[(long)cap + i], slot = o[2L * cap + i];
    while (nm) {
        const int k = __ffsll((long long)nm) - 1;
        nm &= nm - 1ull;
        const int j = nbr[(long)k * cap + i];
        const double4 q = p[j];
        int rec = -1;
        if (owns_record(i, j, pi.w, q.w)) {
            const double dx = pi.x - q.x, dy = pi.y - q.y, dz = pi.z - q.z;
            const double gap = sqrt(dx * dx + dy * dy + dz * dz) - pi.w - q.w;
            rec = (int)(gap < 0.0 ? r1++ : r2++);
            double2 cold;
            bool need = false;
            const PairHot h = pair_compress(hd[hist_idx(am, ao, k, i, cap, Kcap)],
                                            pair_rstar(pi.w, q.w), cold, need);
            q0[rec] = h;
            q1[rec] = h;
            if (need) {
                c0[rec] = cold;
                c1[rec] = cold;
            }
        }
        nl[slot++] = make_int2(j, rec);
    }
}

// mirror slots: the record of the partner's entry for i in its near list.
// Sets *bad when the partner's near list has no entry for i or when the
// directed record of this slot is not the exact mirror of the partner's
// (then the pair records cannot stand for the directed ones).
__global__ void k_pair_fill_mirror(const double4 *p, const int *nnb,
                                   const unsigned long long *nmask, const unsigned *o,
                                   const HistRec *hd, const unsigned long long *amask,
                                   const int *actoff, int2 *nl, const PairHot *q0,
                                   const double2 *c0, int *bad, int cap, long Kcap, int total) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const double4 pi = p[i];
    unsigned long long nm = near_slots(nmask, nnb, i);
    const unsigned long long am = amask[i];
    const int ao = actoff[i];
    const unsigned *noff = o + 2L * cap;
    unsigned slot = noff[i];
    while (nm) {
        const int k = __ffsll((long long)nm) - 1;
        nm &= nm - 1ull;
        const unsigned me = slot++;
        const int j = nl[me].x;
        const double rj = p[j].w;
        if (owns_record(i, j, pi.w, rj)) continue;
        int rec = -1;
        for (unsigned b = noff[j], e = noff[j + 1]; b < e && rec < 0; ++b) {
            const int2 v = nl[b];
            if (v.x == i) rec = v.y >= 0 ? v.y : -2;
        }
        if (rec < 0) {
            *bad = 1;
            nl[me].y = 0;
            continue;
        }
        const PairHot qo = q0[rec];
        const double2 cold = (qo.flags & PR_COLD) ? c0[rec] : make_double2(0.0, 0.0);
        const HistRec ho = pair_expand(qo, cold, pair_rstar(pi.w, rj));
        const HistRec hm = hd[hist_idx(am, ao, k, i, cap, Kcap)];
        if (!(hm.dx == -ho.dx && hm.dy == -ho.dy && hm.dz == -ho.dz && hm.nd == ho.nd &&
              hm.ndm == ho.ndm && hm.prad == ho.prad && hm.pdfm == ho.pdfm && hm.flags == ho.flags))
            *bad = 1;
        nl[me].y = (int)((unsigned)rec | PAIR_MIRROR);
    }
}

// the pair records (buffers q, c: the ones the next step reads) back into
// the directed records of the near slots
__global__ void k_pair_unpack(const double4 *p, const int *nnb, const unsigned long long *nmask,
                              const unsigned *noff, const int2 *nl, const PairHot *q,
                              const double2 *c, HistRec *hd, const unsigned long long *amask,
                              const int *actoff, int cap, long Kcap, int total) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const double ri = p[i].w;
    unsigned long long nm = near_slots(nmask, nnb, i);
    const unsigned long long am = amask[i];
    const int ao = actoff[i];
    unsigned slot = noff[i];
    while (nm) {
        const int k = __ffsll((long long)nm) - 1;
        nm &= nm - 1ull;
        const int2 v = nl[slot++];
        const int rec = v.y & 0x7fffffff;
        const PairHot ph = q[rec];
        const double2 cold = (ph.flags & PR_COLD) ? c[rec] : make_double2(0.0, 0.0);
        HistRec h = pair_expand(ph, cold, pair_rstar(ri, p[v.x].w));
        if (v.y < 0) {
            h.dx = -h.dx; h.dy = -h.dy; h.dz = -h.dz;
        }
        hd[hist_idx(am, ao, k, i, cap, Kcap)] = h;
    }
}

void launch_pair_count(const double4 *p, const int *nbr, const int *nnb,
                       const unsigned long long *nmask, unsigned *c, int cap, int total,
                       hipStream_t s) {
    if (total <= 0) return;
    hipLaunchKernelGGL(k_pair_count, dim3(blocks_for(total, 256)), dim3(256), 0, s, p, nbr, nnb, nmask,
                       c, cap, total);
}

void launch_pair_fill(const double4 *p, double4 *vw, const int *nbr, const int *nnb,
                      const unsigned long long *nmask, const unsigned *o, unsigned base2,
                      const HistRec *hd, const unsigned long long *amask, const int *actoff,
                      int2 *nl, PairHot *q0, PairHot *q1, double2 *c0, double2 *c1, int *bad,
                      int cap, long Kcap, int total, hipStream_t s) {
    if (total <= 0) return;
    hipLaunchKernelGGL(k_pair_fill_owner, dim3(blocks_for(total, 256)), dim3(256), 0, s, p, vw, nbr,
                       nnb, nmask, o, base2, hd, amask, actoff, nl, q0, q1, c0, c1, cap, Kcap,
                       total);
    hipLaunchKernelGGL(k_pair_fill_mirror, dim3(blocks_for(total, 256)), dim3(256), 0, s, p, nnb,
                       nmask, o, hd, amask, actoff, nl, q0, c0, bad, cap, Kcap, total);
}

void launch_pair_unpack(const double4 *p, const int *nnb, const unsigned long long *nmask,
                        const unsigned *noff, const int2 *nl, const PairHot *q, const double2 *c,
                        HistRec *hd, const unsigned long long *amask, const int *actoff, int cap,
                        long Kcap, int total, hipStream_t s) {
    if (total <= 0) return;
    hipLaunchKernelGGL(k_pair_unpack, dim3(blocks_for(total, 256)), dim3(256), 0, s, p, nnb, nmask,
                       noff, nl, q, c, hd, amask, actoff, cap, Kcap, total);
}

// ---------------------------------------------------------------------
// contact laws, FP64 (kernel.cu:89-359), same evaluation order as the
// reference expressions (no FP contraction, see above).
// ---------------------------------------------------------------------

struct V3 { double x, y, z; };
DEV V3 v3(double x, double y, double z) { V3 r = {x, y, z}; return r; }
DEV V3 add(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
DEV V3 sub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
DEV V3 mulv(V3 a, double s) { return v3(a.x * s, a.y * s, a.z * s); }
DEV V3 smulv(double s, V3 a) { return v3(s * a.x, s * a.y, s * a.z); }
DEV V3 divv(V3 a, double s) { return v3(a.x / s, a.y / s, a.z / s); }
DEV double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
DEV double len(V3 a) { return sqrt(dot(a, a)); }
DEV V3 cross(V3 a, V3 b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
DEV double pow15(double x) { return x * sqrt(x); }

// exp, log and x^y for the FP64 path, built from correctly rounded
// +, -, *, /, sqrt and exact scaling only, so the host restatement rounds
// identically (the contact law is discontinuous at the stick/slip and
// yield boundaries, where a last-bit difference between two libm's would
// change a branch).
//   exp: x = k ln2 + r with k = rint(x / ln2) and a two-part ln2 (the high
//        part has 32 bits, so k * hi is exact); e^r by its Taylor series to
//        r^14 (|r| <= ln2/2, truncation < 1e-17); times 2^k.
//   log: x = m 2^e with m in [sqrt(1/2), sqrt(2)); s = (m - 1)/(m + 1);
//        log m = 2 s (1 + z/3 + z^2/5 + ... + z^12/25), z = s^2 <= 0.0295.
DEV double dm_exp(double x) {
    if (x != x) return x;
    if (x > 709.0) return HUGE_VAL;
    if (x < -708.0) return 0.0;
    const double k = rint(x * 1.4426950408889634);
    const double r = (x - k * 0x1.62e42feep-1) - k * 0x1.a39ef35793c76p-33;
    double p = 1.0 / 87178291200.0;
    p = 1.0 / 6227020800.0 + r * p;
    p = 1.0 / 479001600.0 + r * p;
    p = 1.0 / 39916800.0 + r * p;
    p = 1.0 / 3628800.0 + r * p;
    p = 1.0 / 362880.0 + r * p;
    p = 1.0 / 40320.0 + r * p;
    p = 1.0 / 5040.0 + r * p;
    p = 1.0 / 720.0 + r * p;
    p = 1.0 / 120.0 + r * p;
    p = 1.0 / 24.0 + r * p;
    p = 1.0 / 6.0 + r * p;
    p = 0.5 + r * p;
    p = 1.0 + r * p;
    p = 1.0 + r * p;
    return ldexp(p, (int)k);
}

DEV double dm_log(double x) {
    if (!(x > 0.0)) return x == 0.0 ? -HUGE_VAL : (x - x) / 0.0;
    int e;
    double m = frexp(x, &e);
    if (m < 0.70710678118654752) {
        m = m * 2.0;
        e = e - 1;
    }
    const double f = m - 1.0;
    const double s = f / (2.0 + f);
    const double z = s * s;
    double q = 1.0 / 25.0;
    q = 1.0 / 23.0 + z * q;
    q = 1.0 / 21.0 + z * q;
    q = 1.0 / 19.0 + z * q;
    q = 1.0 / 17.0 + z * q;
    q = 1.0 / 15.0 + z * q;
    q = 1.0 / 13.0 + z * q;
    q = 1.0 / 11.0 + z * q;
    q = 1.0 / 9.0 + z * q;
    q = 1.0 / 7.0 + z * q;
    q = 1.0 / 5.0 + z * q;
    q = 1.0 / 3.0 + z * q;
    const double lm = 2.0 * s + 2.0 * s * (z * q);
    const double de = (double)e;
    return de * 0x1.62e42feep-1 + (de * 0x1.a39ef35793c76p-33 + lm);
}

DEV double dm_pow(double x, double y) { return dm_exp(y * dm_log(x)); }

// normal force with the load/unload/reload verdict (kernel.cu:104-144 p-p,
// 250-290 p-w; YIELD_LE selects the p-p yield test "<=")
template <bool YIELD_LE>
DEV double normal_law(const DevLaw &c, HistRec &h, double d, double R) {
    int stage;
    if (d < h.pdfm) stage = 1;
    else if (d >= h.pdfm && d < h.nd) stage = 3;
    else if (d >= h.nd && d < h.ndm) stage = 4;
    else stage = 2;
    double fN = 0.0;
    if (stage == 2) {
        bool elastic = YIELD_LE ? (d <= c.ydsp_t * R) : (d < (c.ydsp_t * R));
        if (elastic) {
            fN = c.c43E * sqrt(R * d) * d;
            h.prad = R;
            h.pdfm = 0.0;
        } else {
            fN = (c.yforce_t * (R * R)) + (c.ystiff_t * R) * (d - (c.ydsp_t * R));
            h.prad = R * (c.c43E * sqrt(R * d) * d) / fN;
            h.pdfm = d - dm_pow((fN / (c.c43E * sqrt(R))), 2.0 / 3.0);
        }
        h.ndm = d;
    } else if (stage != 1) {
        fN = c.c43E * sqrt(h.prad) * pow15(d - h.pdfm);
    }
    h.nd = d;
    h.flags = (h.flags & ~0xff) | stage;
    return fN;
}

// tangential spring update (kernel.cu:172-203 p-p with 1e-7 tolerances,
// 306-337 p-w with 1e-10): rotate the stored spring into the tangent plane
// keeping its length, add v_t dt, cap it at xi_max.  Returns xi_max; t is
// the new spring (also stored in h) and u = 1 - |t|/xi_max where the spring
// sticks below the cap with xi_max above the tolerance (the Mindlin ratio is
// then 1 - u^1.5), else -1 (ratio 1).  tl: |t| where the spring stays
// below the cap (t is then the vector whose length was just taken, so the
// value is the one len(t) would give again), else -1.
template <bool PP>
DEV double spring_update(const DevLaw &c, HistRec &h, double fN, double d, double R, V3 n, V3 vt,
                         double dt, V3 &t, double &u, double &tl) {
    const double tol = PP ? 1.0e-7 : 1.0e-10;
    V3 old = v3(h.dx, h.dy, h.dz);
    t = cross(smulv(-1.0, n), cross(n, old));
    double lt = len(t);
    if (lt > tol)
        t = add(divv(mulv(t, len(old)), lt), mulv(vt, dt));
    else
        t = add(t, mulv(vt, dt));
    double dsmx = c.dsmxA * fN / (c.dsmxB * sqrt(R * d));
    double l2 = len(t);
    u = -1.0;
    tl = -1.0;
    if (l2 < dsmx) {
        if (dsmx > tol) u = 1.0 - l2 / dsmx;
        tl = l2;
    } else {
        if (l2 > tol)
            t = divv(smulv(dsmx, t), l2);
        else
            t = smulv(dsmx, t);
    }
    h.dx = t.x; h.dy = t.y; h.dz = t.z;
    return dsmx;
}

// tangential force (kernel.cu:205-217 p-p, 339-350 p-w): along the spring
// while it is below the cap, against the sliding velocity at the cap
template <bool PP>
DEV V3 tangential_law(const DevLaw &c, HistRec &h, double fN, double d, double R, V3 n, V3 vt,
                      double dt) {
    V3 t;
    double u, tl;
    const double dsmx = spring_update<PP>(c, h, fN, d, R, n, vt, dt, t, u, tl);
    const double ratio = u >= 0.0 ? 1.0 - pow15(u) : 1.0;
    V3 ft = v3(0.0, 0.0, 0.0);
    const double l3 = tl >= 0.0 ? tl : len(t);
    if (l3 < dsmx) {
        if (l3 > 1.0e-10) ft = divv(smulv(-1.0 * c.sfrc * fN * ratio, t), l3);
    } else {
        double lv = len(vt);
        if (lv > 1.0e-10) ft = divv(smulv(-1.0 * c.sfrc * fN * ratio, vt), lv);
    }
    return ft;
}

// one directed contact; cpv = contact-point relative velocity
template <bool PP>
DEV void contact(const DevLaw &c, double rolling, double dt, HistRec &h, double d, V3 n, double R,
                 double ri, V3 cpv, V3 &F, V3 &M, double &Mr) {
    double fN = normal_law<PP>(c, h, d, R);
    double cpvn = dot(cpv, n);
    double fD = 0.0;
    if (c.damp_on) fD = c.damp * sqrt(R * d) * cpvn;
    V3 vt = sub(cpv, smulv(cpvn, n));
    V3 ft = tangential_law<PP>(c, h, fN, d, R, n, vt, dt);
    V3 dF = add(smulv(fN + fD, n), ft);
    double arm = ri - d / 2.0;
    V3 dM = cross(smulv(-arm, n), ft);
    F = add(F, dF);
    M = add(M, dM);
    Mr = Mr + rolling * fN * ri;
}

// The same contact in mixed precision.  Everything that is stored or that
// decides a branch is computed in FP64 exactly as in contact(): the normal
// law with its load/unload verdict and plastic history, the tangential
// spring with its rotation, rescale and Coulomb cap, xi_max and the
// stick/slip test.  The reference's friction force turns from the spring
// direction to the sliding direction on a last-bit comparison at the cap
// (kernel.cu:190-217), so a verdict taken from FP32 values would flip
// against FP64 for every contact resting at the cap.  FP32: the Mindlin
// ratio's power, the damping magnitude and the force, moment and rolling
// terms of the pair, from the FP64 values above.
template <bool PP>
DEV void contact_mixed(const DevLaw &c, double rolling, double dt, HistRec &h, double d, V3 n,
                       double R, double ri, V3 cpv, V3 &F, V3 &M, double &Mr) {
    const double fN = normal_law<PP>(c, h, d, R);
    const double cpvn = dot(cpv, n);
    const V3 vt = sub(cpv, smulv(cpvn, n));
    V3 t;
    double u, tl;
    const double dsmx = spring_update<PP>(c, h, fN, d, R, n, vt, dt, t, u, tl);
    float ratio = 1.0f;
    if (u >= 0.0) {
        const float uf = (float)u;
        ratio = 1.0f - uf * sqrtf(uf);
    }
    const float fNf = (float)fN;
    const float sf = -(float)c.sfrc * fNf * ratio;
    float fx = 0.f, fy = 0.f, fz = 0.f;
    const double l3 = tl >= 0.0 ? tl : len(t);
    if (l3 < dsmx) {
        if (l3 > 1.0e-10) {
            const float k = sf / (float)l3;
            fx = k * (float)t.x; fy = k * (float)t.y; fz = k * (float)t.z;
        }
    } else {
        const double lv = len(vt);
        if (lv > 1.0e-10) {
            const float k = sf / (float)lv;
            fx = k * (float)vt.x; fy = k * (float)vt.y; fz = k * (float)vt.z;
        }
    }
    const float fD = c.damp_on ? (float)c.damp * sqrtf((float)(R * d)) * (float)cpvn : 0.0f;
    const float fn = fNf + fD;
    const float nx = (float)n.x, ny = (float)n.y, nz = (float)n.z;
    F.x += (double)(fn * nx + fx);
    F.y += (double)(fn * ny + fy);
    F.z += (double)(fn * nz + fz);
    const float arm = -(float)(ri - d / 2.0);
    const float ax = arm * nx, ay = arm * ny, az = arm * nz;
    M.x += (double)(ay * fz - az * fy);
    M.y += (double)(az * fx - ax * fz);
    M.z += (double)(ax * fy - ay * fx);
    Mr += (double)((float)rolling * fNf * (float)ri);
}

// Willett bridge force (kernel.cu:65-81) at the instance's contact angle
DEV double fcap(const DevInst &P, double S, double V, double effR) {
    S = fmax(P.gapmn, S);
    S = S / sqrt(V);
    double L = dm_log(V), L2 = L * L, L3 = L2 * L;
    const double *a = P.wc;
    double f1 = a[0] + a[1] * L + a[2] * L2 + a[3] * L3;
    double f2 = a[4] + a[5] * L + a[6] * L2;
    double f3 = a[7] + a[8] * L + a[9] * L2 + a[10] * L3;
    double f4 = a[11] + a[12] * L + a[13] * L2;
    double lS = dm_log(S);
    return P.twoPI * effR * P.stension * dm_exp(f1 - f2 * dm_exp(f3 * lS + f4 * (lS * lS)));
}

// The same force in FP32 for mixed precision: a force, stored nowhere and
// deciding no branch (the bridge's existence and rupture are decided in
// FP64 by the caller), from the same regression with hardware log and exp;
// the scaled gap S / sqrt(V) and the pair's effective radius 2 ri rj /
// (ri + rj), which feed only this force, are formed in FP32 as well
DEV double fcap_f32(const DevInst &P, double S, double V, double ri, double rj) {
    S = fmax(P.gapmn, S);
    const float sv = (float)S / sqrtf((float)V);
    const float rif = (float)ri, rjf = (float)rj;
    const float effR = 2.0f * rif * rjf / (rif + rjf);
    const float L = __logf((float)V), L2 = L * L, L3 = L2 * L;
    const float *a = P.wcf;  // the coefficients rounded to FP32 once, on the host
    const float f1 = a[0] + a[1] * L + a[2] * L2 + a[3] * L3;
    const float f2 = a[4] + a[5] * L + a[6] * L2;
    const float f3 = a[7] + a[8] * L + a[9] * L2 + a[10] * L3;
    const float f4 = a[11] + a[12] * L + a[13] * L2;
    const float lS = __logf(sv);
    const float e = __expf(f1 - f2 * __expf(f3 * lS + f4 * (lS * lS)));
    return (double)((float)P.twoPI * effR * (float)P.stension * e);
}

DEV double rupture(const DevInst &P, double V) {
    return P.ruptureMode == 1 ? (1.0 + 0.5 * P.theta) * dm_pow(V, 1.0 / 3.0)
                              : 0.99 * dm_pow(V, 0.33);
}

// liquid volume of a particle of radius r with nl bridges in the last step
// (kernel.cu:429-434, 511-518: the particle's share, at least 1/12)
DEV double liquid_volume(const DevInst &P, double r, int nl) {
    if (nl < 12) nl = 12;
    const double d = r * 2;
    return P.vol * (d * d * d) / nl;
}

// The capillary block of one p-p slot (kernel.cu:509-550): the bridge
// forms below layer2 (ri + rj), ruptures past the rupture distance of the
// pair's liquid, and pulls i (only) along -n.  fl is the slot's flags word
// (load stage in bits 0-7, bridge in bit 8); the new one is returned.  The
// partner's bridge count (jnl_of) and the liquid volume are evaluated only
// where they decide or act: a bridge that holds unless it ruptures, or a
// forming one.
template <bool MIXED, class JNL>
DEV int bridge_slot(const DevInst &P, int fl, double gap, V3 nv, double ri, double rj, double iVol,
                    JNL jnl_of, V3 &F, int &nbridge) {
    int brg = (fl >> 8) & 1;
    double V = 0.0;
    if (brg || gap < P.layer2 * (ri + rj)) V = iVol + liquid_volume(P, rj, jnl_of());
    if (gap < P.layer2 * (ri + rj)) brg = 1;
    else if (brg && gap > rupture(P, V)) brg = 0;
    if (brg == 1) {
        const double Sp = gap < 0 ? 0.0 : gap;
        V = V * 8.0;
        const double fmag =
            MIXED ? fcap_f32(P, Sp, V, ri, rj) : fcap(P, Sp, V, 2 * ri * rj / (ri + rj));
        const V3 fc = smulv(-1.0 * fmag, nv);
        F = add(F, fc);
        nbridge = nbridge + 1;
    }
    return (fl & 0xff) | (brg << 8);
}

DEV V3 vel_of(const double4 *v, int i) {
    const double4 a = v[2 * i];
    return v3(a.x, a.y, a.z);
}
DEV V3 spin_of(const double4 *v, int i) {
    const double4 a = v[2 * i], b = v[2 * i + 1];
    return v3(a.w, b.x, b.y);
}

// gravity + all p-p slots of particle i (kernel.cu:424-551)
template <bool WET, bool MIXED>
DEV void pp_particle(const DevInst &P, const StepArgs &A, int i, const double4 pi, V3 iv, V3 iw,
                     V3 &F, V3 &M, double &Mr, int &nbridge) {
    const StateBuf &s = A.in;
    const V3 ip = v3(pi.x, pi.y, pi.z);
    const double ri = pi.w;
    const double iVol = WET ? liquid_volume(P, ri, s.lqd[i]) : 0.0;
    double di = 2.0 * ri;
    F.z = F.z - di * di * di;
    const int n = A.nnb[i];
    const unsigned long long am = A.amask[i];
    const int aoff = A.actoff[i];
    const long Kcap = (long)A.K * A.cap;
    for (int k = 0; k < n; ++k) {
        const long sk = (long)k * A.cap + i;
        const int j = A.nbr[sk];
        const double4 pj = s.p[j];
        const V3 jp = v3(pj.x, pj.y, pj.z);
        const double rj = pj.w;
        const V3 dv = sub(ip, jp);
        const double l = len(dv);
        const V3 nv = divv(dv, l);
        const double gap = l - ri - rj;
        HistRec h;
        bool loaded = false;
        const long hi = hist_idx(am, aoff, k, i, A.cap, Kcap);
        if (gap < 0.0) {
            const V3 jv = vel_of(s.v, j);
            const V3 jw = spin_of(s.v, j);
            h = A.h[hi];
            loaded = true;
            double R = ri * rj / (ri + rj);
            V3 cpv = sub(sub(iv, jv), cross(add(mulv(iw, ri), mulv(jw, rj)), nv));
            if (MIXED)
                contact_mixed<true>(P.pp, P.rolling, P.dt, h, -gap, nv, R, ri, cpv, F, M, Mr);
            else
                contact<true>(P.pp, P.rolling, P.dt, h, -gap, nv, R, ri, cpv, F, M, Mr);
        }
        if (WET && P.forceCap == 1) {
            const int fl = loaded ? h.flags : A.h[hi].flags;
            const int nfl = bridge_slot<MIXED>(P, fl, gap, nv, ri, rj, iVol,
                                               [&]() { return s.lqd[j]; }, F, nbridge);
            if (loaded)
                h.flags = nfl;
            else if (nfl != fl)
                A.h[hi].flags = nfl;
        }
        if (loaded) A.h[hi] = h;
    }
}

// contact point, normal and gap of particle (ip, ri) against wall w: lid
// (x, y, Tw), floor (x, y, Bw), or the lateral cylinder point at radius R
// (kernel.cu:987-1014)
DEV void wall_geometry(const DevInst &P, double Tw, const V3 ip, double ri, int w, V3 &nv,
                       double &gap) {
    V3 jp;
    if (w == 0) {
        jp = v3(ip.x, ip.y, Tw);
    } else if (w == 1) {
        jp = v3(ip.x, ip.y, P.Bw);
    } else {
        double ax = P.cirX - ip.x, ay = P.cirY - ip.y;
        double dc = sqrt(ax * ax + ay * ay);
        double cs = (ip.x - P.cirX) / dc, sn = (ip.y - P.cirY) / dc;
        jp.x = P.R * cs + P.cirX;
        jp.y = P.R * sn + P.cirY;
        jp.z = ip.z;
    }
    const V3 dv = sub(ip, jp);
    const double l = len(dv);
    nv = divv(dv, l);
    gap = l - ri;
}

// walls of particle i (kernel.cu:958-1023, 370-403).  The wall velocity is
// not reset between slots (kernel.cu:979-1020): once the lid has been seen,
// later slots use the lid velocity.
template <bool MIXED>
DEV void pw_particle(const DevInst &P, const StepArgs &A, int i, const double4 pi, V3 iv, V3 iw,
                     V3 &F, V3 &M, double &Mr) {
    const int n = A.nwl[i];
    if (n == 0) return;
    const V3 ip = v3(pi.x, pi.y, pi.z);
    const double ri = pi.w;
    V3 jv = v3(0.0, 0.0, 0.0);
    for (int k = 0; k < n; ++k) {
        const long sk = (long)k * A.cap + i;
        const int w = A.wl[sk];
        if (w == 0) jv = v3(0.0, 0.0, P.topv);
        V3 nv;
        double gap;
        wall_geometry(P, A.Tw, ip, ri, w, nv, gap);
        if (gap < 0.0) {
            HistRec h = A.hw[sk];
            V3 cpv = sub(sub(iv, jv), cross(mulv(iw, ri), nv));
            if (MIXED)
                contact_mixed<false>(P.pw, P.rolling, P.dt, h, -gap, nv, ri, ri, cpv, F, M, Mr);
            else
                contact<false>(P.pw, P.rolling, P.dt, h, -gap, nv, ri, ri, cpv, F, M, Mr);
            A.hw[sk] = h;
        }
    }
}

// semi-implicit Euler (kernel.cu:563-606); returns true when the particle's
// path length passed the rebuild threshold
DEV bool integrate(const DevInst &P, const StepArgs &A, int i, const double4 pi, const double4 va,
                   const double4 vb, V3 F, V3 M, double Mr, double tag = 0.0) {
    const double dt = P.dt;
    const double di = 2.0 * pi.w;
    const double mass = di * di * di;
    double vx = va.x + F.x * dt / mass;
    double vy = va.y + F.y * dt / mass;
    double vz = va.z + F.z * dt / mass;
    double4 np;
    np.x = pi.x + vx * dt;
    np.y = pi.y + vy * dt;
    np.z = pi.z + vz * dt;
    np.w = pi.w;
    double pm = vb.z + sqrt(vx * vx + vy * vy + vz * vz) * dt;
    const double d2 = di * di;
    const double inert = 0.1 * (d2 * d2 * di);
    double wx = va.w + M.x * dt / inert;
    double wy = vb.x + M.y * dt / inert;
    double wz = vb.y + M.z * dt / inert;
    double mr = Mr * dt / inert;
    double al = sqrt(wx * wx + wy * wy + wz * wz);
    if (al > 1.0e-7) {
        double f = (1.0 - fmin(1.0, mr / al));
        wx = wx * f; wy = wy * f; wz = wz * f;
    }
    A.out.p[i] = np;
    double4 o0, o1;
    o0.x = vx; o0.y = vy; o0.z = vz; o0.w = wx;
    o1.x = wy; o1.y = wz; o1.z = pm; o1.w = tag;  // tag: the pair kernel's contact mask
    A.out.v[2 * i] = o0;
    A.out.v[2 * i + 1] = o1;
    if (A.store_forces) A.Mr[i] = mr;
    return pm > P.rebuildThr;
}

// A wave that saw any particle over the threshold lowers stop_at to the
// next step: that step and all later ones in the launch chunk skip, and the
// host rebuilds before re-running them.
DEV void request_rebuild(const StepArgs &A, bool mine) {
    unsigned long long b = __ballot(mine);
    if (b != 0ull && (threadIdx.x & 63) == __ffsll((long long)b) - 1)
        atomicMin(A.stop_at, A.step + 1);
}

// one thread per particle: block blk of 256 particles.  The rebuild
// request (stop, loaded by the caller; INT_MAX, "none due", when it
// checked it itself)
// is tested once the particle's own records are in flight, so its round
// trip overlaps theirs.
template <bool WET, bool MIXED>
DEV void step_loop_block(const DevInst &P, const StepArgs &A, unsigned blk, int stop = INT_MAX) {
    const int i = (int)(blk * 256u + threadIdx.x);
    bool over = false;
    double4 pi = make_double4(0.0, 0.0, 0.0, 0.0), va = pi, vb = pi;
    if (i < A.total) {
        pi = A.in.p[i];
        va = A.in.v[2 * i];
        vb = A.in.v[2 * i + 1];
    }
    if (__builtin_amdgcn_readfirstlane(stop) <= A.step) return;
    if (i < A.total) {
        const V3 iv = v3(va.x, va.y, va.z), iw = v3(va.w, vb.x, vb.y);
        V3 F = v3(0.0, 0.0, 0.0), M = v3(0.0, 0.0, 0.0);
        double Mr = 0.0;
        int nbridge = 0;
        pp_particle<WET, MIXED>(P, A, i, pi, iv, iw, F, M, Mr, nbridge);
        if (WET) A.out.lqd[i] = nbridge;
        pw_particle<MIXED>(P, A, i, pi, iv, iw, F, M, Mr);
        if (A.store_forces) {
            A.Fx[i] = F.x; A.Fy[i] = F.y; A.Fz[i] = F.z;
            A.Mx[i] = M.x; A.My[i] = M.y; A.Mz[i] = M.z;
        }
        // wet: the second velocity record is read again (from L2) rather than
        // kept in registers through the slot loops, which would spill it to
        // scratch at 4 waves per SIMD (the pour 3 % faster, a dense wet bed
        // 1-2 %); the dry forms keep it, where the reload costs 1 %
        const double4 vbe = WET ? A.in.v[2 * i + 1] : vb;
        over = integrate(P, A, i, pi, va, vbe, F, M, Mr);
    }
    request_rebuild(A, over);
}

// Held to 4 waves per SIMD (128 registers; no spill code is emitted): the
// kernel is bound by each thread's chain of dependent gathers, and a
// fourth wave hides more of it than a third (a dense wet bed of 200,000
// spheres: 9 % faster in FP64, 15 % in mixed precision, measured when the
// hint still cost a few spilled registers).
template <bool WET, bool MIXED>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
k_step_fused(DevInst P, StepArgs A) {
    // a rebuild is due before this step when stop <= A.step
    const int stop = __hip_atomic_load(A.stop_at, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    step_loop_block<WET, MIXED>(P, A, xcd_block(gridDim.x), stop);
}

void launch_step_fused(const DevInst &P, const StepArgs &A, bool wet, bool mixed, hipStream_t s) {
    if (A.total <= 0) return;
    unsigned nb = blocks_for(A.total, 256);
    if (wet && mixed)
        hipLaunchKernelGGL((k_step_fused<true, true>), dim3(nb), dim3(256), 0, s, P, A);
    else if (wet)
        hipLaunchKernelGGL((k_step_fused<true, false>), dim3(nb), dim3(256), 0, s, P, A);
    else if (mixed)
        hipLaunchKernelGGL((k_step_fused<false, true>), dim3(nb), dim3(256), 0, s, P, A);
    else
        hipLaunchKernelGGL((k_step_fused<false, false>), dim3(nb), dim3(256), 0, s, P, A);
}

// ---------------------------------------------------------------------
// Dry fused step with contact compaction.
// A block owns 256 consecutive particles.  Phase A: each thread tests the
// overlap of all its neighbour slots (cheap) and records the contacting
// slots in a bitmask.  Phase B: a block scan packs every (particle, slot)
// contact into an LDS list, particle by particle in slot order.  Phase C:
// all 256 lanes take list entries in rounds and evaluate the contact law
// (the expensive, branchy part) with every lane busy, writing the pair's
// force, moment and rolling term to LDS; after each round the owner of each
// particle adds its entries in list order, i.e. gravity first, then its
// contacts in slot order, then its walls: the summation order of
// kernel.cu:439-498, so the FP64 path stays bit-identical to the one-thread
// loop.  Requires K <= 64 (slot bitmask).
// ---------------------------------------------------------------------
#define CP_BS 256

// overlap test, gap and normal of a pair, in FP64 in both precisions (the
// normal and the overlap feed the stored history)
DEV bool pair_geometry(const double4 pi, const double4 pj, V3 &nv, double &gap) {
    const V3 dv = v3(pi.x - pj.x, pi.y - pj.y, pi.z - pj.z);
    const double l = len(dv);
    gap = l - pi.w - pj.w;
    if (!(gap < 0.0)) return false;
    nv = divv(dv, l);
    return true;
}

template <bool MIXED>
__global__ void __launch_bounds__(CP_BS) k_step_compact(DevInst P, StepArgs A) {
    if (__hip_atomic_load(A.stop_at, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <= A.step) return;
    extern __shared__ double smem_d[];
    double *res = smem_d;                                     // [7][CP_BS]
    unsigned short *list = (unsigned short *)(smem_d + 7 * CP_BS);
    __shared__ unsigned scan_lds[4];
    __shared__ unsigned total_s;

    const unsigned nb = gridDim.x;
    const int base = (int)(xcd_block(nb) * CP_BS);
    const int tid = threadIdx.x;
    const int i = base + tid;
    const bool own = i < A.total;

    // phase A: own record and the contact bitmask
    double4 pi = make_double4(0.0, 0.0, 0.0, 0.0), va = pi, vb = pi;
    unsigned long long mask = 0ull;
    if (own) {
        pi = A.in.p[i];
        va = A.in.v[2 * i];
        vb = A.in.v[2 * i + 1];
        const int n = A.nnb[i];
        for (int k = 0; k < n; ++k) {
            const int j = A.nbr[(long)k * A.cap + i];
            V3 nv;
            double gap;
            if (pair_geometry(pi, A.in.p[j], nv, gap)) mask |= 1ull << k;
        }
    }
    // phase B: pack (particle, slot) pairs in slot order
    const unsigned cnt = (unsigned)__popcll(mask);
    unsigned tot;
    const unsigned off = block_excl_scan(cnt, &tot, scan_lds);
    {
        unsigned long long m = mask;
        unsigned q = off;
        while (m) {
            const int k = __ffsll((long long)m) - 1;
            m &= m - 1ull;
            list[q++] = (unsigned short)((tid << 8) | k);
        }
    }
    if (tid == 0) total_s = tot;
    __syncthreads();
    const unsigned T = total_s;

    double Fx = 0.0, Fy = 0.0, Fz = 0.0, Mx = 0.0, My = 0.0, Mz = 0.0, Mr = 0.0;
    if (own) {
        const double di = 2.0 * pi.w;
        Fz = Fz - di * di * di;  // gravity first (kernel.cu:439)
    }
    // phase C: evaluate the packed contacts in rounds of CP_BS
    for (unsigned r0 = 0; r0 < T; r0 += CP_BS) {
        const unsigned t = r0 + tid;
        if (t < T) {
            const unsigned e = list[t];
            const int il = (int)(e >> 8), k = (int)(e & 0xff);
            const int ii = base + il;
            const long sk = (long)k * A.cap + ii;
            const int j = A.nbr[sk];
            const double4 qi = A.in.p[ii], qj = A.in.p[j];
            V3 nv;
            double gap;
            pair_geometry(qi, qj, nv, gap);
            const V3 iv = vel_of(A.in.v, ii), iw = spin_of(A.in.v, ii);
            const V3 jv = vel_of(A.in.v, j), jw = spin_of(A.in.v, j);
            const long hi = hist_idx(A.amask[ii], A.actoff[ii], k, ii, A.cap, (long)A.K * A.cap);
            HistRec h = A.h[hi];
            const double ri = qi.w, rj = qj.w;
            const double R = ri * rj / (ri + rj);
            const V3 cpv = sub(sub(iv, jv), cross(add(mulv(iw, ri), mulv(jw, rj)), nv));
            V3 dF = v3(0.0, 0.0, 0.0), dM = v3(0.0, 0.0, 0.0);
            double dMr = 0.0;
            if (MIXED)
                contact_mixed<true>(P.pp, P.rolling, P.dt, h, -gap, nv, R, ri, cpv, dF, dM, dMr);
            else
                contact<true>(P.pp, P.rolling, P.dt, h, -gap, nv, R, ri, cpv, dF, dM, dMr);
            A.h[hi] = h;
            res[0 * CP_BS + tid] = dF.x; res[1 * CP_BS + tid] = dF.y; res[2 * CP_BS + tid] = dF.z;
            res[3 * CP_BS + tid] = dM.x; res[4 * CP_BS + tid] = dM.y; res[5 * CP_BS + tid] = dM.z;
            res[6 * CP_BS + tid] = dMr;
        }
        __syncthreads();
        if (cnt) {
            unsigned a = off > r0 ? off : r0;
            unsigned b = off + cnt < r0 + CP_BS ? off + cnt : r0 + CP_BS;
            for (unsigned q = a; q < b; ++q) {
                const unsigned s = q - r0;
                Fx = Fx + res[0 * CP_BS + s]; Fy = Fy + res[1 * CP_BS + s];
                Fz = Fz + res[2 * CP_BS + s];
                Mx = Mx + res[3 * CP_BS + s]; My = My + res[4 * CP_BS + s];
                Mz = Mz + res[5 * CP_BS + s];
                Mr = Mr + res[6 * CP_BS + s];
            }
        }
        __syncthreads();
    }
    // phase D: walls and integration by the owner
    bool over = false;
    if (own) {
        V3 F = v3(Fx, Fy, Fz), M = v3(Mx, My, Mz);
        const V3 iv = v3(va.x, va.y, va.z), iw = v3(va.w, vb.x, vb.y);
        pw_particle<MIXED>(P, A, i, pi, iv, iw, F, M, Mr);
        if (A.store_forces) {
            A.Fx[i] = F.x; A.Fy[i] = F.y; A.Fz[i] = F.z;
            A.Mx[i] = M.x; A.My[i] = M.y; A.Mz[i] = M.z;
        }
        over = integrate(P, A, i, pi, va, vb, F, M, Mr);
    }
    request_rebuild(A, over);
}

size_t compact_lds_bytes(int K) { return (size_t)7 * CP_BS * sizeof(double) + (size_t)CP_BS * K * 2; }

void launch_step_compact(const DevInst &P, const StepArgs &A, int K, bool mixed, hipStream_t s) {
    if (A.total <= 0) return;
    unsigned nb = blocks_for(A.total, CP_BS);
    size_t lds = compact_lds_bytes(K);
    if (mixed)
        hipLaunchKernelGGL((k_step_compact<true>), dim3(nb), dim3(CP_BS), lds, s, P, A);
    else
        hipLaunchKernelGGL((k_step_compact<false>), dim3(nb), dim3(CP_BS), lds, s, P, A);
}

// ---------------------------------------------------------------------
// Dry fused step with per-wave contact compaction (no block barriers, no
// LDS contact list).  Each wave owns 64 consecutive particles.
//  A: the owner tests its neighbour slots eight at a time (eight index
//     loads, then eight position loads in flight) and its wall slots,
//     keeping a 64-bit pair mask and a 3-bit wall mask.
//  B: a wave prefix sum numbers the wave's contacts 0..T-1, particle by
//     particle, pairs in slot order, then walls in slot order.
//  C: in rounds of 64, lane t finds its contact's owner by binary search
//     over the lanes' offsets and its slot by selecting the r-th set bit of
//     the owner's masks (register shuffles only), evaluates the contact law
//     and puts the force, moment and rolling term in the wave's LDS slice;
//     each owner then adds its entries in list order.  The next round's
//     neighbour indices and history records load while a round computes.
//  D: integration by the owner.
// The owners' records stay in an LDS stash from A to D (not in registers,
// which keeps the kernel at 4 waves per SIMD, and not re-read from memory).
// The summation order per particle is gravity, contacts in slot order,
// walls in slot order (kernel.cu:439-498, 958-1023), as in the one-thread
// loop, so the FP64 results are bit-identical to it.  Requires K <= 64.
// ---------------------------------------------------------------------
#define WV_BS 256
#define WV_NB 8
#define WV_TAIL 192u  // chunks per XCD done as half-waves at the end of a launch
#define WV_QTAIL 32u  // the last of those done as quarter-waves

// per-wave LDS copy of the owners' records (read by the contact lanes of
// phase C and by the owners in phase D instead of re-reading them)
struct WvStash {
    double4 p[64], va[64], vb[64];
};

// wet packings: the owners' liquid volumes (read by the contact lanes) and
// each entry's bridge bit of the current round (summed by its owner)
struct WvWet {
    double vol[64];
    int nb[64];
};

DEV void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// position of the r-th (0-based) set bit of m; m has more than r bits set
DEV int select_bit(unsigned long long m, unsigned r) {
    int pos = 0;
    unsigned c = (unsigned)__popc((unsigned)m);
    if (r >= c) { r -= c; m >>= 32; pos += 32; }
    c = (unsigned)__popc((unsigned)(m & 0xffffull));
    if (r >= c) { r -= c; m >>= 16; pos += 16; }
    c = (unsigned)__popc((unsigned)(m & 0xffull));
    if (r >= c) { r -= c; m >>= 8; pos += 8; }
    c = (unsigned)__popc((unsigned)(m & 0xfull));
    if (r >= c) { r -= c; m >>= 4; pos += 4; }
    c = (unsigned)__popc((unsigned)(m & 0x3ull));
    if (r >= c) { r -= c; m >>= 2; pos += 2; }
    c = (unsigned)(m & 1ull);
    if (r >= c) pos += 1;
    return pos;
}

// the contact law of one particle-particle entry of the compacted list
// (history record h in, updated record out)
template <bool MIXED>
DEV void pair_compute(const DevInst &P, HistRec &h, const double4 qi, const double4 ai,
                      const double4 bi, const double4 qj, const double4 aj, const double4 bj, V3 &dF,
                      V3 &dM, double &dMr) {
    const V3 iv = v3(ai.x, ai.y, ai.z), iw = v3(ai.w, bi.x, bi.y);
    const V3 jv = v3(aj.x, aj.y, aj.z), jw = v3(aj.w, bj.x, bj.y);
    V3 nv;
    double gap;
    pair_geometry(qi, qj, nv, gap);
    const double ri = qi.w, rj = qj.w;
    const double R = ri * rj / (ri + rj);
    const V3 cpv = sub(sub(iv, jv), cross(add(mulv(iw, ri), mulv(jw, rj)), nv));
    if (MIXED)
        contact_mixed<true>(P.pp, P.rolling, P.dt, h, -gap, nv, R, ri, cpv, dF, dM, dMr);
    else
        contact<true>(P.pp, P.rolling, P.dt, h, -gap, nv, R, ri, cpv, dF, dM, dMr);
}

// one wet particle-particle entry: contact law where the spheres overlap,
// then the capillary block of the slot, both as pp_particle evaluates them
// (kernel.cu:458-550).  dF, dM, dMr take the contact's terms, Fc the
// bridge force; the owner adds them in this order.  Returns 2 when the
// record must be stored (a contact), 1 when only its flags changed, else 0.
template <bool MIXED>
DEV int wet_pair_compute(const DevInst &P, HistRec &h, const double4 qi, const double4 ai,
                         const double4 bi, const double4 qj, const double4 aj, const double4 bj,
                         double iVol, int jnl, V3 &dF, V3 &dM, double &dMr, V3 &Fc, int &nb) {
    const V3 dv = v3(qi.x - qj.x, qi.y - qj.y, qi.z - qj.z);
    const double l = len(dv);
    const V3 nv = divv(dv, l);
    const double ri = qi.w, rj = qj.w;
    const double gap = l - ri - rj;
    const bool touch = gap < 0.0;
    if (touch) {
        const V3 iv = v3(ai.x, ai.y, ai.z), iw = v3(ai.w, bi.x, bi.y);
        const V3 jv = v3(aj.x, aj.y, aj.z), jw = v3(aj.w, bj.x, bj.y);
        const double R = ri * rj / (ri + rj);
        const V3 cpv = sub(sub(iv, jv), cross(add(mulv(iw, ri), mulv(jw, rj)), nv));
        if (MIXED)
            contact_mixed<true>(P.pp, P.rolling, P.dt, h, -gap, nv, R, ri, cpv, dF, dM, dMr);
        else
            contact<true>(P.pp, P.rolling, P.dt, h, -gap, nv, R, ri, cpv, dF, dM, dMr);
    }
    const int fl = h.flags;
    h.flags = bridge_slot<MIXED>(P, fl, gap, nv, ri, rj, iVol, [&]() { return jnl; }, Fc, nb);
    return touch ? 2 : (h.flags != fl ? 1 : 0);
}

// one wall entry: slot kw of particle ii, wall w.  The wall velocity is the
// lid's once the lid has been seen in this or an earlier slot
// (kernel.cu:979-1020).
template <bool MIXED>
DEV void wall_entry(const DevInst &P, const StepArgs &A, int ii, int kw, int w, const double4 qi,
                    const double4 ai, const double4 bi, V3 &dF, V3 &dM, double &dMr) {
    const long cap = A.cap;
    const long sk = (long)kw * cap + ii;
    HistRec h = A.hw[sk];
    const V3 iv = v3(ai.x, ai.y, ai.z), iw = v3(ai.w, bi.x, bi.y);
    bool lid = w == 0;
    for (int k = 0; k < kw; ++k) lid = lid || A.wl[(long)k * cap + ii] == 0;
    const V3 jv = v3(0.0, 0.0, lid ? P.topv : 0.0);
    const double ri = qi.w;
    V3 nv;
    double gap;
    wall_geometry(P, A.Tw, v3(qi.x, qi.y, qi.z), ri, w, nv, gap);
    const V3 cpv = sub(sub(iv, jv), cross(mulv(iw, ri), nv));
    if (MIXED)
        contact_mixed<false>(P.pw, P.rolling, P.dt, h, -gap, nv, ri, ri, cpv, dF, dM, dMr);
    else
        contact<false>(P.pw, P.rolling, P.dt, h, -gap, nv, ri, ri, cpv, dF, dM, dMr);
    A.hw[sk] = h;
}

// a compacted entry: particle ii (wave-relative o), slot k, pair or wall
struct WvEntry {
    int ii, k;
    bool wall, valid;
    unsigned long long am;  // the owner's compact-history mask and offset
    int aoff;
};

// entry t of the wave's list; every lane of the wave must call this (it
// reads the other lanes' offsets and masks through shuffles)
DEV WvEntry wave_entry(unsigned t, unsigned T, unsigned off, unsigned mlo, unsigned mhi,
                       unsigned wmask, int wbase, unsigned long long am, int aoff) {
    const unsigned tq = t < T ? t : T - 1;
    int o = 0;
#pragma unroll
    for (int s = 32; s > 0; s >>= 1) {
        const unsigned om = (unsigned)__shfl((int)off, o + s, 64);
        if (om <= tq) o += s;
    }
    const unsigned olo = (unsigned)__shfl((int)mlo, o, 64);
    const unsigned ohi = (unsigned)__shfl((int)mhi, o, 64);
    const unsigned ow = (unsigned)__shfl((int)wmask, o, 64);
    const unsigned ooff = (unsigned)__shfl((int)off, o, 64);
    const unsigned alo = (unsigned)__shfl((int)(unsigned)am, o, 64);
    const unsigned ahi = (unsigned)__shfl((int)(unsigned)(am >> 32), o, 64);
    const int aof = __shfl(aoff, o, 64);
    const unsigned long long om = ((unsigned long long)ohi << 32) | olo;
    const unsigned r = tq - ooff, np = (unsigned)__popcll(om);
    WvEntry e;
    e.ii = wbase + o;
    e.valid = t < T;
    e.wall = r >= np;
    e.k = e.wall ? select_bit(ow, r - np) : select_bit(om, r);
    e.am = ((unsigned long long)ahi << 32) | alo;
    e.aoff = aof;
    return e;
}

// entry t of the wave's list in pair mode (shared pair records, see
// PAIR_MIRROR): per particle its contact pairs in near-list (= slot)
// order, its walls in slot order, then the owned contacts of the previous
// step that ended now, whose records are copied forward once; pair masks
// are over near-list positions.  Every lane of the wave must call this.
struct WpEntry {
    int ii, k, kind;  // kind 0 pair, 1 wall, 2 copy forward; k: near-list position or wall slot
    bool valid;
    int nidx;         // pairs and copies: index of the entry in the near list
};

DEV WpEntry wave_entry_pair(unsigned t, unsigned T, unsigned off, unsigned mlo, unsigned mhi,
                            unsigned wmask, unsigned blo, unsigned bhi, int noff, int wbase) {
    const unsigned tq = t < T ? t : T - 1;
    int o = 0;
#pragma unroll
    for (int s = 32; s > 0; s >>= 1) {
        const unsigned om = (unsigned)__shfl((int)off, o + s, 64);
        if (om <= tq) o += s;
    }
    const unsigned olo = (unsigned)__shfl((int)mlo, o, 64);
    const unsigned ohi = (unsigned)__shfl((int)mhi, o, 64);
    const unsigned ow = (unsigned)__shfl((int)wmask, o, 64);
    const unsigned ooff = (unsigned)__shfl((int)off, o, 64);
    const unsigned oblo = (unsigned)__shfl((int)blo, o, 64);
    const unsigned obhi = (unsigned)__shfl((int)bhi, o, 64);
    const int onoff = __shfl(noff, o, 64);
    const unsigned long long om = ((unsigned long long)ohi << 32) | olo;
    const unsigned long long ob = ((unsigned long long)obhi << 32) | oblo;
    const unsigned r = tq - ooff, np = (unsigned)__popcll(om), nw = (unsigned)__popc(ow);
    WpEntry e;
    e.ii = wbase + o;
    e.valid = t < T;
    if (r < np) {
        e.kind = 0;
        e.k = select_bit(om, r);
    } else if (r < np + nw) {
        e.kind = 1;
        e.k = select_bit(ow, r - np);
    } else {
        e.kind = 2;
        e.k = select_bit(ob, r - np - nw);
    }
    e.nidx = onoff + e.k;
    return e;
}

// block blk of BS particles (one wave per 64; lanes outside [lane_lo,
// lane_hi) own no particle, for half- and quarter-waves); res is this
// wave's [7][64] LDS slice ([10][64] when WET).  PAIR: the history comes
// from the shared pair records (see PAIR_MIRROR) instead of one record per
// directed slot; the forces are the same bit for bit.
// WET (capillary bridges, directed records only): an entry is a slot in
// contact, bridged, or close enough for a bridge to form (its flags are
// read in phase A); it evaluates the contact law where the spheres overlap
// and then the slot's capillary block, and its owner adds the contact's
// terms, then the bridge force, in slot order: pp_particle's order.
// A particle with more than 64 listed slots (polydisperse lists) does not
// fit the slot masks: its owner steps it alone in phase D with the loop
// kernel's code, so the kernel serves any list length.
template <bool MIXED, bool PAIR, int BS = WV_BS, bool WET = false>
DEV void step_wave_block(const DevInst &P, const StepArgs &A, unsigned blk, double (*res)[64],
                         WvStash *st, WvWet *wt = nullptr, int lane_lo = 0, int lane_hi = 64,
                         int stop = INT_MAX) {
    static_assert(!(WET && PAIR), "wet packings keep directed records");
    const int tid = threadIdx.x, lane = tid & 63;
    const int i = (int)(blk * BS) + tid;
    const int wbase = i - lane;
    const bool own = i < A.total && lane >= lane_lo && lane < lane_hi;
    const long cap = A.cap;
    const long Kcap = (long)A.K * cap;
    bool heavy = false;  // more than 64 slots: stepped whole by its owner in phase D

    // phase A: contact masks of the pair slots and the wall slots.  Only
    // the slots of nmask are tested (the others cannot touch before the
    // next rebuild), WV_NB at a time.  Pair mode walks the particle's near
    // list instead (the same slots, in slot order; masks are over its
    // positions) and also notes which contacts own their record (ownm) and
    // which owned contacts of the previous step (the mask in the spare word
    // of the velocity record) ended (brk).
    unsigned long long mask = 0ull, am_own = 0ull, ownm = 0ull, brk = 0ull;
    unsigned wmask = 0u;
    int aoff_own = 0;
    double grav = 0.0;
    // The wave's first loads and the near-slot gathers that depend on them
    // are issued at raised priority, so a new wave reaches its contact
    // rounds sooner while older waves compute.  The rebuild request (stop,
    // loaded by the caller; INT_MAX, "none due", when it checked it itself) is tested
    // only once the particle's own records are in flight, so its round
    // trip overlaps theirs; nothing has been written before the test.
    double4 pi = make_double4(0.0, 0.0, 0.0, 0.0), vai = pi, vbi = pi;
    int nw = 0, nend = 0;
    __builtin_amdgcn_s_setprio(2);
    if (own) {
        if constexpr (PAIR) {
            aoff_own = A.noff[i];
            nend = A.noff[i + 1];
        } else {
            am_own = A.amask[i];
            aoff_own = A.actoff[i];
        }
        pi = A.in.p[i];
        nw = A.nwl[i];
        vbi = A.in.v[2 * i + 1];
        vai = A.in.v[2 * i];
    }
    if (__builtin_amdgcn_readfirstlane(stop) <= A.step) {
        __builtin_amdgcn_s_setprio(0);
        return;
    }
    if (own) {
        st->p[lane] = pi;
        st->va[lane] = vai;
        st->vb[lane] = vbi;
        if constexpr (PAIR) {
            const int nn = nend - aoff_own;
            for (int r0 = 0; r0 < nn; r0 += WV_NB) {
                int jj[WV_NB];
#pragma unroll
                for (int u = 0; u < WV_NB; ++u) jj[u] = r0 + u < nn ? A.nl[aoff_own + r0 + u].x : i;
                double4 q[WV_NB];
#pragma unroll
                for (int u = 0; u < WV_NB; ++u) q[u] = A.in.p[jj[u]];
#pragma unroll
                for (int u = 0; u < WV_NB; ++u) {
                    V3 nv;
                    double gap;
                    if (r0 + u < nn && pair_geometry(pi, q[u], nv, gap)) {
                        mask |= 1ull << (r0 + u);
                        if (owns_record(i, jj[u], pi.w, q[u].w)) ownm |= 1ull << (r0 + u);
                    }
                }
            }
        } else if constexpr (WET) {
            // every listed slot: a bridge holds beyond the near gap
            const int n = A.nnb[i];
            heavy = n > 64;
            const int nt = heavy ? 0 : n;
            for (int r0 = 0; r0 < nt; r0 += WV_NB) {
                int jj[WV_NB], fl[WV_NB];
#pragma unroll
                for (int u = 0; u < WV_NB; ++u) jj[u] = r0 + u < nt ? A.nbr[(long)(r0 + u) * cap + i] : i;
#pragma unroll
                for (int u = 0; u < WV_NB; ++u)
                    fl[u] = r0 + u < nt ? A.h[hist_idx(am_own, aoff_own, r0 + u, i, cap, Kcap)].flags : 0;
                double4 q[WV_NB];
#pragma unroll
                for (int u = 0; u < WV_NB; ++u) q[u] = A.in.p[jj[u]];
#pragma unroll
                for (int u = 0; u < WV_NB; ++u) {
                    const V3 dv = v3(pi.x - q[u].x, pi.y - q[u].y, pi.z - q[u].z);
                    const double gap = len(dv) - pi.w - q[u].w;
                    if (r0 + u < nt &&
                        (gap < 0.0 || ((fl[u] >> 8) & 1) || gap < P.layer2 * (pi.w + q[u].w)))
                        mask |= 1ull << (r0 + u);
                }
            }
        } else {
            unsigned long long nm = A.nmask[i];
            const int n = A.nnb[i];
            heavy = n > 64;
            if (heavy) nm = 0ull;
            if (n < 64) nm &= (1ull << n) - 1ull;
            while (nm) {
                int ks[WV_NB], jj[WV_NB];
#pragma unroll
                for (int u = 0; u < WV_NB; ++u) {
                    ks[u] = nm ? __ffsll((long long)nm) - 1 : -1;
                    nm &= nm - 1ull;
                }
#pragma unroll
                for (int u = 0; u < WV_NB; ++u) jj[u] = ks[u] >= 0 ? A.nbr[(long)ks[u] * cap + i] : i;
                double4 q[WV_NB];
#pragma unroll
                for (int u = 0; u < WV_NB; ++u) q[u] = A.in.p[jj[u]];
#pragma unroll
                for (int u = 0; u < WV_NB; ++u) {
                    V3 nv;
                    double gap;
                    if (ks[u] >= 0 && pair_geometry(pi, q[u], nv, gap)) mask |= 1ull << ks[u];
                }
            }
        }
        __builtin_amdgcn_s_setprio(0);
        if constexpr (PAIR) brk = (unsigned long long)__double_as_longlong(vbi.w) & ~mask;
        const V3 ip = v3(pi.x, pi.y, pi.z);
        for (int k = 0; k < (heavy ? 0 : nw); ++k) {
            V3 nv;
            double gap;
            wall_geometry(P, A.Tw, ip, pi.w, A.wl[(long)k * cap + i], nv, gap);
            if (gap < 0.0) wmask |= 1u << k;
        }
        const double di = 2.0 * pi.w;
        grav = heavy ? 0.0 : di * di * di;
        if constexpr (WET) wt->vol[lane] = liquid_volume(P, pi.w, A.in.lqd[i]);
    }
    wave_sync_lds();  // the stash is read by other lanes in phase C
    // phase B: number the wave's entries, pairs before walls (before
    // copies) per particle
    const unsigned nforce = (unsigned)__popcll(mask) + (unsigned)__popc(wmask);
    const unsigned cnt = nforce + (PAIR ? (unsigned)__popcll(brk) : 0u);
    const unsigned incl = wave_incl_scan(cnt);
    const unsigned off = incl - cnt;
    const unsigned T = __shfl(incl, 63, 64);

    double Fx = 0.0, Fy = 0.0, Fz = 0.0, Mx = 0.0, My = 0.0, Mz = 0.0, Mr = 0.0;
    Fz = Fz - grav;  // gravity first (kernel.cu:439)
    const unsigned mlo = (unsigned)mask, mhi = (unsigned)(mask >> 32);
    const unsigned blo = (unsigned)brk, bhi = (unsigned)(brk >> 32);

    int nbridge = 0;  // WET: bridges of the owner's slots

    // phase C, software-pipelined.  Directed records: a round first issues
    // the loads of its neighbours' state, then those of the next round's
    // neighbour indices (or wall ids) and history records, then computes;
    // the history of round r+1 streams in from HBM while round r computes.
    // Pair records: one round deeper (below).
    typedef typename std::conditional<PAIR, WpEntry, WvEntry>::type Entry;
    typedef typename std::conditional<PAIR, PairHot, HistRec>::type Rec;
    Entry e;
    int j = 0, rn = 0;
    long hin = 0;
    Rec hn;
    auto entry_at = [&](unsigned t) -> Entry {
        if constexpr (PAIR)
            return wave_entry_pair(t, T, off, mlo, mhi, wmask, blo, bhi, aoff_own, wbase);
        else
            return wave_entry(t, T, off, mlo, mhi, wmask, wbase, am_own, aoff_own);
    };
    auto fetch = [&](const Entry &x) {
        if constexpr (!PAIR) {
            j = (x.wall ? A.wl : A.nbr)[(long)x.k * cap + x.ii];
            if (!x.wall) {
                hin = hist_idx(x.am, x.aoff, x.k, x.ii, cap, Kcap);
                hn = A.h[hin];
            }
        }
    };
    // pair mode: a two-deep pipeline; while round r computes, the records
    // and neighbour state of round r+1 and the indices of round r+2 load
    Entry e2;
    int j2 = 0, rn2 = 0;
    double4 qn, an, bn;
    auto fetch_idx = [&](const Entry &x, int &jj, int &rr) {
        if constexpr (PAIR) {
            if (x.kind == 1) {
                jj = A.wl[(long)x.k * cap + x.ii];
            } else {
                const int2 v = A.nl[x.nidx];
                jj = v.x;
                rr = v.y;
            }
        }
    };
    auto fetch_state = [&](const Entry &x) {
        if constexpr (PAIR) {
            if (x.kind != 1) hn = A.hq_in[rn & 0x7fffffff];
            if (x.kind == 0) {
                qn = A.in.p[j];
                an = A.in.v[2 * j];
                bn = A.in.v[2 * j + 1];
            }
        }
    };
    if constexpr (PAIR) {
        e2.valid = false;
        if (T > 0) {
            e = entry_at(lane);
            if (e.valid) fetch_idx(e, j, rn);
            if (64 < T) {
                e2 = entry_at(64 + lane);
                if (e2.valid) fetch_idx(e2, j2, rn2);
            }
            if (e.valid) fetch_state(e);
        }
    } else {
        if (T > 0) {
            e = entry_at(lane);
            if (e.valid) fetch(e);
        }
    }
    for (unsigned r0 = 0; r0 < T; r0 += 64) {
        const Entry ec = e;
        const int jc = j, rc = rn;
        const long hic = hin;
        const Rec hr = hn;
        bool is_pair;
        if constexpr (PAIR) is_pair = ec.kind == 0;
        else is_pair = !ec.wall;
        double4 qj, aj, bj;
        int jl = 0;
        (void)jl;
        double2 cold = make_double2(0.0, 0.0);
        if constexpr (PAIR) {
            qj = qn; aj = an; bj = bn;
            e = e2; j = j2; rn = rn2;
            if (e.valid) fetch_state(e);
            e2.valid = false;
            if (r0 + 128 < T) {
                e2 = entry_at(r0 + 128 + lane);
                if (e2.valid) fetch_idx(e2, j2, rn2);
            }
            if (ec.valid && is_pair && (hr.flags & PR_COLD)) cold = A.hc_in[rc & 0x7fffffff];
        } else {
            if (ec.valid && is_pair) {
                qj = A.in.p[jc];
                aj = A.in.v[2 * jc];
                bj = A.in.v[2 * jc + 1];
                if constexpr (WET) jl = A.in.lqd[jc];
            }
            if (r0 + 64 < T) {
                e = entry_at(r0 + 64 + lane);
                if (e.valid) fetch(e);
            } else {
                e.valid = false;
            }
        }
        bool copy_only = false;
        if constexpr (PAIR) copy_only = ec.kind == 2;
        if (ec.valid && copy_only) {
            if constexpr (PAIR) {
                // an owned record: no mirror bit
                A.hq_out[rc] = hr;
                if (hr.flags & PR_COLD) A.hc_out[rc] = A.hc_in[rc];
            }
        } else if (ec.valid) {
            V3 dF = v3(0.0, 0.0, 0.0), dM = v3(0.0, 0.0, 0.0), Fc = v3(0.0, 0.0, 0.0);
            double dMr = 0.0;
            int nb = 0;
            (void)Fc; (void)nb;
            const int o = ec.ii - wbase;
            const double4 qi = st->p[o], ai = st->va[o], bi = st->vb[o];
            if (is_pair) {
                if constexpr (PAIR) {
                    const bool mirror = rc < 0;
                    const double R = pair_rstar(qi.w, qj.w);
                    HistRec hc = pair_expand(hr, cold, R);
                    if (mirror) {
                        hc.dx = -hc.dx; hc.dy = -hc.dy; hc.dz = -hc.dz;
                    }
                    pair_compute<MIXED>(P, hc, qi, ai, bi, qj, aj, bj, dF, dM, dMr);
                    if (!mirror) {
                        bool need = false;
                        A.hq_out[rc] = pair_compress(hc, R, cold, need);
                        if (need) A.hc_out[rc] = cold;
                    }
                } else if constexpr (WET) {
                    HistRec hc = hr;
                    const int wr = wet_pair_compute<MIXED>(P, hc, qi, ai, bi, qj, aj, bj, wt->vol[o],
                                                           jl, dF, dM, dMr, Fc, nb);
                    if (wr == 2)
                        A.h[hic] = hc;
                    else if (wr == 1)
                        A.h[hic].flags = hc.flags;
                } else {
                    HistRec hc = hr;
                    pair_compute<MIXED>(P, hc, qi, ai, bi, qj, aj, bj, dF, dM, dMr);
                    A.h[hic] = hc;
                }
            } else {
                wall_entry<MIXED>(P, A, ec.ii, ec.k, jc, qi, ai, bi, dF, dM, dMr);
            }
            res[0][lane] = dF.x; res[1][lane] = dF.y; res[2][lane] = dF.z;
            res[3][lane] = dM.x; res[4][lane] = dM.y; res[5][lane] = dM.z;
            res[6][lane] = dMr;
            if constexpr (WET) {
                res[7][lane] = Fc.x; res[8][lane] = Fc.y; res[9][lane] = Fc.z;
                wt->nb[lane] = nb;
            }
        }
        wave_sync_lds();
        if (nforce) {
            const unsigned a = off > r0 ? off : r0;
            const unsigned b = off + nforce < r0 + 64 ? off + nforce : r0 + 64;
            for (unsigned q = a; q < b; ++q) {
                const unsigned s = q - r0;
                Fx = Fx + res[0][s]; Fy = Fy + res[1][s]; Fz = Fz + res[2][s];
                Mx = Mx + res[3][s]; My = My + res[4][s]; Mz = Mz + res[5][s];
                Mr = Mr + res[6][s];
                if constexpr (WET) {  // the bridge force after the slot's contact terms
                    Fx = Fx + res[7][s]; Fy = Fy + res[8][s]; Fz = Fz + res[9][s];
                    nbridge = nbridge + wt->nb[s];
                }
            }
        }
        wave_sync_lds();
    }
    // phase D: integration by the owner
    bool over = false;
    if (own) {
        const double4 pi = st->p[lane], va = st->va[lane], vb = st->vb[lane];
        V3 F = v3(Fx, Fy, Fz), M = v3(Mx, My, Mz);
        if constexpr (!PAIR) {
            if (heavy) {  // the loop kernel's walk over all slots, then the walls
                const V3 iv = v3(va.x, va.y, va.z), iw = v3(va.w, vb.x, vb.y);
                F = v3(0.0, 0.0, 0.0);
                M = v3(0.0, 0.0, 0.0);
                Mr = 0.0;
                nbridge = 0;
                pp_particle<WET, MIXED>(P, A, i, pi, iv, iw, F, M, Mr, nbridge);
                pw_particle<MIXED>(P, A, i, pi, iv, iw, F, M, Mr);
            }
        }
        if constexpr (WET) A.out.lqd[i] = nbridge;
        if (A.store_forces) {
            A.Fx[i] = F.x; A.Fy[i] = F.y; A.Fz[i] = F.z;
            A.Mx[i] = M.x; A.My[i] = M.y; A.Mz[i] = M.z;
        }
        const double tag = PAIR ? __longlong_as_double((long long)(mask & ownm)) : 0.0;
        over = integrate(P, A, i, pi, va, vb, F, M, Mr, tag);
    }
    request_rebuild(A, over);
}

// One wave per block: a wave slot that frees up takes the next 64
// particles at once, instead of waiting until four slots of one CU are
// free.  Each XCD walks a contiguous range of chunks of 64 particles
// (A.xb), so neighbouring particles share its L2; the ranges are cut for
// equal work (k_chunk_work) where the pair records are in use, else for
// equal particle counts.
// The mixed-precision form is held to 3 waves per SIMD (it would take 171
// registers and fit only 2); the FP64 form fits 3 by itself and is
// scheduled better without the hint.
template <bool MIXED, bool PAIR, bool WET = false>
DEV void step_wave_one(const DevInst &P, const StepArgs &A) {
    // a rebuild is due before this step when stop <= A.step (tested in
    // step_wave_block once the first loads are issued)
    const int stop = __hip_atomic_load(A.stop_at, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // blocks are dispatched round-robin over the 8 XCDs: block b runs on
    // XCD b % 8 and takes the (b / 8)-th chunk of that XCD's range.  Odd
    // steps walk each range backwards, so a launch starts where the one
    // before ended, on state and records still in L2 (the order changes
    // nothing else: each particle's step reads only the previous step).
    // The last L chunks an XCD takes are split into two half-waves of 32
    // particles each, and the last Qn of them into four quarter-waves of
    // 16, so the launch ends on short waves.  With A.wsplit > 1 (a packing
    // too small to give every wave slot of the chip several chunks) every
    // chunk is split into wsplit sub-waves of 64 / wsplit particles instead.
    const unsigned x = blockIdx.x % 8u, k = blockIdx.x / 8u;
    const unsigned n = A.xb[x + 1] - A.xb[x], L = n / 4u < WV_TAIL ? n / 4u : WV_TAIL;
    const unsigned Qn = L / 6u < WV_QTAIL ? L / 6u : WV_QTAIL;
    unsigned pos = k;
    int lo = 0, hi = 64;
    if (A.wsplit > 1) {
        const unsigned parts = (unsigned)A.wsplit, q = k % parts;
        pos = k / parts;
        if (pos >= n) return;
        lo = (int)(q * (64u / parts));
        hi = lo + (int)(64u / parts);
    } else if (k >= n - L) {
        const unsigned j = k - (n - L);
        if (j < 2u * (L - Qn)) {
            pos = n - L + j / 2u;
            lo = (int)(j & 1u) * 32;
            hi = lo + 32;
        } else {
            const unsigned q = j - 2u * (L - Qn);
            if (q >= 4u * Qn) return;
            pos = n - Qn + q / 4u;
            lo = (int)(q & 3u) * 16;
            hi = lo + 16;
        }
    }
    const unsigned c = (A.step & 1) ? A.xb[x + 1] - 1u - pos : A.xb[x] + pos;
    __shared__ double res[WET ? 10 : 7][64];
    __shared__ WvStash stash;
    if constexpr (WET) {
        __shared__ WvWet wet;
        step_wave_block<MIXED, PAIR, 64, true>(P, A, c, res, &stash, &wet, lo, hi, stop);
    } else {
        step_wave_block<MIXED, PAIR, 64>(P, A, c, res, &stash, nullptr, lo, hi, stop);
    }
}

template <bool PAIR>
__global__ void __launch_bounds__(64) k_step_wave(DevInst P, StepArgs A) {
    step_wave_one<false, PAIR>(P, A);
}

template <bool PAIR>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3)))
k_step_wave_mixed(DevInst P, StepArgs A) {
    step_wave_one<true, PAIR>(P, A);
}

// wet packings (capillary bridges; directed records)
__global__ void __launch_bounds__(64) k_step_wave_wet(DevInst P, StepArgs A) {
    step_wave_one<false, false, true>(P, A);
}

__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3)))
k_step_wave_wet_mixed(DevInst P, StepArgs A) {
    step_wave_one<true, false, true>(P, A);
}

void launch_step_wave(const DevInst &P, const StepArgs &A, bool mixed, bool pair, bool wet,
                      hipStream_t s) {
    if (A.total <= 0) return;
    unsigned most = 0;
    for (int x = 0; x < 8; ++x) {
        const unsigned n = A.xb[x + 1] - A.xb[x], L = n / 4u < WV_TAIL ? n / 4u : WV_TAIL;
        const unsigned Qn = L / 6u < WV_QTAIL ? L / 6u : WV_QTAIL;
        const unsigned want = A.wsplit > 1 ? n * (unsigned)A.wsplit : n + L + 2u * Qn;
        most = want > most ? want : most;
    }
    const dim3 g(8 * most), b(64);
    if (wet && mixed) hipLaunchKernelGGL(k_step_wave_wet_mixed, g, b, 0, s, P, A);
    else if (wet) hipLaunchKernelGGL(k_step_wave_wet, g, b, 0, s, P, A);
    else if (mixed && pair) hipLaunchKernelGGL((k_step_wave_mixed<true>), g, b, 0, s, P, A);
    else if (mixed) hipLaunchKernelGGL((k_step_wave_mixed<false>), g, b, 0, s, P, A);
    else if (pair) hipLaunchKernelGGL((k_step_wave<true>), g, b, 0, s, P, A);
    else hipLaunchKernelGGL((k_step_wave<false>), g, b, 0, s, P, A);
}

// ---------------------------------------------------------------------
// Batched step over many instances in one launch.  Block t of the grid
// (after the XCD remap) takes task t = (instance, block of that instance);
// tasks list the instances largest first, so an instance's blocks are
// contiguous and run on one XCD.  The instance's constants (DevInst) and
// lists come from per-instance tables in device memory instead of kernel
// arguments; each instance keeps its own rebuild flag and step limit, so
// every instance follows exactly the steps it would take on its own.  The
// instance's rebuild request comes back in stop, for the step block to test
// once its first loads are in flight.
// ---------------------------------------------------------------------
DEV bool batch_step_args(const BatchInst &B, int step, StepArgs &A, int &stop) {
    if (step >= B.limit) return false;
    A = B.A;
    stop = __hip_atomic_load(A.stop_at, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int b = (B.cur + step) & 1, l = (B.Lidx + step) & 1;
    A.in.p = B.P4[b]; A.in.v = B.VW[b]; A.in.lqd = B.lq[l];
    A.out.p = B.P4[1 - b]; A.out.v = B.VW[1 - b]; A.out.lqd = B.lq[1 - l];
    A.step = step;
    double tw = B.Tw0;
    if (B.dTw != 0.0)  // a fixed lid (topv = 0) leaves Tw as it is
        for (int s = 0; s <= step; ++s) tw = tw + B.dTw;
    A.Tw = tw;
    return true;
}

// one block of WVB_BS particles of a batched instance, dry or wet: one wave
// per block, so a wave slot that frees up takes the next chunk at once (as
// in the single-instance launch)
template <bool MIXED>
DEV void step_wave_batched(const DevInst &P, const StepArgs &A, unsigned blk, int stop) {
    __shared__ double res_all[WVB_BS / 64][10][64];
    __shared__ WvStash stash[WVB_BS / 64];
    __shared__ WvWet wet[WVB_BS / 64];
    const int w = threadIdx.x >> 6;
    if (P.forceCap == 1)
        step_wave_block<MIXED, false, WVB_BS, true>(P, A, blk, res_all[w], &stash[w], &wet[w], 0, 64,
                                                    stop);
    else
        step_wave_block<MIXED, false, WVB_BS>(P, A, blk, res_all[w], &stash[w], nullptr, 0, 64, stop);
}

template <bool MIXED>
__global__ void __launch_bounds__(WVB_BS) k_step_batch_wave(const DevInst *insts, const BatchInst *tab,
                                                           const int2 *tasks, int step) {
    const int2 tk = tasks[xcd_block_dir(gridDim.x, step & 1)];
    StepArgs A;
    int stop;
    if (!batch_step_args(tab[tk.x], step, A, stop)) return;
    step_wave_batched<MIXED>(insts[tk.x], A, (unsigned)tk.y, stop);
}

template <bool MIXED>
__global__ void __launch_bounds__(256) k_step_batch_loop(const DevInst *insts, const BatchInst *tab,
                                                         const int2 *tasks, int step) {
    const int2 tk = tasks[xcd_block(gridDim.x)];
    StepArgs A;
    int stop;
    if (!batch_step_args(tab[tk.x], step, A, stop)) return;
    const DevInst &P = insts[tk.x];
    if (P.forceCap == 1)
        step_loop_block<true, MIXED>(P, A, (unsigned)tk.y, stop);
    else
        step_loop_block<false, MIXED>(P, A, (unsigned)tk.y, stop);
}

void launch_step_batch(const DevInst *insts, const BatchInst *tab, const int2 *tasks, int ntasks,
                       bool wave, bool mixed, int step, hipStream_t s) {
    if (ntasks <= 0) return;
    if (wave) {
        if (mixed)
            hipLaunchKernelGGL((k_step_batch_wave<true>), dim3(ntasks), dim3(WVB_BS), 0, s, insts, tab,
                               tasks, step);
        else
            hipLaunchKernelGGL((k_step_batch_wave<false>), dim3(ntasks), dim3(WVB_BS), 0, s, insts, tab,
                               tasks, step);
    } else {
        if (mixed)
            hipLaunchKernelGGL((k_step_batch_loop<true>), dim3(ntasks), dim3(256), 0, s, insts, tab,
                               tasks, step);
        else
            hipLaunchKernelGGL((k_step_batch_loop<false>), dim3(ntasks), dim3(256), 0, s, insts, tab,
                               tasks, step);
    }
}

// ---------------------------------------------------------------------
// Device-driven batched stepping.  Global step g of the run is local step
// s = g - base of an instance's current segment; the instance skips it once
// it has taken its steps or asked for a rebuild at an earlier step (stop_at
// = local step + 1), exactly as a chunk of single-instance steps.  Every few
// steps a rebuild slot runs: kb_arm fills the batched rebuild table for
// the instances that asked (zero particles and cells for the others, so the
// stage kernels skip them), the stage kernels run over task lists of every
// instance, and kb_finish records overflow.
// ---------------------------------------------------------------------
DEV bool dbatch_step_args(const DBatchInst &D, const DBatchCtl &C, int g, StepArgs &A, int &stop) {
    const int s = g - C.base;
    if (C.failed || s < 0 || s >= C.left) return false;
    stop = __hip_atomic_load(D.stop_at, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int b = (C.cur + s) & 1, l = (C.Lidx + s) & 1;
    StepArgs a = {};
    a.in.p = D.P4[b]; a.in.v = D.VW[b]; a.in.lqd = D.lq[l];
    a.out.p = D.P4[1 - b]; a.out.v = D.VW[1 - b]; a.out.lqd = D.lq[1 - l];
    a.nbr = D.nbr[C.hcur]; a.nnb = D.nnb[C.hcur];
    a.h = D.hp[C.hl]; a.amask = D.amask[C.hl]; a.actoff = D.actoff[C.hl];
    a.nmask = D.nmask;
    a.wl = D.wl[C.hcur]; a.nwl = D.nwl[C.hcur]; a.hw = D.hw[C.hcur];
    a.stop_at = D.stop_at;
    a.cap = D.cap; a.K = D.K; a.total = D.total;
    a.step = s;
    double tw = C.Tw;
    if (D.dTw != 0.0)  // a fixed lid (topv = 0) leaves Tw as it is
        for (int q = 0; q <= s; ++q) tw = tw + D.dTw;
    a.Tw = tw;
    A = a;
    return true;
}

template <bool MIXED>
__global__ void __launch_bounds__(WVB_BS) k_step_dbatch_wave(const DevInst *insts, const DBatchInst *dins,
                                                            const DBatchCtl *ctl, const int2 *tasks,
                                                            int g) {
    const int2 tk = tasks[xcd_block_dir(gridDim.x, g & 1)];
    StepArgs A;
    int stop;
    if (!dbatch_step_args(dins[tk.x], ctl[tk.x], g, A, stop)) return;
    step_wave_batched<MIXED>(insts[tk.x], A, (unsigned)tk.y, stop);
}

template <bool MIXED>
__global__ void __launch_bounds__(256) k_step_dbatch_loop(const DevInst *insts, const DBatchInst *dins,
                                                          const DBatchCtl *ctl, const int2 *tasks,
                                                          int g) {
    const int2 tk = tasks[xcd_block(gridDim.x)];
    StepArgs A;
    int stop;
    if (!dbatch_step_args(dins[tk.x], ctl[tk.x], g, A, stop)) return;
    const DevInst &P = insts[tk.x];
    if (P.forceCap == 1)
        step_loop_block<true, MIXED>(P, A, (unsigned)tk.y, stop);
    else
        step_loop_block<false, MIXED>(P, A, (unsigned)tk.y, stop);
}

void launch_step_dbatch(const DevInst *insts, const DBatchInst *dins, const DBatchCtl *ctl,
                        const int2 *tasks, int ntasks, bool wave, bool mixed, int g, hipStream_t s) {
    if (ntasks <= 0) return;
    if (wave) {
        if (mixed)
            hipLaunchKernelGGL((k_step_dbatch_wave<true>), dim3(ntasks), dim3(WVB_BS), 0, s, insts, dins,
                               ctl, tasks, g);
        else
            hipLaunchKernelGGL((k_step_dbatch_wave<false>), dim3(ntasks), dim3(WVB_BS), 0, s, insts,
                               dins, ctl, tasks, g);
    } else {
        if (mixed)
            hipLaunchKernelGGL((k_step_dbatch_loop<true>), dim3(ntasks), dim3(256), 0, s, insts, dins,
                               ctl, tasks, g);
        else
            hipLaunchKernelGGL((k_step_dbatch_loop<false>), dim3(ntasks), dim3(256), 0, s, insts, dins,
                               ctl, tasks, g);
    }
}

// one thread per instance, after global steps up to g_end: the rebuild
// table entry of each instance that asked for a rebuild and has steps left
// (the buffer flips of a rebuild as dem_do_rebuild makes them), and in
// nxt the control record it resumes with at global step resume.  The steps
// of the next R global steps run concurrently on the other stream and read
// ctl, so ctl is left alone here: kb_commit installs nxt once the slot is
// done and those steps have run.  An armed instance idles until then (its
// stop_at stays at or below the steps it has seen).
__global__ void kb_arm(const DBatchInst *dins, const DBatchCtl *ctl, DBatchCtl *nxt, RbInst *tab,
                       int n, int g_end, int resume) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const DBatchInst &D = dins[k];
    DBatchCtl C = ctl[k];
    RbInst R;
    memset(&R, 0, sizeof(R));
    R.inst = k;
    R.overflow = D.overflow;
    R.keys = D.keys; R.vals = D.vals; R.keys2 = D.keys2; R.vals2 = D.vals2;
    R.cap = D.cap;
    R.K = D.K;
    const int seen = g_end + 1 - C.base;
    // steps running now may lower stop_at, never to seen or below
    const int stop = __hip_atomic_load(D.stop_at, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int done = stop < seen ? stop : seen;
    const int executed = done < C.left ? done : C.left;
    const bool need = !C.failed && stop <= seen && executed < C.left;
    if (!need) {  // R.total = R.ncells = 0: every stage skips this entry
        tab[k] = R;
        nxt[k].armed = 0;
        return;
    }
    const int cur = (C.cur + executed) & 1, lx = (C.Lidx + executed) & 1;
    const int hc = 1 - C.hcur, hl = 1 - C.hl;  // the lists become the old ones
    R.pa = D.P4[cur]; R.va = D.VW[cur]; R.lqa = D.lq[lx];
    R.pb = D.P4[1 - cur]; R.vb = D.VW[1 - cur]; R.lqb = D.lq[1 - lx];
    R.lvla = D.level[C.pcur]; R.ida = D.id[C.pcur];
    R.lvlb = D.level[1 - C.pcur]; R.idb = D.id[1 - C.pcur];
    R.passes = D.passes;
    R.cstart = D.cstart; R.cend = D.cend; R.ncells = D.ncells;
    R.nbr = D.nbr[hc]; R.nnb = D.nnb[hc]; R.nbr_old = D.nbr[1 - hc]; R.nnb_old = D.nnb[1 - hc];
    R.hp = D.hp[hl]; R.hp_old = D.hp[1 - hl];
    R.amask = D.amask[hl]; R.amask_old = D.amask[1 - hl];
    R.actoff = D.actoff[hl]; R.actoff_old = D.actoff[1 - hl];
    R.nmask = D.nmask;
    R.wl = D.wl[hc]; R.nwl = D.nwl[hc]; R.wl_old = D.wl[1 - hc]; R.nwl_old = D.nwl[1 - hc];
    R.hw = D.hw[hc]; R.hw_old = D.hw[1 - hc];
    R.total = D.total;
    R.old_total = D.total;
    double tw = C.Tw;
    if (D.dTw != 0.0)
        for (int q = 0; q < executed; ++q) tw = tw + D.dTw;
    R.Tw = tw;
    tab[k] = R;
    C.cur = 1 - cur;
    C.Lidx = 1 - lx;
    C.pcur = 1 - C.pcur;
    C.hcur = hc;
    C.hl = hl;
    C.base = resume;
    C.left -= executed;
    C.executed += executed;
    C.seg_steps = executed;
    C.Tw = tw;
    C.nreb += 1;
    C.armed = 1;
    nxt[k] = C;
}

// after the slot's stages: an instance whose lists overflowed halts
__global__ void kb_finish(const DBatchInst *dins, DBatchCtl *nxt, int n) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n || !nxt[k].armed) return;
    const int ov = *dins[k].overflow;
    if (ov > 0) nxt[k].failed = ov;
}

// on the step stream, after the slot and the steps that ran beside it:
// the rebuilt instances take their new control records and clear their
// rebuild requests
__global__ void kb_commit(const DBatchInst *dins, DBatchCtl *ctl, const DBatchCtl *nxt, int n) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n || !nxt[k].armed) return;
    DBatchCtl C = nxt[k];
    C.armed = 0;
    ctl[k] = C;
    *dins[k].stop_at = INT_MAX;
}

void launch_rebuild_slot(const DevInst *insts, const DBatchInst *dins, const DBatchCtl *ctl,
                         DBatchCtl *nxt, RbInst *tab, int n, int g_end, int resume, const int2 *ptasks,
                         int nptasks, const int2 *ctasks, int nctasks, hipStream_t s) {
    hipLaunchKernelGGL(kb_arm, dim3((n + 63) / 64), dim3(64), 0, s, dins, ctl, nxt, tab, n, g_end,
                       resume);
    launch_rebuild_batch(insts, tab, n, ptasks, nptasks, ctasks, nctasks, s);
    hipLaunchKernelGGL(kb_finish, dim3((n + 63) / 64), dim3(64), 0, s, dins, nxt, n);
}

void launch_commit_slot(const DBatchInst *dins, DBatchCtl *ctl, const DBatchCtl *nxt, int n,
                        hipStream_t s) {
    hipLaunchKernelGGL(kb_commit, dim3((n + 63) / 64), dim3(64), 0, s, dins, ctl, nxt, n);
}

// ---------------------------------------------------------------------
// Packing analysis: solid fraction phi of the core cylinder and mean
// contact number Z, one block per instance, deterministic reductions.
// ---------------------------------------------------------------------
// positive nodes and weights of 32-point Gauss-Legendre on [-1, 1]
__constant__ double c_gl_x[16] = {0.04830766568773831, 0.14447196158279649, 0.23928736225213706, 0.33186860228212767, 0.42135127613063533, 0.50689990893222936, 0.5877157572407623, 0.66304426693021523, 0.73218211874028971, 0.79448379596794239, 0.84936761373256997, 0.8963211557660522, 0.93490607593773967, 0.96476225558750639, 0.98561151154526838, 0.99726386184948157};
__constant__ double c_gl_w[16] = {0.096540088514727812, 0.095638720079274833, 0.093844399080804566, 0.091173878695763863, 0.087652093004403908, 0.083311924226946846, 0.078193895787070311, 0.072345794108848449, 0.065822222776361752, 0.058684093478535704, 0.050998059262376244, 0.042835898022226426, 0.034273862913021626, 0.025392065309262427, 0.016274394730905965, 0.0070186100094692984};

// area of the intersection of a disc of radius s (centre at distance d from
// the axis) with the disc of radius Rc around the axis
DEV double lens_area(double s, double Rc, double d) {
    if (s <= 0.0) return 0.0;
    if (d >= s + Rc) return 0.0;
    if (d <= fabs(Rc - s)) {
        const double m = fmin(s, Rc);
        return M_PI * m * m;
    }
    const double c1 = fmin(1.0, fmax(-1.0, (d * d + s * s - Rc * Rc) / (2.0 * d * s)));
    const double c2 = fmin(1.0, fmax(-1.0, (d * d + Rc * Rc - s * s) / (2.0 * d * Rc)));
    const double k = (-d + s + Rc) * (d + s - Rc) * (d - s + Rc) * (d + s + Rc);
    return s * s * acos(c1) + Rc * Rc * acos(c2) - 0.5 * sqrt(fmax(0.0, k));
}

// volume of the sphere (centre at axial distance rho, height z, radius r)
// inside {axial distance <= Rc, zlo <= z <= zhi}
DEV double core_volume(double rho, double z, double r, double Rc, double zlo, double zhi) {
    const double a = fmax(z - r, zlo), b = fmin(z + r, zhi);
    if (!(a < b) || rho - r >= Rc) return 0.0;
    const double u1 = a - z, u2 = b - z;
    if (rho + r <= Rc)
        return M_PI * ((r * r * u2 - u2 * u2 * u2 / 3.0) - (r * r * u1 - u1 * u1 * u1 / 3.0));
    const double h = 0.5 * (u2 - u1), m = 0.5 * (u2 + u1);
    double v = 0.0;
    for (int q = 0; q < 16; ++q) {
        const double up = m + h * c_gl_x[q], um = m - h * c_gl_x[q];
        const double sp = sqrt(fmax(0.0, r * r - up * up)), sm = sqrt(fmax(0.0, r * r - um * um));
        v += c_gl_w[q] * (lens_area(sp, Rc, rho) + lens_area(sm, Rc, rho));
    }
    return h * v;
}

DEV double block_sum_f64(double v, double *sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] = sh[threadIdx.x] + sh[threadIdx.x + s];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

DEV double block_max_f64(double v, double *sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

__global__ void __launch_bounds__(256) k_analyze(const DevInst *insts, const AnaArgs *args,
                                                 double *out /* [n][3] */) {
    __shared__ double sh[256];
    const DevInst &P = insts[blockIdx.x];
    const AnaArgs &G = args[blockIdx.x];
    const long cap = G.cap;
    double contacts = 0.0, top = -1.0e300;
    for (int i = threadIdx.x; i < G.total; i += blockDim.x) {
        const double4 pi = G.p[i];
        const int n = G.nnb[i], nw = G.nwl[i];
        int c = 0;
        for (int k = 0; k < n; ++k) {
            V3 nv;
            double gap;
            if (pair_geometry(pi, G.p[G.nbr[(long)k * cap + i]], nv, gap)) ++c;
        }
        bool touch = c > 0;
        for (int k = 0; k < nw; ++k) {
            V3 nv;
            double gap;
            wall_geometry(P, G.Tw, v3(pi.x, pi.y, pi.z), pi.w, G.wl[(long)k * cap + i], nv, gap);
            touch = touch || gap < 0.0;
        }
        contacts += (double)c;
        if (touch) top = fmax(top, pi.z + pi.w);
    }
    const double C = block_sum_f64(contacts, sh);
    const double T = block_max_f64(top, sh);
    const double Rc = P.R - 2.0, zlo = P.Bw + 2.0, zhi = T - 2.0;
    double vol = 0.0;
    if (zhi > zlo) {
        for (int i = threadIdx.x; i < G.total; i += blockDim.x) {
            const double4 pi = G.p[i];
            const double dx = pi.x - P.cirX, dy = pi.y - P.cirY;
            vol += core_volume(sqrt(dx * dx + dy * dy), pi.z, pi.w, Rc, zlo, zhi);
        }
    }
    const double V = block_sum_f64(vol, sh);
    if (threadIdx.x == 0) {
        const double core = (zhi > zlo && Rc > 0.0) ? M_PI * Rc * Rc * (zhi - zlo) : 0.0;
        out[3 * blockIdx.x + 0] = core > 0.0 ? V / core : 0.0;
        out[3 * blockIdx.x + 1] = G.total > 0 ? C / (double)G.total : 0.0;
        out[3 * blockIdx.x + 2] = T;
    }
}

void launch_analyze(const DevInst *insts, const AnaArgs *args, double *out, int n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_analyze, dim3(n), dim3(256), 0, s, insts, args, out);
}

// ---------------------------------------------------------------------
// Batched rebuild: every stage of the rebuild (hash, sort, permutation,
// cell ranges, neighbour and wall lists, history remap) runs as one launch
// over all instances due for a rebuild.  Particle stages take tasks
// (table entry, block of 256); the sort gives each instance one block.
// Each stage calls the device function of the single-instance kernel, so
// the lists and histories are identical to a separate rebuild.
// ---------------------------------------------------------------------
#define RB_BS 256
#define BSORT_BS 1024
// The particle and cell stages run in one-wave blocks, KB_PER per task of
// RB_BS: beside the step kernel a block of one wave starts wherever one
// wave fits.
#define KB_WV 64
#define KB_PER (RB_BS / KB_WV)

DEV int kb_item(const int2 tk) {
    return tk.y * RB_BS + (int)(blockIdx.x % KB_PER) * KB_WV + (int)threadIdx.x;
}

__global__ void __launch_bounds__(KB_WV) kb_calc_hash(const DevInst *insts, const RbInst *tab,
                                                      const int2 *tasks) {
    const int2 tk = tasks[blockIdx.x / KB_PER];
    const RbInst &R = tab[tk.x];
    const int i = kb_item(tk);
    if (i < R.total) calc_hash_one(insts[R.inst], R.pa, R.lvla, R.ida, R.keys, R.vals, i);
}

// stable LSD radix sort of one instance's (key, value) pairs by one block:
// per pass a digit histogram, its prefix sum, and a tile-by-tile stable
// scatter (ranks within a tile from wave ballots, as k_rs_scatter)
__global__ void __launch_bounds__(BSORT_BS) kb_sort(const RbInst *tab) {
    const RbInst &R = tab[blockIdx.x];
    const int n = R.total, tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    __shared__ unsigned base[256];
    __shared__ unsigned wcnt[BSORT_BS / 64][256];
    unsigned *ka = R.keys, *va = R.vals, *kb = R.keys2, *vb = R.vals2;
    const unsigned long long lt = (1ull << lane) - 1ull;
    if (n <= 1) return;
    for (int pass = 0; pass < R.passes; ++pass) {
        const int shift = 8 * pass;
        if (tid < 256) base[tid] = 0;
        __syncthreads();
        for (int e = tid; e < n; e += BSORT_BS) atomicAdd(&base[(ka[e] >> shift) & 255u], 1u);
        __syncthreads();
        if (tid < 64) {
            const unsigned c0 = base[4 * tid], c1 = base[4 * tid + 1], c2 = base[4 * tid + 2],
                           c3 = base[4 * tid + 3];
            const unsigned sum = c0 + c1 + c2 + c3;
            const unsigned ex = wave_incl_scan(sum) - sum;
            base[4 * tid] = ex;
            base[4 * tid + 1] = ex + c0;
            base[4 * tid + 2] = ex + c0 + c1;
            base[4 * tid + 3] = ex + c0 + c1 + c2;
        }
        __syncthreads();
        for (int t0 = 0; t0 < n; t0 += BSORT_BS) {
            for (int q = tid; q < (BSORT_BS / 64) * 256; q += BSORT_BS) (&wcnt[0][0])[q] = 0;
            __syncthreads();
            const int e = t0 + tid;
            const bool valid = e < n;
            unsigned key = 0, val = 0, d = 256u;
            if (valid) {
                key = ka[e];
                val = va[e];
                d = (key >> shift) & 255u;
            }
            const unsigned long long peers = wave_match9(d);
            const unsigned rank = (unsigned)__popcll(peers & lt);
            if (valid && (peers & lt) == 0ull) wcnt[w][d] = (unsigned)__popcll(peers);
            __syncthreads();
            if (valid) {
                unsigned pos = base[d] + rank;
                for (int q = 0; q < w; ++q) pos += wcnt[q][d];
                kb[pos] = key;
                vb[pos] = val;
            }
            __syncthreads();
            if (tid < 256) {
                unsigned add = 0;
                for (int q = 0; q < BSORT_BS / 64; ++q) add += wcnt[q][tid];
                base[tid] += add;
            }
            __syncthreads();
        }
        unsigned *t = ka; ka = kb; kb = t;
        t = va; va = vb; vb = t;
    }
    if (ka != R.keys) {
        for (int e = tid; e < n; e += BSORT_BS) {
            R.keys[e] = ka[e];
            R.vals[e] = va[e];
        }
    }
}

__global__ void __launch_bounds__(KB_WV) kb_permute(const RbInst *tab, const int2 *tasks) {
    const int2 tk = tasks[blockIdx.x / KB_PER];
    const RbInst &R = tab[tk.x];
    const int i = kb_item(tk);
    if (i >= R.total) return;
    StateBuf a, b;
    a.p = (double4 *)R.pa; a.v = (double4 *)R.va; a.lqd = (int *)R.lqa;
    b.p = R.pb; b.v = R.vb; b.lqd = R.lqb;
    permute_one(R.vals, a, b, R.lvla, R.lvlb, R.ida, R.idb, i);
}

// empty cells and a cleared overflow flag; tasks over blocks of cells
__global__ void __launch_bounds__(KB_WV) kb_fill_cells(const RbInst *tab, const int2 *tasks) {
    const int2 tk = tasks[blockIdx.x / KB_PER];
    const RbInst &R = tab[tk.x];
    const int c = kb_item(tk);
    if (c < R.ncells) {
        R.cstart[c] = DEM_EMPTY;
        R.cend[c] = DEM_EMPTY;
    }
    if (c == 0 && R.total > 0) *R.overflow = 0;
}

__global__ void __launch_bounds__(KB_WV) kb_cell_range(const DevInst *insts, const RbInst *tab,
                                                       const int2 *tasks) {
    const int2 tk = tasks[blockIdx.x / KB_PER];
    const RbInst &R = tab[tk.x];
    const int k = kb_item(tk);
    if (k < R.total) cell_range_one(R.keys, R.pb, R.lvlb, insts[R.inst], R.cstart, R.cend, R.total, k);
}

// one-wave blocks, 16 lanes per particle: 64 blocks per task of 256
// particles.  The slots run beside the step kernel, whose waves hold most of
// every SIMD's registers; a block of one wave starts wherever one wave fits,
// where a block of 1024 lanes would wait for four free wave slots on every
// SIMD of one CU.
#define KBL_BS 64
#define KBL_PER (RB_BS * 16 / KBL_BS)
__global__ void __launch_bounds__(KBL_BS) kb_lists(const DevInst *insts, const RbInst *tab,
                                                   const int2 *tasks) {
    const int2 tk = tasks[blockIdx.x / KBL_PER];
    const RbInst &R = tab[tk.x];
    const int first = tk.y * RB_BS + (int)(blockIdx.x % KBL_PER) * (KBL_BS / 16);
    if (first >= R.total) return;  // the whole block: no lane is needed for a shuffle
    const int i = first + (int)(threadIdx.x >> 4), g = threadIdx.x & 15;
    const bool valid = i < R.total;
    const DevInst &P = insts[R.inst];
    neighbor_find_group16(R.pb, R.lvlb, P, R.cstart, R.cend, R.nbr, R.nnb, R.overflow, R.cap, i, g,
                          valid);
    if (valid && g == 1) wall_list_one(R.pb, P, R.Tw, R.wl, R.nwl, R.cap, i);
}

__global__ void __launch_bounds__(KB_WV) kb_active(const DevInst *insts, const RbInst *tab,
                                                  const int2 *tasks) {
    const int2 tk = tasks[blockIdx.x / KB_PER];
    const RbInst &R = tab[tk.x];
    const int i = kb_item(tk);
    if (i < R.total)
        active_slots_one(R.pb, R.nbr, R.nnb, R.amask, R.keys2, R.nmask,
                         near_gap_of(insts[R.inst]), R.cap, i);
}

// actoff = exclusive prefix sum of the per-particle counts, one block per
// instance, tiles of SCAN_ITEMS consecutive counts per lane
__global__ void __launch_bounds__(SCAN_BS) kb_actoff(const RbInst *tab) {
    const RbInst &R = tab[blockIdx.x];
    __shared__ unsigned lds[4];
    unsigned carry = 0;
    for (int t0 = 0; t0 < R.total; t0 += SCAN_TILE) {
        const int e = t0 + threadIdx.x * SCAN_ITEMS;
        unsigned v[SCAN_ITEMS], sum = 0;
#pragma unroll
        for (int q = 0; q < SCAN_ITEMS; ++q) {
            v[q] = e + q < R.total ? R.keys2[e + q] : 0u;
            sum += v[q];
        }
        unsigned tot;
        unsigned ex = carry + block_excl_scan(sum, &tot, lds);
#pragma unroll
        for (int q = 0; q < SCAN_ITEMS; ++q) {
            if (e + q < R.total) R.actoff[e + q] = (int)ex;
            ex += v[q];
        }
        carry += tot;
    }
}

__global__ void __launch_bounds__(KB_WV) kb_hist_remap(const RbInst *tab, const int2 *tasks) {
    const int2 tk = tasks[blockIdx.x / KB_PER];
    const RbInst &R = tab[tk.x];
    const int i = kb_item(tk);
    if (i >= R.total) return;
    hist_remap_one(R.vals, R.idb, R.ida, R.nbr, R.nnb, R.amask, R.actoff, R.nbr_old, R.nnb_old,
                   R.amask_old, R.actoff_old, R.hp, R.hp_old, R.cap, (long)R.K * R.cap, R.old_total, i);
    hist_remap_pw_one(R.vals, R.wl, R.nwl, R.wl_old, R.nwl_old, R.hw, R.hw_old, R.cap, R.old_total, i);
}

void launch_rebuild_batch(const DevInst *insts, const RbInst *tab, int ninst, const int2 *ptasks,
                          int nptasks, const int2 *ctasks, int nctasks, hipStream_t s) {
    if (ninst <= 0) return;
    const dim3 pg(KB_PER * nptasks), cg(KB_PER * nctasks), wv(KB_WV);
    if (nptasks > 0) hipLaunchKernelGGL(kb_calc_hash, pg, wv, 0, s, insts, tab, ptasks);
    hipLaunchKernelGGL(kb_sort, dim3(ninst), dim3(BSORT_BS), 0, s, tab);
    if (nptasks > 0) hipLaunchKernelGGL(kb_permute, pg, wv, 0, s, tab, ptasks);
    if (nctasks > 0) hipLaunchKernelGGL(kb_fill_cells, cg, wv, 0, s, tab, ctasks);
    if (nptasks > 0) {
        hipLaunchKernelGGL(kb_cell_range, pg, wv, 0, s, insts, tab, ptasks);
        hipLaunchKernelGGL(kb_lists, dim3(KBL_PER * nptasks), dim3(KBL_BS), 0, s, insts, tab, ptasks);
        hipLaunchKernelGGL(kb_active, pg, wv, 0, s, insts, tab, ptasks);
        hipLaunchKernelGGL(kb_actoff, dim3(ninst), dim3(SCAN_BS), 0, s, tab);
        hipLaunchKernelGGL(kb_hist_remap, pg, wv, 0, s, tab, ptasks);
    }
}

// --- the reference's separate kernels (PpInterCalc, CirPWCalc,
// UpdateSystem), accumulating into the force arrays -----------------------
template <bool WET, bool MIXED>
__global__ void __launch_bounds__(256) k_pp_force(DevInst P, StepArgs A) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.total) return;
    const double4 pi = A.in.p[i];
    const V3 iv = vel_of(A.in.v, i), iw = spin_of(A.in.v, i);
    V3 F = v3(A.Fx[i], A.Fy[i], A.Fz[i]), M = v3(A.Mx[i], A.My[i], A.Mz[i]);
    double Mr = A.Mr[i];
    int nbridge = 0;
    pp_particle<WET, MIXED>(P, A, i, pi, iv, iw, F, M, Mr, nbridge);
    if (WET) A.out.lqd[i] = A.out.lqd[i] + nbridge;
    A.Fx[i] = F.x; A.Fy[i] = F.y; A.Fz[i] = F.z;
    A.Mx[i] = M.x; A.My[i] = M.y; A.Mz[i] = M.z;
    A.Mr[i] = Mr;
}

template <bool MIXED>
__global__ void __launch_bounds__(256) k_pw_force(DevInst P, StepArgs A) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.total) return;
    const double4 pi = A.in.p[i];
    const V3 iv = vel_of(A.in.v, i), iw = spin_of(A.in.v, i);
    V3 F = v3(A.Fx[i], A.Fy[i], A.Fz[i]), M = v3(A.Mx[i], A.My[i], A.Mz[i]);
    double Mr = A.Mr[i];
    pw_particle<MIXED>(P, A, i, pi, iv, iw, F, M, Mr);
    A.Fx[i] = F.x; A.Fy[i] = F.y; A.Fz[i] = F.z;
    A.Mx[i] = M.x; A.My[i] = M.y; A.Mz[i] = M.z;
    A.Mr[i] = Mr;
}

__global__ void __launch_bounds__(256) k_update(DevInst P, StepArgs A, int *count) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool over = false;
    if (i < A.total) {
        const double4 pi = A.in.p[i];
        const double4 va = A.in.v[2 * i], vb = A.in.v[2 * i + 1];
        V3 F = v3(A.Fx[i], A.Fy[i], A.Fz[i]), M = v3(A.Mx[i], A.My[i], A.Mz[i]);
        over = integrate(P, A, i, pi, va, vb, F, M, A.Mr[i]);
    }
    unsigned long long b = __ballot(over);
    if (b != 0ull && (threadIdx.x & 63) == __ffsll((long long)b) - 1)
        atomicAdd(count, (int)__popcll(b));
}

void launch_pp_force(const DevInst &P, const StepArgs &A, bool wet, bool mixed, hipStream_t s) {
    if (A.total <= 0) return;
    dim3 g(blocks_for(A.total, 256)), b(256);
    if (wet && mixed) hipLaunchKernelGGL((k_pp_force<true, true>), g, b, 0, s, P, A);
    else if (wet) hipLaunchKernelGGL((k_pp_force<true, false>), g, b, 0, s, P, A);
    else if (mixed) hipLaunchKernelGGL((k_pp_force<false, true>), g, b, 0, s, P, A);
    else hipLaunchKernelGGL((k_pp_force<false, false>), g, b, 0, s, P, A);
}
void launch_pw_force(const DevInst &P, const StepArgs &A, bool mixed, hipStream_t s) {
    if (A.total <= 0) return;
    dim3 g(blocks_for(A.total, 256)), b(256);
    if (mixed) hipLaunchKernelGGL((k_pw_force<true>), g, b, 0, s, P, A);
    else hipLaunchKernelGGL((k_pw_force<false>), g, b, 0, s, P, A);
}
void launch_update(const DevInst &P, const StepArgs &A, int *count, hipStream_t s) {
    if (A.total <= 0) return;
    hipLaunchKernelGGL(k_update, dim3(blocks_for(A.total, 256)), dim3(256), 0, s, P, A, count);
}

// ---------------------------------------------------------------------
// import / export (original particle order)
// ---------------------------------------------------------------------
// what: 0 position, 1 velocity, 2 spin, 3 radius (one value)
__global__ void k_gather_state(const int *id, StateBuf s, int what, double *out, int total) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const long o = (long)id[i];
    if (what == 0) {
        const double4 q = s.p[i];
        out[3 * o] = q.x; out[3 * o + 1] = q.y; out[3 * o + 2] = q.z;
    } else if (what == 1) {
        const double4 a = s.v[2 * i];
        out[3 * o] = a.x; out[3 * o + 1] = a.y; out[3 * o + 2] = a.z;
    } else if (what == 2) {
        const double4 a = s.v[2 * i], b = s.v[2 * i + 1];
        out[3 * o] = a.w; out[3 * o + 1] = b.x; out[3 * o + 2] = b.y;
    } else if (what == 3) {
        out[o] = s.p[i].w;
    } else {
        out[o] = s.v[2 * i + 1].z;
    }
}
__global__ void k_gather3(const int *id, const double *a, const double *b, const double *c,
                          double *out, int total) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    long o = (long)id[i] * 3;
    out[o] = a[i]; out[o + 1] = b[i]; out[o + 2] = c[i];
}
__global__ void k_gather1(const int *id, const double *a, double *out, int total) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    out[id[i]] = a[i];
}
__global__ void k_gather1i(const int *id, const int *a, int *out, int total) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    out[id[i]] = a[i];
}
// particles off..off+n-1 from host AoS arrays (vel/angv may be null)
__global__ void k_scatter_state(const double *pos, const double *vel, const double *angv,
                                const double *rad, StateBuf s, int n, int off) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int i = off + k;
    double4 q;
    q.x = pos[3 * (long)k]; q.y = pos[3 * (long)k + 1]; q.z = pos[3 * (long)k + 2];
    q.w = rad[k];
    s.p[i] = q;
    double4 a, b;
    a.x = vel ? vel[3 * (long)k] : 0.0;
    a.y = vel ? vel[3 * (long)k + 1] : 0.0;
    a.z = vel ? vel[3 * (long)k + 2] : 0.0;
    a.w = angv ? angv[3 * (long)k] : 0.0;
    b.x = angv ? angv[3 * (long)k + 1] : 0.0;
    b.y = angv ? angv[3 * (long)k + 2] : 0.0;
    b.z = 0.0;
    b.w = 0.0;
    s.v[2 * i] = a;
    s.v[2 * i + 1] = b;
}
void launch_gather_state(const int *id, const StateBuf &s, int what, double *out, int total,
                         hipStream_t st) {
    if (total > 0)
        hipLaunchKernelGGL(k_gather_state, dim3(blocks_for(total, 256)), dim3(256), 0, st, id, s,
                           what, out, total);
}
void launch_gather3(const int *id, const double *a, const double *b, const double *c, double *out,
                    int total, hipStream_t s) {
    if (total > 0)
        hipLaunchKernelGGL(k_gather3, dim3(blocks_for(total, 256)), dim3(256), 0, s, id, a, b, c,
                           out, total);
}
void launch_gather1(const int *id, const double *a, double *out, int total, hipStream_t s) {
    if (total > 0)
        hipLaunchKernelGGL(k_gather1, dim3(blocks_for(total, 256)), dim3(256), 0, s, id, a, out,
                           total);
}
void launch_gather1i(const int *id, const int *a, int *out, int total, hipStream_t s) {
    if (total > 0)
        hipLaunchKernelGGL(k_gather1i, dim3(blocks_for(total, 256)), dim3(256), 0, s, id, a, out,
                           total);
}
void launch_scatter_state(const double *pos, const double *vel, const double *angv,
                          const double *rad, const StateBuf &s, int n, int off, hipStream_t st) {
    if (n > 0)
        hipLaunchKernelGGL(k_scatter_state, dim3(blocks_for(n, 256)), dim3(256), 0, st, pos, vel,
                           angv, rad, s, n, off);
}

// ---------------------------------------------------------------------
// plain device copy: the attainable HBM bandwidth, and the known-size
// kernel that calibrates the FETCH_SIZE / WRITE_SIZE counters
// ---------------------------------------------------------------------
// Each block copies tiles of 4 x 256 consecutive 16-byte elements: four
// independent loads per lane, then four stores, all non-temporal (the data
// is touched once).  Of the forms timed on an MI355X (grid-stride single
// loads, 2-16 loads per lane, contiguous or strided, temporal or not) this
// one moved the most: 5.86 TB/s of read + write at 64 blocks per CU.
__global__ void __launch_bounds__(256) k_stream_copy(const double2 *__restrict__ a,
                                                     double2 *__restrict__ b, long n) {
    for (long base = (long)blockIdx.x * 1024; base < n; base += (long)gridDim.x * 1024) {
        double2 v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const long i = base + u * 256 + threadIdx.x;
            if (i < n) {
                v[u].x = __builtin_nontemporal_load(&a[i].x);
                v[u].y = __builtin_nontemporal_load(&a[i].y);
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const long i = base + u * 256 + threadIdx.x;
            if (i < n) {
                __builtin_nontemporal_store(v[u].x, &b[i].x);
                __builtin_nontemporal_store(v[u].y, &b[i].y);
            }
        }
    }
}

// blocks_per_cu x 256 CUs of 256 threads
void launch_stream_copy(const void *a, void *b, long nbytes, int blocks_per_cu, hipStream_t s) {
    long n = nbytes / 16;
    if (n > 0)
        hipLaunchKernelGGL(k_stream_copy, dim3(256 * blocks_per_cu), dim3(256), 0, s,
                           (const double2 *)a, (double2 *)b, n);
}

// ---------------------------------------------------------------------
// velocity and spin scaling (a settling aid: a quench to rest keeps the
// contact history)
// ---------------------------------------------------------------------
__global__ void k_scale_velocities(double4 *v, double f, int total) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    double4 a = v[2 * i], b = v[2 * i + 1];
    a.x *= f; a.y *= f; a.z *= f; a.w *= f;  // vx vy vz wx
    b.x *= f; b.y *= f;                      // wy wz (path length and tag kept)
    v[2 * i] = a;
    v[2 * i + 1] = b;
}

void launch_scale_velocities(const StateBuf &s, double f, int total, hipStream_t st) {
    if (total <= 0) return;
    hipLaunchKernelGGL(k_scale_velocities, dim3(blocks_for(total, 256)), dim3(256), 0, st, s.v, f,
                       total);
}

// ---------------------------------------------------------------------
// contact statistics (for the roofline bytes)
// ---------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_stats(DevInst P, StepArgs A, unsigned long long *out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long nn = 0, nc = 0, nw = 0, nwc = 0, nbr = 0;
    if (i < A.total) {
        const StateBuf &s = A.in;
        const double4 pi = s.p[i];
        nn = A.nnb[i];
        for (int k = 0; k < (int)nn; ++k) {
            long sk = (long)k * A.cap + i;
            int j = A.nbr[sk];
            const double4 pj = s.p[j];
            double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
            double gap = sqrt(dx * dx + dy * dy + dz * dz) - pi.w - pj.w;
            if (gap < 0.0) nc++;
            if (A.h &&
                (A.h[hist_idx(A.amask[i], A.actoff[i], k, i, A.cap, (long)A.K * A.cap)].flags >> 8) & 1)
                nbr++;
        }
        nw = A.nwl[i];
        for (int k = 0; k < (int)nw; ++k) {
            int w = A.wl[(long)k * A.cap + i];
            double gap;
            if (w == 0) gap = (A.Tw - pi.z) - pi.w;
            else if (w == 1) gap = (pi.z - P.Bw) - pi.w;
            else {
                double ax = P.cirX - pi.x, ay = P.cirY - pi.y;
                gap = (P.R - sqrt(ax * ax + ay * ay)) - pi.w;
            }
            if (gap < 0.0) nwc++;
        }
    }
    __shared__ unsigned long long red[5][4];
    unsigned long long v[5] = {nn, nc, nw, nwc, nbr};
    for (int q = 0; q < 5; ++q) {
        unsigned long long t = v[q];
        for (int d = 32; d >= 1; d >>= 1) t += __shfl_xor(t, d, 64);
        if ((threadIdx.x & 63) == 0) red[q][threadIdx.x >> 6] = t;
    }
    __syncthreads();
    if (threadIdx.x < 5) {
        unsigned long long t = 0;
        for (int w = 0; w < 4; ++w) t += red[threadIdx.x][w];
        atomicAdd(&out[threadIdx.x], t);
    }
}

void launch_stats(const DevInst &P, const StepArgs &A, unsigned long long *out, hipStream_t s) {
    if (A.total <= 0) return;
    hipLaunchKernelGGL(k_stats, dim3(blocks_for(A.total, 256)), dim3(256), 0, s, P, A, out);
}
