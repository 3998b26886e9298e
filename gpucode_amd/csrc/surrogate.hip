// surrogate.hip -- the sweep surrogate MLP (8 -> 128 -> 128 -> 128 -> 2,
// SiLU): layer products on the MFMA units (BF16 operands, FP32
// accumulation), backward pass, Adam, and the C ABI of dem_surrogate.h.
//
// Every product of the forward and backward pass is one call of a strided
// product kernel C(i,j) = sum_k A(i,k) B(k,j): one wave owns a 16 x 16
// tile of C and walks k in steps of 16 with
// v_mfma_f32_16x16x16_bf16.  For that instruction lane l holds
// A(l%16, 4(l/16)+r) and B(4(l/16)+r, l%16), r = 0..3, and receives
// C(4(l/16)+r, l%16).
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "../../include/dem_capi.h"
#include "../../include/dem_surrogate.h"

int dem_fail(int code, const char *fmt, ...);

#define SCHK(call)                                                                       \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess)                                                            \
            return dem_fail(DEM_ERR_HIP, "%s:%d %s: %s", __FILE__, __LINE__, #call,       \
                            hipGetErrorString(e_));                                      \
    } while (0)

typedef unsigned short bf16;
typedef short sh4 __attribute__((ext_vector_type(4)));
typedef float fl4 __attribute__((ext_vector_type(4)));

// parameter layout
static const int H = SUR_HIDDEN;
static const int kW1 = 0, kB1 = kW1 + H * SUR_IN, kW2 = kB1 + H, kB2 = kW2 + H * H, kW3 = kB2 + H,
                 kB3 = kW3 + H * H, kW4 = kB3 + H, kB4 = kW4 + SUR_OUT * H,
                 kNP = kB4 + SUR_OUT;

__device__ inline bf16 f2bf(float f) {
    unsigned u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (bf16)0x7fc0;  // NaN
    u += 0x7fffu + ((u >> 16) & 1u);                           // round to nearest even
    return (bf16)(u >> 16);
}
__device__ inline float bf2f(bf16 h) { return __uint_as_float((unsigned)h << 16); }
__device__ inline float sigmoidf_(float z) { return 1.0f / (1.0f + expf(-z)); }

struct GemmArgs {
    const bf16 *A;
    long as0, as1;        // A(i,k) = A[i*as0 + k*as1]
    const bf16 *B;
    long bs0, bs1;        // B(k,j) = B[k*bs0 + j*bs1]
    int M, N, K;
    float *C;             // C(i,j) = C[i*cs0 + j], FP32 (may be null)
    long cs0;
    const float *bias;    // added per column j (may be null)
    bf16 *Hs;             // SiLU(C) as BF16, row stride hs0 (may be null)
    long hs0;
    bf16 *Cb;             // C as BF16, row stride cs0 (may be null)
};

__global__ void __launch_bounds__(256) k_gemm_bf16(GemmArgs g) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int ti = (blockIdx.x * 4 + w) * 16, tj = blockIdx.y * 16;
    if (ti >= g.M) return;  // whole wave
    fl4 acc = {0.f, 0.f, 0.f, 0.f};
    const int i = ti + (lane & 15), j = tj + (lane & 15), kq = 4 * (lane >> 4);
    for (int k0 = 0; k0 < g.K; k0 += 16) {
        sh4 a, b;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int k = k0 + kq + r;
            a[r] = (short)((i < g.M && k < g.K) ? g.A[(long)i * g.as0 + (long)k * g.as1] : 0);
            b[r] = (short)((j < g.N && k < g.K) ? g.B[(long)k * g.bs0 + (long)j * g.bs1] : 0);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int oi = ti + kq + r, oj = tj + (lane & 15);
        if (oi >= g.M || oj >= g.N) continue;
        float v = acc[r];
        if (g.bias) v += g.bias[oj];
        if (g.C) g.C[(long)oi * g.cs0 + oj] = v;
        if (g.Cb) g.Cb[(long)oi * g.cs0 + oj] = f2bf(v);
        if (g.Hs) g.Hs[(long)oi * g.hs0 + oj] = f2bf(v * sigmoidf_(v));
    }
}

static void gemm(const GemmArgs &g, hipStream_t s) {
    if (g.M <= 0 || g.N <= 0) return;
    dim3 grid((unsigned)((g.M + 63) / 64), (unsigned)((g.N + 15) / 16));
    hipLaunchKernelGGL(k_gemm_bf16, grid, dim3(256), 0, s, g);
}

__global__ void k_to_bf16(const float *x, bf16 *y, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = f2bf(x[i]);
}

// dY = 2 (Y - T) / (n * 2), FP32 and BF16
__global__ void k_mse_grad(const float *Y, const float *T, float *dY, bf16 *dYb, long n2) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n2) {
        const float g = 2.0f * (Y[i] - T[i]) / (float)n2;
        dY[i] = g;
        dYb[i] = f2bf(g);
    }
}

// dZ = dH * SiLU'(Z), SiLU'(z) = s (1 + z (1 - s)), s = sigmoid(z)
__global__ void k_silu_back(const float *dH, const float *Z, float *dZ, bf16 *dZb, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const float z = Z[i], s = sigmoidf_(z);
        const float g = dH[i] * (s * (1.0f + z * (1.0f - s)));
        dZ[i] = g;
        dZb[i] = f2bf(g);
    }
}

// db[j] = sum_i dZ[i][j] (rows in order)
__global__ void k_bias_grad(const float *dZ, int n, int N, float *db) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N) return;
    float s = 0.f;
    for (int i = 0; i < n; ++i) s += dZ[(long)i * N + j];
    db[j] = s;
}

__global__ void k_adam(float *p, const float *g, float *m, float *v, bf16 *pb, int n, float lr,
                       float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = 0.9f * m[i] + 0.1f * gi;
    const float vi = 0.999f * v[i] + 0.001f * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + 1e-8f);
    p[i] = pi;
    pb[i] = f2bf(pi);
}

__global__ void k_gather_rows(const float *src, const int *idx, float *dst, int n, int w) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n * w) dst[t] = src[(long)idx[t / w] * w + t % w];
}

static unsigned blocks(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

struct sur_mlp {
    hipStream_t stream = nullptr;
    float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;  // FP32 parameters, grads, Adam
    bf16 *pb = nullptr;                                           // BF16 parameters
    int cap = 0;                                                  // rows of the buffers below
    float *X = nullptr, *T = nullptr, *Y = nullptr, *dY = nullptr;
    bf16 *Xb = nullptr, *dYb = nullptr;
    float *Z[3] = {}, *dH = nullptr, *dZ = nullptr;
    bf16 *Hb[3] = {}, *dZb = nullptr;
    float *Xall = nullptr, *Tall = nullptr;                       // training set
    int *idx = nullptr;
    int tcap = 0;
    long step = 0;
    unsigned char *scratch = nullptr;  // this GPU's copy of `scratch` (below)
};

// Activations of sur_forward_noalloc: a fixed array in device memory, one
// copy per GPU the library runs on, held until the process ends.  Rows of
// at most SUR_NOALLOC_MAX_ROWS inputs: the BF16 inputs (16 B per row), then
// two BF16 activation buffers that the hidden layers use in turn (256 B per
// row each).  Every element is written before it is read, so nothing is
// cleared; every call on a GPU uses the same array, so calls that overlap on
// one GPU race.
static const size_t kNoAllocXb = (size_t)SUR_NOALLOC_MAX_ROWS * SUR_IN * sizeof(bf16);
static const size_t kNoAllocH = (size_t)SUR_NOALLOC_MAX_ROWS * SUR_HIDDEN * sizeof(bf16);
static __device__ __align__(256) unsigned char scratch[kNoAllocXb + 2 * kNoAllocH];

// Device address of `scratch` on each GPU, looked up once per device with
// that device current (the array's name is a host-side stand-in, not a
// device address).
static const int kMaxDevices = 64;
static unsigned char *g_scratch_dev[kMaxDevices];
static std::mutex g_scratch_mu;

static int scratch_on_current_device(unsigned char **out) {
    int dev = -1;
    SCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return dem_fail(DEM_ERR_ARG, "device %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        SCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_dev[dev] = (unsigned char *)p;
    }
    *out = g_scratch_dev[dev];
    return DEM_OK;
}

static void free_rows(sur_mlp *s) {
    (void)hipFree(s->X); (void)hipFree(s->T); (void)hipFree(s->Y); (void)hipFree(s->dY);
    (void)hipFree(s->Xb); (void)hipFree(s->dYb);
    for (int l = 0; l < 3; ++l) { (void)hipFree(s->Z[l]); (void)hipFree(s->Hb[l]); }
    (void)hipFree(s->dH); (void)hipFree(s->dZ); (void)hipFree(s->dZb);
    s->X = s->T = s->Y = s->dY = s->dH = s->dZ = nullptr;
    s->Xb = s->dYb = s->dZb = nullptr;
    for (int l = 0; l < 3; ++l) { s->Z[l] = nullptr; s->Hb[l] = nullptr; }
    s->cap = 0;
}

static int ensure_rows(sur_mlp *s, int n) {
    if (n <= s->cap) return DEM_OK;
    free_rows(s);
    const size_t r = (size_t)n;
    SCHK(hipMalloc((void **)&s->X, r * SUR_IN * sizeof(float)));
    SCHK(hipMalloc((void **)&s->T, r * SUR_OUT * sizeof(float)));
    SCHK(hipMalloc((void **)&s->Y, r * SUR_OUT * sizeof(float)));
    SCHK(hipMalloc((void **)&s->dY, r * SUR_OUT * sizeof(float)));
    SCHK(hipMalloc((void **)&s->Xb, r * SUR_IN * sizeof(bf16)));
    SCHK(hipMalloc((void **)&s->dYb, r * SUR_OUT * sizeof(bf16)));
    for (int l = 0; l < 3; ++l) {
        SCHK(hipMalloc((void **)&s->Z[l], r * H * sizeof(float)));
        SCHK(hipMalloc((void **)&s->Hb[l], r * H * sizeof(bf16)));
    }
    SCHK(hipMalloc((void **)&s->dH, r * H * sizeof(float)));
    SCHK(hipMalloc((void **)&s->dZ, r * H * sizeof(float)));
    SCHK(hipMalloc((void **)&s->dZb, r * H * sizeof(bf16)));
    s->cap = n;
    return DEM_OK;
}

// Buffers of one forward pass: input rows X (FP32), their BF16 copy Xb,
// the pre-activations Z[l] (FP32, may be null when no backward pass
// follows), the SiLU activations Hb[l] (BF16) and the outputs Y.
struct FwdBufs {
    const float *X;
    bf16 *Xb;
    float *Z[3];
    bf16 *Hb[3];
    float *Y;
};

static void forward_into(const sur_mlp *s, const FwdBufs &b, int n, hipStream_t st) {
    hipLaunchKernelGGL(k_to_bf16, dim3(blocks((long)n * SUR_IN, 256)), dim3(256), 0, st, b.X, b.Xb,
                       (long)n * SUR_IN);
    const bf16 *in = b.Xb;
    int K = SUR_IN;
    const int wo[3] = {kW1, kW2, kW3}, bo[3] = {kB1, kB2, kB3};
    for (int l = 0; l < 3; ++l) {
        GemmArgs g = {};
        g.A = in; g.as0 = K; g.as1 = 1;
        g.B = s->pb + wo[l]; g.bs0 = 1; g.bs1 = K;  // B(k,j) = W[j][k]
        g.M = n; g.N = H; g.K = K;
        g.C = b.Z[l]; g.cs0 = H; g.bias = s->p + bo[l];
        g.Hs = b.Hb[l]; g.hs0 = H;
        gemm(g, st);
        in = b.Hb[l];
        K = H;
    }
    GemmArgs g = {};
    g.A = in; g.as0 = H; g.as1 = 1;
    g.B = s->pb + kW4; g.bs0 = 1; g.bs1 = H;
    g.M = n; g.N = SUR_OUT; g.K = H;
    g.C = b.Y; g.cs0 = SUR_OUT; g.bias = s->p + kB4;
    gemm(g, st);
}

// forward pass of n rows already in s->X (FP32): Xb, Z[l], Hb[l], Y
static void forward(sur_mlp *s, int n) {
    FwdBufs b = {s->X, s->Xb, {s->Z[0], s->Z[1], s->Z[2]}, {s->Hb[0], s->Hb[1], s->Hb[2]}, s->Y};
    forward_into(s, b, n, s->stream);
}

// backward pass after forward(n) with targets in s->T: gradients into s->g
static void backward(sur_mlp *s, int n) {
    const long n2 = (long)n * SUR_OUT;
    hipLaunchKernelGGL(k_mse_grad, dim3(blocks(n2, 256)), dim3(256), 0, s->stream, s->Y, s->T, s->dY,
                       s->dYb, n2);
    // output layer: dW4 = dY^T H3, db4, dH3 = dY W4
    {
        GemmArgs g = {};
        g.A = s->dYb; g.as0 = 1; g.as1 = SUR_OUT;   // A(o,i) = dY[i][o]
        g.B = s->Hb[2]; g.bs0 = H; g.bs1 = 1;       // B(i,k) = H3[i][k]
        g.M = SUR_OUT; g.N = H; g.K = n;
        g.C = s->g + kW4; g.cs0 = H;
        gemm(g, s->stream);
        hipLaunchKernelGGL(k_bias_grad, dim3(1), dim3(64), 0, s->stream, s->dY, n, SUR_OUT, s->g + kB4);
        GemmArgs d = {};
        d.A = s->dYb; d.as0 = SUR_OUT; d.as1 = 1;   // A(i,o) = dY[i][o]
        d.B = s->pb + kW4; d.bs0 = H; d.bs1 = 1;    // B(o,k) = W4[o][k]
        d.M = n; d.N = H; d.K = SUR_OUT;
        d.C = s->dH; d.cs0 = H;
        gemm(d, s->stream);
    }
    const int wo[3] = {kW1, kW2, kW3}, bo[3] = {kB1, kB2, kB3};
    for (int l = 2; l >= 0; --l) {
        const long nh = (long)n * H;
        hipLaunchKernelGGL(k_silu_back, dim3(blocks(nh, 256)), dim3(256), 0, s->stream, s->dH, s->Z[l],
                           s->dZ, s->dZb, nh);
        const int K = l == 0 ? SUR_IN : H;
        const bf16 *in = l == 0 ? s->Xb : s->Hb[l - 1];
        GemmArgs g = {};
        g.A = s->dZb; g.as0 = 1; g.as1 = H;         // A(o,i) = dZ[i][o]
        g.B = in; g.bs0 = K; g.bs1 = 1;             // B(i,k) = in[i][k]
        g.M = H; g.N = K; g.K = n;
        g.C = s->g + wo[l]; g.cs0 = K;
        gemm(g, s->stream);
        hipLaunchKernelGGL(k_bias_grad, dim3(blocks(H, 128)), dim3(128), 0, s->stream, s->dZ, n, H,
                           s->g + bo[l]);
        if (l > 0) {
            GemmArgs d = {};
            d.A = s->dZb; d.as0 = H; d.as1 = 1;     // A(i,o) = dZ[i][o]
            d.B = s->pb + wo[l]; d.bs0 = K; d.bs1 = 1;  // B(o,k) = W[o][k]
            d.M = n; d.N = K; d.K = H;
            d.C = s->dH; d.cs0 = K;
            gemm(d, s->stream);
        }
    }
}

static void refresh_bf16(sur_mlp *s) {
    hipLaunchKernelGGL(k_to_bf16, dim3(blocks(kNP, 256)), dim3(256), 0, s->stream, s->p, s->pb,
                       (long)kNP);
}

extern "C" int sur_nparams(void) { return kNP; }

static double splitmix_u01(unsigned long long &s) {
    unsigned long long z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z = z ^ (z >> 31);
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

extern "C" int sur_destroy(sur_mlp *s) {
    if (!s) return DEM_OK;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    free_rows(s);
    (void)hipFree(s->p); (void)hipFree(s->g); (void)hipFree(s->m); (void)hipFree(s->v);
    (void)hipFree(s->pb); (void)hipFree(s->Xall); (void)hipFree(s->Tall); (void)hipFree(s->idx);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
    return DEM_OK;
}

extern "C" int sur_set_params(sur_mlp *s, const float *params) {
    if (!s || !params) return dem_fail(DEM_ERR_ARG, "null argument");
    SCHK(hipMemcpyAsync(s->p, params, kNP * sizeof(float), hipMemcpyHostToDevice, s->stream));
    refresh_bf16(s);
    SCHK(hipMemsetAsync(s->m, 0, kNP * sizeof(float), s->stream));
    SCHK(hipMemsetAsync(s->v, 0, kNP * sizeof(float), s->stream));
    SCHK(hipStreamSynchronize(s->stream));
    s->step = 0;
    return DEM_OK;
}

extern "C" int sur_get_params(sur_mlp *s, float *params) {
    if (!s || !params) return dem_fail(DEM_ERR_ARG, "null argument");
    SCHK(hipMemcpyAsync(params, s->p, kNP * sizeof(float), hipMemcpyDeviceToHost, s->stream));
    SCHK(hipStreamSynchronize(s->stream));
    return DEM_OK;
}

extern "C" int sur_create(unsigned long long seed, sur_mlp **out) {
    if (!out) return dem_fail(DEM_ERR_ARG, "null argument");
    *out = nullptr;
    sur_mlp *s = new sur_mlp();
    // the network lives on the current device; so does its scratch array
    int src = scratch_on_current_device(&s->scratch);
    if (src) {
        delete s;
        return src;
    }
    hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc((void **)&s->p, kNP * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void **)&s->g, kNP * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void **)&s->m, kNP * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void **)&s->v, kNP * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void **)&s->pb, kNP * sizeof(bf16));
    if (e != hipSuccess) {
        sur_destroy(s);
        return dem_fail(DEM_ERR_HIP, "surrogate allocation: %s", hipGetErrorString(e));
    }
    // Glorot-uniform weights, zero biases
    std::vector<float> h(kNP, 0.0f);
    unsigned long long st = seed;
    const int wo[4] = {kW1, kW2, kW3, kW4}, fin[4] = {SUR_IN, H, H, H}, fout[4] = {H, H, H, SUR_OUT};
    for (int l = 0; l < 4; ++l) {
        const double a = sqrt(6.0 / (fin[l] + fout[l]));
        for (int q = 0; q < fin[l] * fout[l]; ++q) h[wo[l] + q] = (float)(a * (2.0 * splitmix_u01(st) - 1.0));
    }
    int rc = sur_set_params(s, h.data());
    if (rc) {
        sur_destroy(s);
        return rc;
    }
    *out = s;
    return DEM_OK;
}

extern "C" int sur_forward(sur_mlp *s, int n, const float *x, float *y) {
    if (!s || n < 0 || (n > 0 && (!x || !y))) return dem_fail(DEM_ERR_ARG, "bad argument");
    if (n == 0) return DEM_OK;
    int rc = ensure_rows(s, n);
    if (rc) return rc;
    SCHK(hipMemcpyAsync(s->X, x, (size_t)n * SUR_IN * sizeof(float), hipMemcpyHostToDevice, s->stream));
    forward(s, n);
    SCHK(hipGetLastError());
    SCHK(hipMemcpyAsync(y, s->Y, (size_t)n * SUR_OUT * sizeof(float), hipMemcpyDeviceToHost, s->stream));
    SCHK(hipStreamSynchronize(s->stream));
    return DEM_OK;
}

extern "C" int sur_forward_noalloc(sur_mlp *s, int n, const float *x, float *y) {
    if (!s || n < 0 || (n > 0 && (!x || !y))) return dem_fail(DEM_ERR_ARG, "bad argument");
    if (n > SUR_NOALLOC_MAX_ROWS)
        return dem_fail(DEM_ERR_ARG, "sur_forward_noalloc: %d rows, at most %d per call", n,
                        SUR_NOALLOC_MAX_ROWS);
    if (n == 0) return DEM_OK;
    // the device address found at sur_create, never the array's name
    unsigned char *base = s->scratch;
    bf16 *Xb = (bf16 *)base;
    bf16 *H0 = (bf16 *)(base + kNoAllocXb), *H1 = (bf16 *)(base + kNoAllocXb + kNoAllocH);
    FwdBufs b = {x, Xb, {nullptr, nullptr, nullptr}, {H0, H1, H0}, y};
    forward_into(s, b, n, s->stream);
    SCHK(hipGetLastError());
    SCHK(hipStreamSynchronize(s->stream));
    return DEM_OK;
}

extern "C" int sur_loss_grad(sur_mlp *s, int n, const float *x, const float *t, float *loss, float *grad) {
    if (!s || n <= 0 || !x || !t) return dem_fail(DEM_ERR_ARG, "bad argument");
    int rc = ensure_rows(s, n);
    if (rc) return rc;
    SCHK(hipMemcpyAsync(s->X, x, (size_t)n * SUR_IN * sizeof(float), hipMemcpyHostToDevice, s->stream));
    SCHK(hipMemcpyAsync(s->T, t, (size_t)n * SUR_OUT * sizeof(float), hipMemcpyHostToDevice, s->stream));
    forward(s, n);
    backward(s, n);
    SCHK(hipGetLastError());
    std::vector<float> yh((size_t)n * SUR_OUT);
    SCHK(hipMemcpyAsync(yh.data(), s->Y, yh.size() * sizeof(float), hipMemcpyDeviceToHost, s->stream));
    if (grad) SCHK(hipMemcpyAsync(grad, s->g, kNP * sizeof(float), hipMemcpyDeviceToHost, s->stream));
    SCHK(hipStreamSynchronize(s->stream));
    if (loss) {
        double acc = 0.0;
        for (size_t q = 0; q < yh.size(); ++q) acc += ((double)yh[q] - t[q]) * ((double)yh[q] - t[q]);
        *loss = (float)(acc / (double)yh.size());
    }
    return DEM_OK;
}

extern "C" int sur_train(sur_mlp *s, int n, const float *x, const float *t, int epochs, int batch, float lr,
                         unsigned long long seed, float *loss) {
    if (!s || n <= 0 || !x || !t || epochs < 0 || batch <= 0) return dem_fail(DEM_ERR_ARG, "bad argument");
    if (batch > n) batch = n;
    if (n > s->tcap) {
        (void)hipFree(s->Xall); (void)hipFree(s->Tall); (void)hipFree(s->idx);
        s->Xall = s->Tall = nullptr;
        s->idx = nullptr;
        s->tcap = 0;
        SCHK(hipMalloc((void **)&s->Xall, (size_t)n * SUR_IN * sizeof(float)));
        SCHK(hipMalloc((void **)&s->Tall, (size_t)n * SUR_OUT * sizeof(float)));
        SCHK(hipMalloc((void **)&s->idx, (size_t)n * sizeof(int)));
        s->tcap = n;
    }
    int rc = ensure_rows(s, batch > n ? n : batch);
    if (rc) return rc;
    SCHK(hipMemcpyAsync(s->Xall, x, (size_t)n * SUR_IN * sizeof(float), hipMemcpyHostToDevice, s->stream));
    SCHK(hipMemcpyAsync(s->Tall, t, (size_t)n * SUR_OUT * sizeof(float), hipMemcpyHostToDevice, s->stream));
    std::vector<int> perm(n);
    for (int q = 0; q < n; ++q) perm[q] = q;
    unsigned long long st = seed;
    for (int ep = 0; ep < epochs; ++ep) {
        for (int q = n - 1; q > 0; --q) {  // Fisher-Yates
            const int r = (int)(splitmix_u01(st) * (q + 1));
            std::swap(perm[q], perm[r < q ? r : q]);
        }
        SCHK(hipMemcpyAsync(s->idx, perm.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice, s->stream));
        for (int b0 = 0; b0 + batch <= n; b0 += batch) {
            hipLaunchKernelGGL(k_gather_rows, dim3(blocks((long)batch * SUR_IN, 256)), dim3(256), 0, s->stream,
                               s->Xall, s->idx + b0, s->X, batch, SUR_IN);
            hipLaunchKernelGGL(k_gather_rows, dim3(blocks((long)batch * SUR_OUT, 256)), dim3(256), 0,
                               s->stream, s->Tall, s->idx + b0, s->T, batch, SUR_OUT);
            forward(s, batch);
            backward(s, batch);
            s->step++;
            const float c1 = 1.0f - powf(0.9f, (float)s->step), c2 = 1.0f - powf(0.999f, (float)s->step);
            hipLaunchKernelGGL(k_adam, dim3(blocks(kNP, 256)), dim3(256), 0, s->stream, s->p, s->g, s->m,
                               s->v, s->pb, kNP, lr, c1, c2);
        }
        // the permutation buffer is rewritten next epoch: finish this one
        SCHK(hipStreamSynchronize(s->stream));
    }
    SCHK(hipGetLastError());
    SCHK(hipStreamSynchronize(s->stream));
    if (loss) {
        std::vector<float> y((size_t)n * SUR_OUT);
        rc = sur_forward(s, n, x, y.data());
        if (rc) return rc;
        double acc = 0.0;
        for (size_t q = 0; q < y.size(); ++q) acc += ((double)y[q] - t[q]) * ((double)y[q] - t[q]);
        *loss = (float)(acc / (double)y.size());
    }
    return DEM_OK;
}
