// Learned per-item cost model for the force kernel's work queue.
//
// The reference has no cost model: its only scheduling is a static equal
// split of cells or particles (python-bh/barneshut/implementations/parla.py:
// 188,248; gpu/grid.py:122-156 one cell per block row).  Here a small MLP
// (8 -> 32 -> 32 -> 1, FP32, ReLU) predicts log(item duration) from item
// features; the queue is then ordered longest-first.  The physics does not
// depend on the order: every particle's sum is formed inside one wave.
#include <cmath>

#include "../../include/bh_hip.h"
#include "bh_common.h"
#include "bh_kernels.h"

Item* bh_ctx_items(void* ctx);
int32_t* bh_ctx_stats(void* ctx);
int64_t bh_ctx_max_items(void* ctx);
int bh_ctx_num_cu(void* ctx);
int bh_ctx_p(void* ctx);
const SysDev* bh_ctx_sys(void* ctx);
int bh_ctx_S(void* ctx);
const int32_t* bh_ctx_offsets(void* ctx);
const int32_t* bh_ctx_neprefix(void* ctx);
const Com64* bh_ctx_l64(void* ctx);
int32_t* bh_ctx_cost_scratch(void* ctx);
float* bh_ctx_feat(void* ctx);
float* bh_ctx_pred(void* ctx);
int bh_ctx_fail(int code, const char* msg);

#define NH 32
#define NF BH_MLP_NFEAT
#define OFF_W1 0
#define OFF_B1 (NH * NF)
#define OFF_W2 (OFF_B1 + NH)
#define OFF_B2 (OFF_W2 + NH * NH)
#define OFF_W3 (OFF_B2 + NH)
#define OFF_B3 (OFF_W3 + NH)
static_assert(OFF_B3 + 1 == BH_MLP_NPARAM, "parameter layout");

#define COST_BINS 4096
#define COST_CHUNKS 256   // item chunks of the order's histogram and scatter

// ---------------------------------------------------------------- features
// f0 log1p(np)              particles in the item
// f1 np / (64 P)            lane fill
// f2 log1p(near rows)       particles staged for the near field
// f3 log1p(masked entries)  COM entries in the masked (3x3) ranges
// f4 log1p(K)               non-empty cells of the system (far-field length)
// f5 log1p(rows spanned)
// f6 log1p(neighbour mass)  COM mass in the masked ranges
// f7 log1p(cells in item)   non-empty cells the item covers
__global__ __launch_bounds__(256) void k_item_features(const Item* __restrict__ items,
                                                       const int32_t* __restrict__ stats,
                                                       const SysDev* __restrict__ sys,
                                                       const int32_t* __restrict__ offsets,
                                                       const int32_t* __restrict__ neprefix,
                                                       const Com64* __restrict__ l64, int pitem,
                                                       float* __restrict__ feat)
{
    const int n = stats[ST_NITEMS];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Item it = items[i];
    const SysDev& d = sys[it.sys];
    const int G = d.G;
    const int64_t coff = d.c_off;
    const int ylo = max(it.y0 - 1, 0), yhi = min(it.y1 + 1, G - 1);
    long long nearn = 0, masked = 0;
    double nbm = 0.0;
    for (int dx = -1; dx <= 1; ++dx) {
        int xc = it.cx + dx;
        if (xc < 0 || xc >= G) continue;
        int64_t cl = coff + (int64_t)xc * G;
        nearn += offsets[cl + yhi + 1] - offsets[cl + ylo];
        int a = neprefix[cl + ylo], b = neprefix[cl + yhi + 1];
        masked += b - a;
        for (int k = a; k < b; ++k) nbm += l64[k].M;
    }
    const long long K = neprefix[coff + (int64_t)G * G] - neprefix[coff];
    const int64_t cb = coff + (int64_t)it.cx * G;
    const int cells = neprefix[cb + it.y1 + 1] - neprefix[cb + it.y0];
    float* f = feat + (int64_t)i * NF;
    f[0] = log1pf((float)it.np);
    f[1] = (float)it.np / (float)pitem;
    f[2] = log1pf((float)nearn);
    f[3] = log1pf((float)masked);
    f[4] = log1pf((float)K);
    f[5] = log1pf((float)(it.y1 - it.y0 + 1));
    f[6] = log1pf((float)nbm);
    f[7] = log1pf((float)cells);
}

// ---------------------------------------------------------------- MLP
__device__ __forceinline__ float mlp_row(const float* P, const float* x, float* h1, float* h2)
{
#pragma unroll 4
    for (int j = 0; j < NH; ++j) {
        float a = P[OFF_B1 + j];
#pragma unroll
        for (int i = 0; i < NF; ++i) a = fmaf(P[OFF_W1 + j * NF + i], x[i], a);
        h1[j] = a > 0.0f ? a : 0.0f;
    }
#pragma unroll 4
    for (int j = 0; j < NH; ++j) {
        float a = P[OFF_B2 + j];
        for (int i = 0; i < NH; ++i) a = fmaf(P[OFF_W2 + j * NH + i], h1[i], a);
        h2[j] = a > 0.0f ? a : 0.0f;
    }
    float y = P[OFF_B3];
    for (int j = 0; j < NH; ++j) y = fmaf(P[OFF_W3 + j], h2[j], y);
    return y;
}

// rows < min(n, *dcount) (dcount may be NULL)
__global__ __launch_bounds__(256) void k_mlp_fwd(const float* __restrict__ params,
                                                 const float* __restrict__ X, int64_t n,
                                                 const int32_t* __restrict__ dcount,
                                                 float* __restrict__ Y)
{
    __shared__ float P[BH_MLP_NPARAM];
    for (int t = threadIdx.x; t < BH_MLP_NPARAM; t += blockDim.x) P[t] = params[t];
    __syncthreads();
    int64_t lim = dcount ? min(n, (int64_t)*dcount) : n;
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= lim) return;
    float x[NF], h1[NH], h2[NH];
#pragma unroll
    for (int i = 0; i < NF; ++i) x[i] = X[r * NF + i];
    Y[r] = mlp_row(P, x, h1, h2);
}

#define BWD_ROWS 64
// Per-block gradient of mean((y - t)^2): rows' activations and deltas go to
// LDS, then each thread sums its parameters' terms over the block's rows in
// a fixed order (deterministic; no float atomics).
__global__ __launch_bounds__(256) void k_mlp_bwd_partial(const float* __restrict__ params,
                                                         const float* __restrict__ X,
                                                         const float* __restrict__ T, int64_t n,
                                                         float* __restrict__ partial,
                                                         float* __restrict__ loss_partial)
{
    __shared__ float P[BH_MLP_NPARAM];
    __shared__ float sx[BWD_ROWS][NF];
    __shared__ float sh1[BWD_ROWS][NH + 1];
    __shared__ float sh2[BWD_ROWS][NH + 1];
    __shared__ float sd1[BWD_ROWS][NH + 1];
    __shared__ float sd2[BWD_ROWS][NH + 1];
    __shared__ float sdy[BWD_ROWS];
    __shared__ float sl[BWD_ROWS];
    for (int t = threadIdx.x; t < BH_MLP_NPARAM; t += blockDim.x) P[t] = params[t];
    __syncthreads();
    const float inv_n = 1.0f / (float)n;
    if (threadIdx.x < BWD_ROWS) {
        const int r = threadIdx.x;
        const int64_t row = (int64_t)blockIdx.x * BWD_ROWS + r;
        float x[NF], h1[NH], h2[NH];
        if (row < n) {
#pragma unroll
            for (int i = 0; i < NF; ++i) x[i] = X[row * NF + i];
            float y = mlp_row(P, x, h1, h2);
            float e = y - T[row];
            float dy = 2.0f * e * inv_n;
            sdy[r] = dy;
            sl[r] = e * e * inv_n;
            float d2[NH];
            for (int j = 0; j < NH; ++j) d2[j] = h2[j] > 0.0f ? dy * P[OFF_W3 + j] : 0.0f;
            for (int i = 0; i < NH; ++i) {
                float a = 0.0f;
                for (int j = 0; j < NH; ++j) a = fmaf(P[OFF_W2 + j * NH + i], d2[j], a);
                sd1[r][i] = h1[i] > 0.0f ? a : 0.0f;
            }
            for (int j = 0; j < NH; ++j) {
                sd2[r][j] = d2[j];
                sh1[r][j] = h1[j];
                sh2[r][j] = h2[j];
            }
#pragma unroll
            for (int i = 0; i < NF; ++i) sx[r][i] = x[i];
        } else {
            sdy[r] = 0.0f;
            sl[r] = 0.0f;
            for (int j = 0; j < NH; ++j) {
                sd1[r][j] = 0.0f;
                sd2[r][j] = 0.0f;
                sh1[r][j] = 0.0f;
                sh2[r][j] = 0.0f;
            }
            for (int i = 0; i < NF; ++i) sx[r][i] = 0.0f;
        }
    }
    __syncthreads();
    float* out = partial + (int64_t)blockIdx.x * BH_MLP_NPARAM;
    for (int p = threadIdx.x; p < BH_MLP_NPARAM; p += blockDim.x) {
        float g = 0.0f;
        if (p < OFF_B1) {
            const int j = p / NF, i = p % NF;
            for (int r = 0; r < BWD_ROWS; ++r) g = fmaf(sd1[r][j], sx[r][i], g);
        } else if (p < OFF_W2) {
            const int j = p - OFF_B1;
            for (int r = 0; r < BWD_ROWS; ++r) g += sd1[r][j];
        } else if (p < OFF_B2) {
            const int q = p - OFF_W2, j = q / NH, i = q % NH;
            for (int r = 0; r < BWD_ROWS; ++r) g = fmaf(sd2[r][j], sh1[r][i], g);
        } else if (p < OFF_W3) {
            const int j = p - OFF_B2;
            for (int r = 0; r < BWD_ROWS; ++r) g += sd2[r][j];
        } else if (p < OFF_B3) {
            const int j = p - OFF_W3;
            for (int r = 0; r < BWD_ROWS; ++r) g = fmaf(sdy[r], sh2[r][j], g);
        } else {
            for (int r = 0; r < BWD_ROWS; ++r) g += sdy[r];
        }
        out[p] = g;
    }
    if (threadIdx.x == 0) {
        float l = 0.0f;
        for (int r = 0; r < BWD_ROWS; ++r) l += sl[r];
        loss_partial[blockIdx.x] = l;
    }
}

__global__ __launch_bounds__(256) void k_mlp_bwd_reduce(const float* __restrict__ partial, int nb,
                                                        float* __restrict__ grad,
                                                        const float* __restrict__ loss_partial,
                                                        float* __restrict__ loss)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < BH_MLP_NPARAM) {
        float g = 0.0f;
        for (int b = 0; b < nb; ++b) g += partial[(int64_t)b * BH_MLP_NPARAM + p];
        grad[p] = g;
    }
    if (p == 0 && loss) {
        float l = 0.0f;
        for (int b = 0; b < nb; ++b) l += loss_partial[b];
        *loss = l;
    }
}

__global__ void k_mlp_adam(float* __restrict__ P, const float* __restrict__ g,
                           float* __restrict__ m, float* __restrict__ v, int t, float lr)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= BH_MLP_NPARAM) return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    float gi = g[p];
    float mi = b1 * m[p] + (1.0f - b1) * gi;
    float vi = b2 * v[p] + (1.0f - b2) * gi * gi;
    m[p] = mi;
    v[p] = vi;
    float mh = mi / (1.0f - powf(b1, (float)t));
    float vh = vi / (1.0f - powf(b2, (float)t));
    P[p] -= lr * mh / (sqrtf(vh) + eps);
}

// ---------------------------------------------------------------- cost order
// scratch layout (int32): [0..1] ordered min/max keys, [2] non-finite flag,
// [16 ..) hist[COST_BINS], offsets[COST_BINS + 1], cursor[COST_BINS]
__device__ __forceinline__ uint32_t f_ord(float f)
{
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float f_val(uint32_t k)
{
    uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    return __uint_as_float(u);
}

__global__ void k_cost_init(int32_t* sc)
{
    int t = threadIdx.x + blockIdx.x * blockDim.x;
    if (t == 0) {
        sc[0] = (int32_t)0xffffffffu;
        sc[1] = 0;
        sc[2] = 0;
    }
    int32_t* hist = sc + 16;
    int32_t* cur = sc + 16 + 2 * COST_BINS + 1;
    for (int b = t; b < COST_BINS; b += blockDim.x * gridDim.x) {
        hist[b] = 0;
        cur[b] = 0;
    }
}

// Min and max key of the finite predictions, and a flag if any prediction is
// not finite: a grid-stride loop, a wave reduction, then one pair of atomics
// per wave (one pair per item would serialise on the two words).
__global__ __launch_bounds__(256) void k_cost_minmax(const float* __restrict__ cost,
                                                     const int32_t* __restrict__ stats, int32_t* sc)
{
    const int n = stats[ST_NITEMS];
    uint32_t lo = 0xffffffffu, hi = 0u;
    bool bad = false;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float c = cost[i];
        if (!isfinite(c)) {
            bad = true;
            continue;
        }
        const uint32_t k = f_ord(c);
        lo = min(lo, k);
        hi = max(hi, k);
    }
#pragma unroll
    for (int o = BH_WAVE / 2; o > 0; o >>= 1) {
        lo = min(lo, (uint32_t)__shfl_xor((int)lo, o, BH_WAVE));
        hi = max(hi, (uint32_t)__shfl_xor((int)hi, o, BH_WAVE));
    }
    const bool any_bad = __ballot(bad) != 0ull;
    if ((threadIdx.x & (BH_WAVE - 1)) == 0) {
        if (lo <= hi) {
            atomicMin((unsigned int*)&sc[0], lo);
            atomicMax((unsigned int*)&sc[1], hi);
        }
        if (any_bad) atomicOr(&sc[2], 1);
    }
}

__device__ __forceinline__ int cost_bin(float c, float mn, float mx)
{
    float w = mx - mn;
    int b = w > 0.0f ? (int)((c - mn) / w * (float)(COST_BINS - 1)) : 0;
    b = b < 0 ? 0 : (b > COST_BINS - 1 ? COST_BINS - 1 : b);
    return COST_BINS - 1 - b;   // descending cost
}

// The items are cut into gridDim.x contiguous chunks.  A block counts its
// chunk's bins in LDS and adds only the non-zero ones to the global
// histogram (one global atomic per item would contend on the bins).
__global__ __launch_bounds__(1024) void k_cost_hist(const float* __restrict__ cost,
                                                    const int32_t* __restrict__ stats, int32_t* sc)
{
    __shared__ int32_t h[COST_BINS];
    const int n = stats[ST_NITEMS];
    if (sc[2]) return;   // written by the previous kernel: the same for every thread
    const float mn = f_val((uint32_t)sc[0]), mx = f_val((uint32_t)sc[1]);
    for (int b = threadIdx.x; b < COST_BINS; b += blockDim.x) h[b] = 0;
    __syncthreads();
    const int per = (n + (int)gridDim.x - 1) / (int)gridDim.x;
    const int i0 = (int)blockIdx.x * per, i1 = min(n, i0 + per);
    for (int i = i0 + (int)threadIdx.x; i < i1; i += blockDim.x)
        atomicAdd(&h[cost_bin(cost[i], mn, mx)], 1);
    __syncthreads();
    for (int b = threadIdx.x; b < COST_BINS; b += blockDim.x)
        if (h[b]) atomicAdd(&sc[16 + b], h[b]);
}

__global__ __launch_bounds__(1024) void k_cost_scan(int32_t* sc)
{
    __shared__ int part[1024];
    int32_t* hist = sc + 16;
    int32_t* off = sc + 16 + COST_BINS;
    constexpr int R = COST_BINS / 1024;
    int v[R], s = 0;
    for (int r = 0; r < R; ++r) {
        v[r] = hist[threadIdx.x * R + r];
        s += v[r];
    }
    part[threadIdx.x] = s;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        int a = threadIdx.x >= o ? part[threadIdx.x - o] : 0;
        __syncthreads();
        part[threadIdx.x] += a;
        __syncthreads();
    }
    int base = part[threadIdx.x] - s;
    for (int r = 0; r < R; ++r) {
        off[threadIdx.x * R + r] = base;
        base += v[r];
    }
    if (threadIdx.x == 1023) off[COST_BINS] = base;
}

// Same chunks as k_cost_hist: a block counts its chunk's bins in LDS,
// reserves one range per non-zero bin in the global order (one atomic per
// bin), and places each item at its range's base plus its rank in the block.
// Items of one bin may land in any order; each item lands exactly once.
__global__ __launch_bounds__(1024) void k_cost_scatter(const float* __restrict__ cost,
                                                       const int32_t* __restrict__ stats, int32_t* sc,
                                                       int32_t* __restrict__ order)
{
    __shared__ int32_t h[COST_BINS];   // the chunk's count, then its base position, per bin
    __shared__ int32_t r[COST_BINS];   // rank of the next item per bin
    const int n = stats[ST_NITEMS];
    const int per = (n + (int)gridDim.x - 1) / (int)gridDim.x;
    const int i0 = (int)blockIdx.x * per, i1 = min(n, i0 + per);
    if (sc[2]) {   // a prediction is not finite: static order
        for (int i = i0 + (int)threadIdx.x; i < i1; i += blockDim.x) order[i] = i;
        return;
    }
    const float mn = f_val((uint32_t)sc[0]), mx = f_val((uint32_t)sc[1]);
    for (int b = threadIdx.x; b < COST_BINS; b += blockDim.x) {
        h[b] = 0;
        r[b] = 0;
    }
    __syncthreads();
    for (int i = i0 + (int)threadIdx.x; i < i1; i += blockDim.x)
        atomicAdd(&h[cost_bin(cost[i], mn, mx)], 1);
    __syncthreads();
    for (int b = threadIdx.x; b < COST_BINS; b += blockDim.x)
        if (h[b]) h[b] = sc[16 + COST_BINS + b] + atomicAdd(&sc[16 + 2 * COST_BINS + 1 + b], h[b]);
    __syncthreads();
    for (int i = i0 + (int)threadIdx.x; i < i1; i += blockDim.x) {
        const int b = cost_bin(cost[i], mn, mx);
        order[h[b] + atomicAdd(&r[b], 1)] = i;
    }
}

static int launch_order(void* ctx, const float* cost, int32_t* order, hipStream_t st)
{
    int32_t* sc = bh_ctx_cost_scratch(ctx);
    const int32_t* stats = bh_ctx_stats(ctx);
    const int64_t mi = bh_ctx_max_items(ctx);
    unsigned g = (unsigned)((mi + 255) / 256);
    hipLaunchKernelGGL(k_cost_init, dim3(4), dim3(1024), 0, st, sc);
    hipLaunchKernelGGL(k_cost_minmax, dim3(g < 512u ? g : 512u), dim3(256), 0, st, cost, stats, sc);
    // COST_CHUNKS blocks of 1024 threads, one per chunk of the items
    // (k_cost_hist and k_cost_scatter must cut the same chunks)
    hipLaunchKernelGGL(k_cost_hist, dim3(COST_CHUNKS), dim3(1024), 0, st, cost, stats, sc);
    hipLaunchKernelGGL(k_cost_scan, dim3(1), dim3(1024), 0, st, sc);
    hipLaunchKernelGGL(k_cost_scatter, dim3(COST_CHUNKS), dim3(1024), 0, st, cost, stats, sc, order);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return bh_ctx_fail(BH_EHIP, hipGetErrorString(e));
    return BH_OK;
}

extern "C" {

int bh_item_features(void* ctx, float* feat, void* stream)
{
    if (!ctx || !feat) return bh_ctx_fail(BH_EINVAL, "NULL argument");
    hipStream_t st = (hipStream_t)stream;
    const int64_t mi = bh_ctx_max_items(ctx);
    hipLaunchKernelGGL(k_item_features, dim3((unsigned)((mi + 255) / 256)), dim3(256), 0, st,
                       (const Item*)bh_ctx_items(ctx), (const int32_t*)bh_ctx_stats(ctx),
                       bh_ctx_sys(ctx), bh_ctx_offsets(ctx), bh_ctx_neprefix(ctx), bh_ctx_l64(ctx),
                       64 * bh_ctx_p(ctx), feat);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return bh_ctx_fail(BH_EHIP, hipGetErrorString(e));
    return BH_OK;
}

int bh_mlp_forward(const float* params, const float* x, int64_t n, float* y, void* stream)
{
    if (!params || !x || !y || n < 0) return bh_ctx_fail(BH_EINVAL, "bad argument");
    if (n == 0) return BH_OK;
    hipLaunchKernelGGL(k_mlp_fwd, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, params, x, n, (const int32_t*)nullptr, y);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return bh_ctx_fail(BH_EHIP, hipGetErrorString(e));
    return BH_OK;
}

int bh_mlp_backward(const float* params, const float* x, const float* target, int64_t n,
                    float* grad, float* loss, void* stream)
{
    if (!params || !x || !target || !grad || n <= 0) return bh_ctx_fail(BH_EINVAL, "bad argument");
    hipStream_t st = (hipStream_t)stream;
    int nb = (int)((n + BWD_ROWS - 1) / BWD_ROWS);
    float* partial = nullptr;
    hipError_t e = hipMalloc((void**)&partial, sizeof(float) * ((size_t)nb * BH_MLP_NPARAM + nb));
    if (e != hipSuccess) return bh_ctx_fail(BH_EHIP, hipGetErrorString(e));
    float* lp = partial + (size_t)nb * BH_MLP_NPARAM;
    hipLaunchKernelGGL(k_mlp_bwd_partial, dim3(nb), dim3(256), 0, st, params, x, target, n,
                       partial, lp);
    hipLaunchKernelGGL(k_mlp_bwd_reduce, dim3((BH_MLP_NPARAM + 255) / 256), dim3(256), 0, st,
                       (const float*)partial, nb, grad, (const float*)lp, loss);
    e = hipGetLastError();
    hipError_t e2 = hipStreamSynchronize(st);
    (void)hipFree(partial);
    if (e != hipSuccess) return bh_ctx_fail(BH_EHIP, hipGetErrorString(e));
    if (e2 != hipSuccess) return bh_ctx_fail(BH_EHIP, hipGetErrorString(e2));
    return BH_OK;
}

int bh_mlp_adam(float* params, const float* grad, float* m, float* v, int32_t t, float lr,
                void* stream)
{
    if (!params || !grad || !m || !v || t < 1) return bh_ctx_fail(BH_EINVAL, "bad argument");
    hipLaunchKernelGGL(k_mlp_adam, dim3((BH_MLP_NPARAM + 255) / 256), dim3(256), 0,
                       (hipStream_t)stream, params, grad, m, v, (int)t, lr);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return bh_ctx_fail(BH_EHIP, hipGetErrorString(e));
    return BH_OK;
}

int bh_order_by_cost(void* ctx, const float* cost, int32_t* order, void* stream)
{
    if (!ctx || !cost || !order) return bh_ctx_fail(BH_EINVAL, "NULL argument");
    return launch_order(ctx, cost, order, (hipStream_t)stream);
}

int bh_learned_order(void* ctx, const float* params, int32_t* order, void* stream)
{
    if (!ctx || !params || !order) return bh_ctx_fail(BH_EINVAL, "NULL argument");
    hipStream_t st = (hipStream_t)stream;
    float* feat = bh_ctx_feat(ctx);
    float* pred = bh_ctx_pred(ctx);
    int r = bh_item_features(ctx, feat, stream);
    if (r) return r;
    const int64_t mi = bh_ctx_max_items(ctx);
    hipLaunchKernelGGL(k_mlp_fwd, dim3((unsigned)((mi + 255) / 256)), dim3(256), 0, st, params,
                       (const float*)feat, mi, (const int32_t*)(bh_ctx_stats(ctx) + ST_NITEMS),
                       pred);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return bh_ctx_fail(BH_EHIP, hipGetErrorString(e));
    return launch_order(ctx, pred, order, st);
}

}  // extern "C"
