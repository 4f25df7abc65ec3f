// Placement phase: bounding box -> cell keys + histogram -> offsets ->
// ordered permutation -> gather.  Replaces python-bh's
//   set_particles_bounding_box   barneshut/implementations/base.py:211-239
//   g_place_particles            barneshut/kernels/grid_decomposition/gpu/grid.py:15-50
//   g_calculate_box_cumm         .../gpu/grid.py:52-60
//   g_sort_particles             .../gpu/grid.py:62-76
// with the ordering of the sequential path (implementations/sequential.py:32-50):
// x-major cell key, and inside a cell numpy's structured-sort tie rule
// (px, py, mass, vx, vy, ax, ay, pid), so the permutation is deterministic.
#include <limits.h>

#include "bh_common.h"
#include "bh_kernels.h"

// ---------------------------------------------------------------- fills
__global__ void k_fill_i32(int32_t* p, int64_t n, int32_t v)
{
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

__global__ void k_fill_f64(double* p, int64_t n, double v)
{
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

// ---------------------------------------------------------------- bbox
__global__ void k_bbox_init(uint64_t* bb, int S)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < S) {
        bb[4 * s + 0] = ~0ull;   // min x (ordered key)
        bb[4 * s + 1] = ~0ull;   // min y
        bb[4 * s + 2] = 0ull;    // max x
        bb[4 * s + 3] = 0ull;    // max y
    }
}

// Atomic min/max of the ordered keys; skipped when the stored value already
// bounds ours (stored minima only fall and maxima only rise, so a stale read
// can only cause an unneeded atomic, never a missed one).
__device__ __forceinline__ void bbox_flush(uint64_t* bb, int s, double mnx, double mny, double mxx,
                                           double mxy)
{
    unsigned long long* b = (unsigned long long*)&bb[4 * s];
    const unsigned long long k0 = bh_ord_key(mnx), k1 = bh_ord_key(mny);
    const unsigned long long k2 = bh_ord_key(mxx), k3 = bh_ord_key(mxy);
    if (k0 < __atomic_load_n(&b[0], __ATOMIC_RELAXED)) atomicMin(&b[0], k0);
    if (k1 < __atomic_load_n(&b[1], __ATOMIC_RELAXED)) atomicMin(&b[1], k1);
    if (k2 > __atomic_load_n(&b[2], __ATOMIC_RELAXED)) atomicMax(&b[2], k2);
    if (k3 > __atomic_load_n(&b[3], __ATOMIC_RELAXED)) atomicMax(&b[3], k3);
}

// Each block covers PLACE_CHUNK consecutive particles; min/max are exact so
// the reduction order does not matter.
__global__ __launch_bounds__(256) void k_bbox(const double* __restrict__ px,
                                              const double* __restrict__ py,
                                              const int32_t* __restrict__ sys_of, int64_t N,
                                              uint64_t* bb, int32_t* stats)
{
    const double inf = __longlong_as_double(0x7ff0000000000000ll);
    int64_t base = (int64_t)blockIdx.x * PLACE_CHUNK;
    int cur = -1;
    double mnx = inf, mny = inf, mxx = -inf, mxy = -inf;
    bool bad = false;
    for (int r = 0; r < PLACE_CHUNK / 256; ++r) {
        int64_t i = base + r * 256 + threadIdx.x;
        if (i >= N) break;
        int s = sys_of[i];
        if (s != cur) {
            if (cur >= 0) bbox_flush(bb, cur, mnx, mny, mxx, mxy);
            cur = s;
            mnx = inf; mny = inf; mxx = -inf; mxy = -inf;
        }
        double x = px[i], y = py[i];
        if (!isfinite(x) || !isfinite(y)) {
            bad = true;
            continue;
        }
        mnx = fmin(mnx, x);
        mxx = fmax(mxx, x);
        mny = fmin(mny, y);
        mxy = fmax(mxy, y);
    }
    // Reduce over the block when all its rows belong to one system (one
    // flush per block), else per wave, else per lane.
    __shared__ double red[4][4];
    __shared__ int wsys[4];
    const int wid = threadIdx.x >> 6;
    int c0 = __shfl(cur, 0, BH_WAVE);
    bool uni = __all(cur == c0);
    if (uni && c0 >= 0) {
        mnx = wave_min(mnx);
        mny = wave_min(mny);
        mxx = wave_max(mxx);
        mxy = wave_max(mxy);
    }
    if ((threadIdx.x & 63) == 0) {
        wsys[wid] = (uni && c0 >= 0) ? c0 : -1;
        red[wid][0] = mnx;
        red[wid][1] = mny;
        red[wid][2] = mxx;
        red[wid][3] = mxy;
    }
    __syncthreads();
    const int nwv = (int)(blockDim.x >> 6);
    bool block_uni = true;
    for (int v = 0; v < nwv; ++v) block_uni = block_uni && wsys[v] >= 0 && wsys[v] == wsys[0];
    if (block_uni) {
        if (threadIdx.x == 0) {
            for (int v = 1; v < nwv; ++v) {
                mnx = fmin(mnx, red[v][0]);
                mny = fmin(mny, red[v][1]);
                mxx = fmax(mxx, red[v][2]);
                mxy = fmax(mxy, red[v][3]);
            }
            bbox_flush(bb, wsys[0], mnx, mny, mxx, mxy);
        }
    } else if (uni && c0 >= 0) {
        if ((threadIdx.x & 63) == 0) bbox_flush(bb, c0, mnx, mny, mxx, mxy);
    } else if (cur >= 0) {
        bbox_flush(bb, cur, mnx, mny, mxx, mxy);
    }
    if (bad) atomicOr(&stats[ST_ERR], BH_ERR_NONFINITE);
}

// helpers/grid.py:15-22 square-up, base.py:237 step, exactly in FP64.
__global__ void k_bbox_finalize(SysDev* sys, const uint64_t* bb, int S, int32_t* stats)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    double mnx = bh_ord_val(bb[4 * s + 0]);
    double mny = bh_ord_val(bb[4 * s + 1]);
    double mxx = bh_ord_val(bb[4 * s + 2]);
    double mxy = bh_ord_val(bb[4 * s + 3]);
    double xe = mxx - mnx, ye = mxy - mny;
    if (xe > ye) {
        mxy += (xe - ye);
    } else if (ye > xe) {
        mxx += (ye - xe);
    }
    double step = (mxx - mnx) / (double)sys[s].G;
    sys[s].min_x = mnx;
    sys[s].min_y = mny;
    sys[s].max_x = mxx;
    sys[s].max_y = mxy;
    sys[s].step = step;
    if (!(step > 0.0) || !isfinite(step)) atomicOr(&stats[ST_ERR], BH_ERR_DEGENERATE);
}

// ---------------------------------------------------------------- keys + histogram
// sequential.py:32-34: g = clip(floor((p - min) / step), 0, G-1), IEEE division.
__device__ __forceinline__ int32_t cell_key(double x, double y, const SysDev& d)
{
    double fx = floor((x - d.min_x) / d.step);
    double fy = floor((y - d.min_y) / d.step);
    double gm = (double)(d.G - 1);
    fx = fx < 0.0 ? 0.0 : (fx > gm ? gm : fx);
    fy = fy < 0.0 ? 0.0 : (fy > gm ? gm : fy);
    int gx = (int)fx, gy = (int)fy;
    // second clamp keeps a NaN coordinate (already flagged) inside the grid
    gx = gx < 0 ? 0 : (gx > d.G - 1 ? d.G - 1 : gx);
    gy = gy < 0 ? 0 : (gy > d.G - 1 ? d.G - 1 : gy);
    return (int32_t)(d.c_off + (int64_t)gx * d.G + gy);
}

__device__ __forceinline__ void block_minmax(int& kmin, int& kmax, int* red)
{
    kmin = wave_min(kmin);
    kmax = wave_max(kmax);
    int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[w] = kmin;
        red[4 + w] = kmax;
    }
    __syncthreads();
    kmin = min(min(red[0], red[1]), min(red[2], red[3]));
    kmax = max(max(red[4], red[5]), max(red[6], red[7]));
    __syncthreads();
}

// Keys for PLACE_CHUNK consecutive particles per block, then an LDS
// histogram over the block's key window when it fits (the input is the
// previous step's sorted order, so a block's keys are nearly contiguous);
// only non-zero bins go to global atomics.
__global__ __launch_bounds__(256) void k_keys_hist(const double* __restrict__ px,
                                                   const double* __restrict__ py,
                                                   const int32_t* __restrict__ sys_of,
                                                   const SysDev* __restrict__ sys, int64_t N,
                                                   int32_t* __restrict__ key_out,
                                                   int32_t* __restrict__ counts)
{
    __shared__ int32_t hist[HIST_BINS];
    __shared__ int red[8];
    constexpr int R = PLACE_CHUNK / 256;
    int64_t base = (int64_t)blockIdx.x * PLACE_CHUNK;
    int kk[R];
    int kmin = INT_MAX, kmax = INT_MIN;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        int64_t i = base + r * 256 + threadIdx.x;
        kk[r] = -1;
        if (i < N) {
            const SysDev& d = sys[sys_of[i]];
            int k = cell_key(px[i], py[i], d);
            key_out[i] = k;
            kk[r] = k;
            kmin = min(kmin, k);
            kmax = max(kmax, k);
        }
    }
    block_minmax(kmin, kmax, red);
    if (kmax < kmin) return;   // empty block
    if ((int64_t)kmax - kmin < HIST_BINS) {
        int span = kmax - kmin + 1;
        for (int b = threadIdx.x; b < span; b += 256) hist[b] = 0;
        __syncthreads();
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (kk[r] >= 0) atomicAdd(&hist[kk[r] - kmin], 1);
        __syncthreads();
        for (int b = threadIdx.x; b < span; b += 256)
            if (hist[b]) atomicAdd(&counts[kmin + b], hist[b]);
    } else {
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (kk[r] >= 0) atomicAdd(&counts[kk[r]], 1);
    }
}

// ---------------------------------------------------------------- scan
// Block-wide exclusive scan of two int sequences (256 or 1024 threads).
__device__ __forceinline__ void block_excl_scan2(int v, int f, int& ev, int& ef, int& tv, int& tf,
                                                 int* sh /* >= 2*16 + 2 ints */)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    int iv = v, jf = f;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int a = __shfl_up(iv, o, BH_WAVE);
        int b = __shfl_up(jf, o, BH_WAVE);
        if (lane >= o) {
            iv += a;
            jf += b;
        }
    }
    if (lane == 63) {
        sh[wid] = iv;
        sh[16 + wid] = jf;
    }
    __syncthreads();
    if (wid == 0) {
        int a = lane < nw ? sh[lane] : 0;
        int b = lane < nw ? sh[16 + lane] : 0;
        int ia = a, ib = b;
#pragma unroll
        for (int o = 1; o < 16; o <<= 1) {
            int x = __shfl_up(ia, o, BH_WAVE);
            int y = __shfl_up(ib, o, BH_WAVE);
            if (lane >= o) {
                ia += x;
                ib += y;
            }
        }
        if (lane < nw) {
            sh[lane] = ia - a;
            sh[16 + lane] = ib - b;
        }
        if (lane == nw - 1) {
            sh[32] = ia;
            sh[33] = ib;
        }
    }
    __syncthreads();
    ev = sh[wid] + iv - v;
    ef = sh[16 + wid] + jf - f;
    tv = sh[32];
    tf = sh[33];
    __syncthreads();
}

// Phase 1: per-tile sums of v = in[c] and f = (in[c] > 0), plus the max.
template <bool CNT>
__global__ __launch_bounds__(256) void k_scan_reduce(const int32_t* __restrict__ in, int64_t n,
                                                     int32_t* bsum, int32_t* bne, int32_t* stats)
{
    __shared__ int sh[40];
    int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * (SCAN_TILE / 256);
    int v = 0, f = 0, mx = 0;
#pragma unroll
    for (int r = 0; r < SCAN_TILE / 256; ++r) {
        int64_t i = base + r;
        if (i < n) {
            int c = in[i];
            v += c;
            f += c > 0;
            mx = max(mx, c);
        }
    }
    int ev, ef, tv, tf;
    block_excl_scan2(v, f, ev, ef, tv, tf, sh);
    if (CNT) {
        mx = wave_max(mx);
        if ((threadIdx.x & 63) == 0) atomicMax(&stats[ST_MAXCOUNT], mx);
    }
    if (threadIdx.x == 0) {
        bsum[blockIdx.x] = tv;
        if (CNT) bne[blockIdx.x] = tf;
    }
}

// Phase 2: one block scans the tile sums in place (exclusive).
template <bool CNT>
__global__ __launch_bounds__(1024) void k_scan_blocks(int32_t* bsum, int32_t* bne, int64_t nb,
                                                      int32_t* stats, int total_slot)
{
    __shared__ int sh[40];
    int cv = 0, cf = 0;
    for (int64_t c0 = 0; c0 < nb; c0 += 1024) {
        int64_t i = c0 + threadIdx.x;
        int v = i < nb ? bsum[i] : 0;
        int f = (CNT && i < nb) ? bne[i] : 0;
        int ev, ef, tv, tf;
        block_excl_scan2(v, f, ev, ef, tv, tf, sh);
        if (i < nb) {
            bsum[i] = cv + ev;
            if (CNT) bne[i] = cf + ef;
        }
        cv += tv;
        cf += tf;
    }
    if (threadIdx.x == 0) {
        if (CNT) stats[ST_NCELLS_NE] = cf;
        if (total_slot >= 0) stats[total_slot] = cv;
    }
}

// Phase 3: exclusive prefix per element; out has n + 1 entries.
template <bool CNT>
__global__ __launch_bounds__(256) void k_scan_final(const int32_t* __restrict__ in, int64_t n,
                                                    const int32_t* bsum, const int32_t* bne,
                                                    int32_t* out_v, int32_t* out_f)
{
    __shared__ int sh[40];
    constexpr int R = SCAN_TILE / 256;
    int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * R;
    int cv[R];
    int v = 0, f = 0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        int64_t i = base + r;
        cv[r] = i < n ? in[i] : 0;
        v += cv[r];
        f += cv[r] > 0;
    }
    int ev, ef, tv, tf;
    block_excl_scan2(v, f, ev, ef, tv, tf, sh);
    int av = bsum[blockIdx.x] + ev;
    int af = CNT ? bne[blockIdx.x] + ef : 0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        int64_t i = base + r;
        if (i < n) {
            out_v[i] = av;
            if (CNT) out_f[i] = af;
            av += cv[r];
            af += cv[r] > 0;
            if (i == n - 1) {
                out_v[n] = av;
                if (CNT) out_f[n] = af;
            }
        }
    }
}

template __global__ void k_scan_reduce<true>(const int32_t*, int64_t, int32_t*, int32_t*, int32_t*);
template __global__ void k_scan_reduce<false>(const int32_t*, int64_t, int32_t*, int32_t*, int32_t*);
template __global__ void k_scan_blocks<true>(int32_t*, int32_t*, int64_t, int32_t*, int);
template __global__ void k_scan_blocks<false>(int32_t*, int32_t*, int64_t, int32_t*, int);
template __global__ void k_scan_final<true>(const int32_t*, int64_t, const int32_t*, const int32_t*,
                                            int32_t*, int32_t*);
template __global__ void k_scan_final<false>(const int32_t*, int64_t, const int32_t*,
                                             const int32_t*, int32_t*, int32_t*);

// 64-bit key ordered as numpy orders doubles: -0.0 equals +0.0, every NaN is
// equal to every other NaN and sorts after all numbers.
__device__ __forceinline__ uint64_t sort_key(double d)
{
    if (isnan(d)) return ~0ull;
    return bh_ord_key(d == 0.0 ? 0.0 : d);
}

// ---------------------------------------------------------------- scatter
// Counting-sort scatter into cell order.  A block reserves one contiguous
// range per distinct key with a single global atomic (LDS-privatised
// cursor); the order inside a cell is fixed afterwards by k_sort_units.
// Each row's px and py sort keys travel with it, so that pass reads them
// contiguously (rows clipped to the box edge tie in px by the thousand).
__global__ __launch_bounds__(256) void k_scatter(const int32_t* __restrict__ key, int64_t N,
                                                 const int32_t* __restrict__ offsets,
                                                 int32_t* __restrict__ cursor,
                                                 int32_t* __restrict__ idx_out,
                                                 int32_t* __restrict__ key_sorted,
                                                 const double* __restrict__ px,
                                                 const double* __restrict__ py,
                                                 uint64_t* __restrict__ pxkey_out,
                                                 uint64_t* __restrict__ pykey_out)
{
    __shared__ int32_t hist[HIST_BINS];
    __shared__ int red[8];
    constexpr int R = PLACE_CHUNK / 256;
    int64_t base = (int64_t)blockIdx.x * PLACE_CHUNK;
    int kk[R];
    int kmin = INT_MAX, kmax = INT_MIN;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        int64_t i = base + r * 256 + threadIdx.x;
        kk[r] = i < N ? key[i] : -1;
        if (kk[r] >= 0) {
            kmin = min(kmin, kk[r]);
            kmax = max(kmax, kk[r]);
        }
    }
    block_minmax(kmin, kmax, red);
    if (kmax < kmin) return;
    if ((int64_t)kmax - kmin < HIST_BINS) {
        int span = kmax - kmin + 1;
        int lr[R];
        for (int b = threadIdx.x; b < span; b += 256) hist[b] = 0;
        __syncthreads();
#pragma unroll
        for (int r = 0; r < R; ++r)
            lr[r] = kk[r] >= 0 ? atomicAdd(&hist[kk[r] - kmin], 1) : 0;
        __syncthreads();
        for (int b = threadIdx.x; b < span; b += 256) {
            int c = hist[b];
            if (c) hist[b] = offsets[kmin + b] + atomicAdd(&cursor[kmin + b], c);
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (kk[r] >= 0) {
                int pos = hist[kk[r] - kmin] + lr[r];
                const int64_t i = base + r * 256 + threadIdx.x;
                idx_out[pos] = (int32_t)i;
                key_sorted[pos] = kk[r];
                pxkey_out[pos] = sort_key(px[i]);
                pykey_out[pos] = sort_key(py[i]);
            }
        }
    } else {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (kk[r] >= 0) {
                int pos = offsets[kk[r]] + atomicAdd(&cursor[kk[r]], 1);
                const int64_t i = base + r * 256 + threadIdx.x;
                idx_out[pos] = (int32_t)i;
                key_sorted[pos] = kk[r];
                pxkey_out[pos] = sort_key(px[i]);
                pykey_out[pos] = sort_key(py[i]);
            }
        }
    }
}

// ---------------------------------------------------------------- within-cell order
// numpy's structured-sort tie rule after (gx, gy): px, py, mass, vx, vy, ax,
// ay, pid (internals/particle.py:3-9); equal rows keep input order (stable).
// Doubles compare as numpy sorts them: NaN after every number, NaNs equal,
// so the order is total and the ranks always form a permutation.
__device__ __forceinline__ int dcmp(double u, double v)
{
    const bool un = isnan(u), vn = isnan(v);
    if (un || vn) return (int)un - (int)vn;
    return (u < v) ? -1 : (u > v ? 1 : 0);
}

__device__ __forceinline__ bool tie_less(const SortFields& F, int a, int b)
{
    int c;
    if ((c = dcmp(F.py[a], F.py[b]))) return c < 0;
    if ((c = dcmp(F.m[a], F.m[b]))) return c < 0;
    if ((c = dcmp(F.vx[a], F.vx[b]))) return c < 0;
    if ((c = dcmp(F.vy[a], F.vy[b]))) return c < 0;
    if ((c = dcmp(F.ax[a], F.ax[b]))) return c < 0;
    if ((c = dcmp(F.ay[a], F.ay[b]))) return c < 0;
    if (F.pid[a] != F.pid[b]) return F.pid[a] < F.pid[b];
    return a < b;
}

__device__ __forceinline__ bool row_less(const SortFields& F, int a, int b)
{
    int c = dcmp(F.px[a], F.px[b]);
    if (c) return c < 0;
    return tie_less(F, a, b);
}

// A "unit" is a whole cell when the cell holds <= SORT_W particles, else one
// SORT_W-long chunk of it (chunks counted from the cell start).  Every unit
// starts in exactly one aligned SORT_W window; the block owning that window
// ranks the unit's rows in LDS (all of them lie in [w0, w0 + 2*SORT_W)) on
// their px keys, then py keys; the remaining fields are read only for rows
// equal in both.
__global__ __launch_bounds__(SORT_THREADS) void k_sort_units(SortFields F, int64_t N,
                                                    const int32_t* __restrict__ idx_in,
                                                    const int32_t* __restrict__ key_sorted,
                                                    const uint64_t* __restrict__ pxkey,
                                                    const uint64_t* __restrict__ pykey,
                                                    const int32_t* __restrict__ offsets,
                                                    int32_t* __restrict__ idx_out)
{
    __shared__ uint64_t lkx[2 * SORT_W];
    __shared__ uint64_t lky[2 * SORT_W];
    __shared__ int64_t lend;
    const int64_t w0 = (int64_t)blockIdx.x * SORT_W;
    if (threadIdx.x == 0) {
        // rows past the window belong to this block only up to the end of the
        // unit that holds the window's last row
        const int64_t ql = min((int64_t)N, w0 + SORT_W) - 1;
        const int k = key_sorted[ql];
        const int64_t a = offsets[k], b = offsets[k + 1];
        int64_t ub = b;
        if (b - a > SORT_W) ub = min(b, a + ((ql - a) / SORT_W) * SORT_W + SORT_W);
        lend = max(ql + 1, ub);
    }
    __syncthreads();
    const int64_t wend = min(lend, w0 + 2 * SORT_W);   // > w0: the window holds a row
    // Each thread's rows t = threadIdx.x + SORT_THREADS r.  Every load below reads a
    // row clamped into [w0, wend), so it is always in bounds and needs no
    // branch: the loads of all rows are issued before the first is waited on
    // (a guarded load per row would wait for each one in turn).
    constexpr int RPT = 2 * SORT_W / SORT_THREADS;
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        const int t = threadIdx.x + SORT_THREADS * r;
        const int64_t q = min(w0 + t, wend - 1);
        // slots past wend get the last row's keys; no rank reads them (every
        // unit this block ranks ends at or before wend)
        lkx[t] = pxkey[q];
        lky[t] = pykey[q];
    }
    __syncthreads();
    // the rows' cells and source rows, then the cells' bounds
    int kr[RPT], er[RPT];
    int64_t ar[RPT], br[RPT];
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        const int64_t q = min(w0 + threadIdx.x + SORT_THREADS * r, wend - 1);
        kr[r] = key_sorted[q];
        er[r] = idx_in[q];
    }
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        ar[r] = offsets[kr[r]];
        br[r] = offsets[kr[r] + 1];
    }
    for (int r = 0; r < RPT; ++r) {
        const int t = threadIdx.x + SORT_THREADS * r;
        const int64_t q = w0 + t;
        if (q >= wend) break;
        const int64_t a = ar[r], b = br[r];
        int64_t ua, ub;
        if (b - a <= SORT_W) {
            ua = a;
            ub = b;
        } else {
            int64_t c = (q - a) / SORT_W;
            ua = a + c * SORT_W;
            ub = min(b, ua + SORT_W);
        }
        if (ua < w0 || ua >= w0 + SORT_W) continue;   // unit owned by another block
        const int e = er[r];
        const uint64_t x = lkx[t], xy = lky[t];
        int rank = 0;
        for (int j = (int)(ua - w0); j < (int)(ub - w0); ++j) {
            const uint64_t y = lkx[j];
            if (y < x) {
                ++rank;
            } else if (y == x) {
                const uint64_t yy = lky[j];
                if (yy < xy) ++rank;
                else if (yy == xy && j != t && tie_less(F, idx_in[w0 + j], e)) ++rank;
            }
        }
        idx_out[ua + rank] = e;
    }
}

// Merge pass for cells longer than SORT_W: runs of length R = SORT_W << L,
// counted from the cell start, are merged pairwise by binary search.  Only
// rows of such cells are touched; the pass is a no-op once R >= max count.
__global__ __launch_bounds__(256) void k_sort_merge(SortFields F, int64_t N,
                                                    const int32_t* __restrict__ idx_in,
                                                    const int32_t* __restrict__ key_sorted,
                                                    const int32_t* __restrict__ offsets,
                                                    int32_t* __restrict__ idx_out, int L,
                                                    const int32_t* __restrict__ stats)
{
    const int64_t R = (int64_t)SORT_W << L;
    if (R >= stats[ST_MAXCOUNT]) return;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < N; q += stride) {
        int k = key_sorted[q];
        int64_t a = offsets[k], b = offsets[k + 1];
        if (b - a <= SORT_W) continue;
        int e = idx_in[q];
        int64_t W0 = a + ((q - a) / (2 * R)) * (2 * R);
        int64_t mid = W0 + R, hi = min(b, W0 + 2 * R);
        int64_t own0, p0, p1;
        if (q < mid) {
            own0 = W0;
            p0 = mid;
            p1 = hi;
        } else {
            own0 = mid;
            p0 = W0;
            p1 = mid;
        }
        if (p0 >= p1) {
            idx_out[q] = e;
            continue;
        }
        int64_t lo2 = p0, hi2 = p1;
        while (lo2 < hi2) {
            int64_t m2 = (lo2 + hi2) >> 1;
            if (row_less(F, idx_in[m2], e)) lo2 = m2 + 1; else hi2 = m2;
        }
        idx_out[W0 + (q - own0) + (lo2 - p0)] = e;
    }
}

// ---------------------------------------------------------------- gather
// dst[q] = src[perm[q]] for every field.  Rows of cells <= SORT_W are final
// in idx_b; longer cells went through npass merge passes b->a->b...
__global__ __launch_bounds__(256) void k_gather(SortFields S, BhParticles D, int64_t N,
                                                const int32_t* __restrict__ idx_a,
                                                const int32_t* __restrict__ idx_b,
                                                const int32_t* __restrict__ key_sorted,
                                                const int32_t* __restrict__ offsets,
                                                const int32_t* __restrict__ stats,
                                                int32_t* __restrict__ perm_out)
{
    int mc = stats[ST_MAXCOUNT];
    int npass = 0;
    while (((int64_t)SORT_W << npass) < mc) ++npass;
    const int32_t* big = (npass & 1) ? idx_a : idx_b;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < N; q += stride) {
        int k = key_sorted[q];
        int e = (offsets[k + 1] - offsets[k] <= SORT_W) ? idx_b[q] : big[q];
        if ((uint32_t)e >= (uint32_t)N) e = 0;   // never taken: the ranks form a permutation
        D.px[q] = S.px[e];
        D.py[q] = S.py[e];
        D.m[q] = S.m[e];
        D.vx[q] = S.vx[e];
        D.vy[q] = S.vy[e];
        D.ax[q] = S.ax[e];
        D.ay[q] = S.ay[e];
        D.pid[q] = S.pid[e];
        if (perm_out) perm_out[q] = e;
    }
}
