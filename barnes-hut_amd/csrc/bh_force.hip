// Force evaluation: work items + persistent force kernels (FP64 and FP32).
//
// Replaces g_evaluate_boxes (python-bh/barneshut/kernels/grid_decomposition/
// gpu/grid.py:116-156, with d_is_neighbor :187-192 and the d_*_grav device
// functions :211-268).  For a particle i in cell (cx, cy):
//   near: every particle j != i of the 3x3 cell block around (cx, cy),
//   far:  the COM of every NON-EMPTY cell outside that block (the empty-cell
//         skip of seq/evaluation.py:86-95),
//   a_i += Gc * sum m_j (x_j - x_i) w(r2),  w(s) = 1/(s + eps2)  [LAW 0]
//                                           w(s) = s > r2cut ? 1/(s + eps2) : 0  [LAW 1]
// and the result is added to the existing acceleration, as the reference
// accumulates in place (gpu/grid.py:227-230).
//
// Work decomposition: a work item is a run of <= 64*P consecutive sorted
// particles of one column cx spanning rows [y0, y1] (y1 - y0 < yspan).  One wave (a 64-thread
// workgroup) takes items from a queue (atomic cursor; optional order array),
// stages the item's neighbour particles and then the system's compact COM
// list through LDS in 64-entry tiles, and every lane evaluates its P
// particles against each staged entry (LDS broadcast reads).  The COM
// entries that fall in some lane's 3x3 block lie in three compact ranges
// (columns cx-1..cx+1, rows y0-1..y1+1); only tiles meeting those ranges
// take the per-lane masked path.
#include "bh_common.h"
#include "bh_kernels.h"

// ---------------------------------------------------------------- items
// One wave per grid column (all systems).  The column's sorted particles
// are cut, in order, into runs of pitem particles (a run may start or end
// inside a cell); a run is closed early when its next particle lies yspan or
// more rows below its first one, and at the end of the column.  Pass 1
// (WRITE = false) counts items per column; pass 2 writes them at the scanned
// column offsets.
template <bool WRITE>
__global__ __launch_bounds__(64) void k_items(const SysDev* __restrict__ sys, int S, int64_t ncols,
                                              const int32_t* __restrict__ counts,
                                              const int32_t* __restrict__ offsets, int pitem,
                                              int yspan, int32_t* __restrict__ col_items,
                                              Item* __restrict__ items, int64_t max_items,
                                              int32_t* stats)
{
    // One wave per column: the column's counts and offsets are read 64 rows at
    // a time, and the greedy cut walks the non-empty rows with wave-uniform
    // values (every lane holds the same state; lane 0 writes).
    const int64_t col = blockIdx.x;
    if (col >= ncols) return;
    const int lane = threadIdx.x;
    const int s = sys_of_col(sys, S, col);
    const SysDev& d = sys[s];
    const int G = d.G;
    const int x = (int)(col - d.col_off);
    const int64_t cb = d.c_off + (int64_t)x * G;
    int64_t out = WRITE ? col_items[col] : 0;
    int n_out = 0;
    bool open = false;
    int p0 = 0, np = 0, y0 = 0, y1 = 0;
    auto emit = [&]() {
        if (WRITE && lane == 0) {
            int64_t k = out + n_out;
            if (k < max_items) {
                Item it;
                it.p0 = p0;
                it.np = np;
                it.sys = s;
                it.cx = x;
                it.y0 = y0;
                it.y1 = y1;
                it.pad0 = 0;
                it.pad1 = 0;
                items[k] = it;
            } else {
                atomicOr(&stats[ST_ERR], BH_ERR_ITEMS);
            }
        }
        ++n_out;
        open = false;
    };
    for (int yb = 0; yb < G; yb += BH_WAVE) {
        const int yl = yb + lane;
        const int ncl = yl < G ? counts[cb + yl] : 0;
        const int pcl = yl < G ? offsets[cb + yl] : 0;
        unsigned long long nz = __ballot(ncl > 0);
        while (nz) {
            const int l = __ffsll((long long)nz) - 1;
            nz &= nz - 1;
            const int yy = yb + l;
            int nc = __shfl(ncl, l, BH_WAVE);
            int pc = __shfl(pcl, l, BH_WAVE);
            if (open && yy - y0 >= yspan) emit();
            while (nc > 0) {
                if (!open) {
                    open = true;
                    p0 = pc;
                    np = 0;
                    y0 = yy;
                }
                int take = min(nc, pitem - np);
                np += take;
                pc += take;
                nc -= take;
                y1 = yy;
                if (np == pitem) emit();
            }
        }
    }
    if (open) emit();
    if (!WRITE && lane == 0) col_items[col] = n_out;
}

template __global__ void k_items<false>(const SysDev*, int, int64_t, const int32_t*, const int32_t*,
                                        int, int, int32_t*, Item*, int64_t, int32_t*);
template __global__ void k_items<true>(const SysDev*, int, int64_t, const int32_t*, const int32_t*,
                                       int, int, int32_t*, Item*, int64_t, int32_t*);

// Keeps the item count within the allocated capacity (flags the overflow).
__global__ void k_clamp_items(int32_t* stats, int64_t max_items)
{
    if (threadIdx.x == 0 && blockIdx.x == 0 && stats[ST_NITEMS] > max_items) {
        stats[ST_NITEMS] = (int32_t)max_items;
        stats[ST_ERR] |= BH_ERR_ITEMS;
    }
}

// ---------------------------------------------------------------- common pieces
struct NearE {
    double x, y, m;
    int32_t row, idx;
};

__device__ __forceinline__ void kahan_add(float& s, float& c, float v)
{
    float y = v - c;
    float t = s + y;
    c = (t - s) - y;
    s = t;
}

__device__ __forceinline__ int next_item(int32_t* stats, int lane)
{
    int q = 0;
    if (lane == 0) q = atomicAdd(&stats[ST_QUEUE], 1);
    q = __shfl(q, 0, BH_WAVE);
    return uniform_int(q);
}

// Queue positions [begin, end) of launch `part` of `nparts`: contiguous
// ranges of the `total` positions (the launches of one pass run one after
// the other on one stream; each particle's sums do not depend on the split).
__device__ __forceinline__ void part_range(int total, int part, int nparts, int& begin, int& end)
{
    begin = (int)(((int64_t)total * part) / nparts);
    end = (int)(((int64_t)total * (part + 1)) / nparts);
}

__device__ __forceinline__ bool meets(int b0, int b1, int lo, int hi)
{
    return lo < hi && b0 < hi && b1 > lo;
}

// M w(s2) of a far entry from its reciprocal mass rM = 1 / M (+inf for the
// padding and zero-mass cells, whose terms are then exactly 0).
template <int LAW>
__device__ __forceinline__ float far_t32(float wx, float wy, float rM, float eps2c, float r2c)
{
    if (LAW == 0) return __builtin_amdgcn_rcpf(fmaf(wx, wx, fmaf(wy, wy, eps2c)) * rM);
    const float s2 = fmaf(wx, wx, wy * wy);
    return (s2 > r2c) ? __builtin_amdgcn_rcpf((s2 + eps2c) * rM) : 0.0f;
}

// ---------------------------------------------------------------- FP64
// Far-field weights M0 / q0 and M1 / q1 of two COM entries for one particle
// from one reciprocal: Q = q0 q1, R = 1/Q from v_rcp_f64 refined as in
// wdiv64, then M0 / q0 = (M0 R) q1 and M1 / q1 = (M1 R) q0.  That is the VALU
// count of two separate refined weights with half the reciprocals (v_rcp_f64
// takes four issue cycles of its own), at four roundings per weight instead
// of two.  Needs every kept q >= 1, which the FP64 kernel's far entries have in
// units of 1 / sig (their cell is at least one cell away); an entry the
// particle does not keep (k0 / k1 false: inside its 3x3 block) enters the pair
// as q = 1 and gets weight 0, so the other weight is the same as in any pair
// with it.
template <int LAW>
__device__ __forceinline__ void pair_w64(double wx0, double wy0, double M0, bool k0, double wx1,
                                         double wy1, double M1, bool k1, double eps2,
                                         double r2cut, double& w0, double& w1)
{
    const double s0 = fma(wx0, wx0, wy0 * wy0), s1 = fma(wx1, wx1, wy1 * wy1);
    double q0 = (LAW == 0) ? fma(wx0, wx0, fma(wy0, wy0, eps2)) : s0 + eps2;
    double q1 = (LAW == 0) ? fma(wx1, wx1, fma(wy1, wy1, eps2)) : s1 + eps2;
    q0 = k0 ? q0 : 1.0;
    q1 = k1 ? q1 : 1.0;
    const double Q = q0 * q1;
    const double R0 = __builtin_amdgcn_rcp(Q);
    const double e = fma(-Q, R0, 1.0);
    const double R = fma(R0, fma(e, e, e), R0);
    w0 = (M0 * R) * q1;
    w1 = (M1 * R) * q0;
    if (LAW == 1) {
        w0 = s0 > r2cut ? w0 : 0.0;
        w1 = s1 > r2cut ? w1 : 0.0;
    }
    w0 = k0 ? w0 : 0.0;
    w1 = k1 ? w1 : 0.0;
}

// FAST 1: far-field weights in pairs (pair_w64), near-field weights by
// wdiv64, positions in units of a power of two near the cell size; FAST 0:
// IEEE division per weight, physical units (fp64_ieee).
template <int P, int LAW, int FAST>
__global__ __launch_bounds__(64) void k_force64(ForceArgs A)
{
    __shared__ Com64 ftile[BH_WAVE];
    __shared__ NearE ntile[BH_WAVE];
    const int lane = threadIdx.x;
    const int nitems = A.stats[ST_NITEMS];
    int qb, qe;
    part_range(nitems, A.qpart, A.nqparts, qb, qe);
    for (;;) {
        const int q = qb + next_item(A.stats, lane);
        if (q >= qe) break;
        const long long t0 = A.cost_log ? (long long)wall_clock64() : 0;
        const int itx = A.order ? uniform_int(A.order[q]) : q;
        if ((unsigned)itx >= (unsigned)nitems) continue;   // malformed order entry
        const Item it = A.items[itx];
        const SysDev& d = A.sys[it.sys];
        const int G = d.G;
        const int64_t coff = d.c_off;
        const int cx = it.cx;
        const int ylo = max(it.y0 - 1, 0), yhi = min(it.y1 + 1, G - 1);
        // FAST: positions in units of 1 / sig, sig a power of two with
        // step * sig in [1, 2).  The scaling is exact, so every difference and
        // product rounds as in physical units, and a far-field q is >= 1.
        double sig = 1.0;
        if (FAST && d.step > 0.0 && d.step < __builtin_huge_val()) {
            int e2;
            (void)frexp(d.step, &e2);
            sig = ldexp(1.0, 1 - e2);
        }
        const double eps2 = A.eps2 * sig * sig, r2cut = A.r2cut * sig * sig;
        // past the list: zero mass far away (the term is exactly 0).  In units of
        // 1 / sig a padding q is ~2e120, so a pair of two padding entries keeps
        // Q ~ 4e240 finite (an infinite Q would give R = NaN through
        // fma(-inf, 0, 1)), and particles up to 1e59 cells from the origin stay
        // far from it
        const double pad = FAST ? 1e60 : 1e150;

        double xi[P], yi[P], sx[P], sy[P];
        int cy[P], gi[P];
        bool act[P];
#pragma unroll
        for (int k = 0; k < P; ++k) {
            int slot = lane + BH_WAVE * k;
            act[k] = slot < it.np;
            gi[k] = it.p0 + slot;
            xi[k] = 0.0;
            yi[k] = 0.0;
            cy[k] = -100000;
            sx[k] = 0.0;
            sy[k] = 0.0;
            if (act[k]) {
                xi[k] = A.x[gi[k]] * sig;
                yi[k] = A.y[gi[k]] * sig;
                cy[k] = (int)(A.key_sorted[gi[k]] - coff) - cx * G;
            }
        }

        // near field: particles of columns cx-1..cx+1, rows ylo..yhi
        for (int dx = -1; dx <= 1; ++dx) {
            const int xc = cx + dx;
            if (xc < 0 || xc >= G) continue;
            const int64_t cl = coff + (int64_t)xc * G;
            const int pa = A.offsets[cl + ylo], pb = A.offsets[cl + yhi + 1];
            for (int base = pa; base < pb; base += BH_WAVE) {
                const int j = base + lane;
                NearE e;
                if (j < pb) {
                    e.x = A.x[j] * sig;
                    e.y = A.y[j] * sig;
                    e.m = A.m[j];
                    e.row = (int)(A.key_sorted[j] - cl);
                    e.idx = j;
                } else {
                    e.x = 0.0;
                    e.y = 0.0;
                    e.m = 0.0;
                    e.row = -100000;
                    e.idx = -1;
                }
                ntile[lane] = e;
                __syncthreads();
                const int cnt = min(BH_WAVE, pb - base);
                for (int jj = 0; jj < cnt; ++jj) {
                    const NearE f = ntile[jj];
#pragma unroll
                    for (int k = 0; k < P; ++k) {
                        const bool keep = (f.idx != gi[k]) && (abs(f.row - cy[k]) <= 1);
                        const double wx = f.x - xi[k], wy = f.y - yi[k];
                        double w = near_w64<LAW, FAST>(wx, wy, f.m, eps2, r2cut);
                        w = keep ? w : 0.0;
                        sx[k] = fma(wx, w, sx[k]);
                        sy[k] = fma(wy, w, sy[k]);
                    }
                }
                __syncthreads();
            }
        }

        // far field: compact COM list of the system, minus the 3x3 blocks
        const int ne0 = A.neprefix[coff], ne1 = A.neprefix[coff + (int64_t)G * G];
        int mlo[3], mhi[3];
#pragma unroll
        for (int t = 0; t < 3; ++t) {
            const int xc = cx + t - 1;
            if (xc >= 0 && xc < G) {
                const int64_t cl = coff + (int64_t)xc * G;
                mlo[t] = A.neprefix[cl + ylo];
                mhi[t] = A.neprefix[cl + yhi + 1];
            } else {
                mlo[t] = 0;
                mhi[t] = 0;
            }
        }
        auto stage = [&](int e) {
            Com64 c;
            if (e < ne1) {
                c = A.l64[e];
                c.X *= sig;
                c.Y *= sig;
            } else {
                c.X = pad; c.Y = pad; c.M = 0.0; c.kx = -100000; c.ky = -100000;
            }
            return c;
        };
        Com64 nxt = stage(ne0 + lane);
        for (int base = ne0; base < ne1; base += BH_WAVE) {
            ftile[lane] = nxt;
            __syncthreads();
            nxt = stage(base + BH_WAVE + lane);
            const int b1 = base + BH_WAVE;
            const bool masked = meets(base, b1, mlo[0], mhi[0]) || meets(base, b1, mlo[1], mhi[1]) ||
                                meets(base, b1, mlo[2], mhi[2]);
            if (FAST) {
                // entries in pairs (fixed by their index, so every particle
                // gets the same pairs under any launch shape) sharing one
                // reciprocal; masked tiles give an entry a particle does not
                // keep q = 1 inside the pair, so the kept one rounds exactly
                // as in an unmasked tile
                if (!masked) {
#pragma unroll 2
                    for (int j = 0; j < BH_WAVE; j += 2) {
                        const Com64 f0 = ftile[j], f1 = ftile[j + 1];
#pragma unroll
                        for (int k = 0; k < P; ++k) {
                            const double wx0 = f0.X - xi[k], wy0 = f0.Y - yi[k];
                            const double wx1 = f1.X - xi[k], wy1 = f1.Y - yi[k];
                            double w0, w1;
                            pair_w64<LAW>(wx0, wy0, f0.M, true, wx1, wy1, f1.M, true, eps2, r2cut,
                                          w0, w1);
                            sx[k] = fma(wx0, w0, sx[k]);
                            sy[k] = fma(wy0, w0, sy[k]);
                            sx[k] = fma(wx1, w1, sx[k]);
                            sy[k] = fma(wy1, w1, sy[k]);
                        }
                    }
                } else {
                    for (int j = 0; j < BH_WAVE; j += 2) {
                        const Com64 f0 = ftile[j], f1 = ftile[j + 1];
                        const bool cn0 = abs(f0.kx - cx) <= 1, cn1 = abs(f1.kx - cx) <= 1;
#pragma unroll
                        for (int k = 0; k < P; ++k) {
                            const bool k0 = !(cn0 && abs(f0.ky - cy[k]) <= 1);
                            const bool k1 = !(cn1 && abs(f1.ky - cy[k]) <= 1);
                            const double wx0 = f0.X - xi[k], wy0 = f0.Y - yi[k];
                            const double wx1 = f1.X - xi[k], wy1 = f1.Y - yi[k];
                            double w0, w1;
                            pair_w64<LAW>(wx0, wy0, f0.M, k0, wx1, wy1, f1.M, k1, eps2, r2cut, w0,
                                          w1);
                            sx[k] = fma(wx0, w0, sx[k]);
                            sy[k] = fma(wy0, w0, sy[k]);
                            sx[k] = fma(wx1, w1, sx[k]);
                            sy[k] = fma(wy1, w1, sy[k]);
                        }
                    }
                }
            } else if (!masked) {
#pragma unroll 4
                for (int j = 0; j < BH_WAVE; ++j) {
                    const Com64 f = ftile[j];
#pragma unroll
                    for (int k = 0; k < P; ++k) {
                        const double wx = f.X - xi[k], wy = f.Y - yi[k];
                        const double w = far_w64<LAW, 0>(wx, wy, f.M, eps2, r2cut);
                        sx[k] = fma(wx, w, sx[k]);
                        sy[k] = fma(wy, w, sy[k]);
                    }
                }
            } else {
                for (int j = 0; j < BH_WAVE; ++j) {
                    const Com64 f = ftile[j];
                    const bool colnear = abs(f.kx - cx) <= 1;
#pragma unroll
                    for (int k = 0; k < P; ++k) {
                        const bool keep = !(colnear && abs(f.ky - cy[k]) <= 1);
                        const double wx = f.X - xi[k], wy = f.Y - yi[k];
                        double w = far_w64<LAW, 0>(wx, wy, f.M, eps2, r2cut);
                        w = keep ? w : 0.0;
                        sx[k] = fma(wx, w, sx[k]);
                        sy[k] = fma(wy, w, sy[k]);
                    }
                }
            }
            __syncthreads();
        }

#pragma unroll
        for (int k = 0; k < P; ++k) {
            if (act[k]) {
                A.ax[gi[k]] += A.Gc * (sx[k] * sig);
                A.ay[gi[k]] += A.Gc * (sy[k] * sig);
            }
        }
        if (A.cost_log && lane == 0) A.cost_log[itx] = (long long)wall_clock64() - t0;
    }
}

// ---------------------------------------------------------------- FP32
// Positions in cell units relative to the item origin (cx, oy), oy = the
// item's middle row: the integer cell offsets are exact in FP32, so a far
// term's separation carries a relative error of a few ulp.  Near-field
// separations are formed in FP64 and rounded once.  KAHAN: 0 plain FP32
// sums; 1 FP32 partial per 64-entry tile, tiles Kahan-summed; 2 Kahan per term.
//
// Far field without a cutoff (LAW 0) in mass-scaled form: with c = s / M per
// COM entry (s a power of two per item),
//     t = c (a - u),   M (a - u) / (|a - u|^2 + eps2) = s * t / (|t|^2 + c^2 eps2),
// so the tile holds (c a_x, c a_y, -c, c^2 eps2) and each term costs two fma
// for t, two for |t|^2, v_rcp_f32 and two fma into the sums: one VALU
// instruction fewer than forming M * w.  Scaling by the power of two s is
// exact, so s only keeps c^2 |a - u|^2 inside the FP32 range (s = 2^e with
// M <= 2^e < 2 M for the item's own cell; any power of two gives the same
// sums bit for bit); near-field terms are divided by s (exactly) to share the
// sums, and the sums are multiplied by s at the end.  c = s * rM from the
// list's reciprocal mass equals the correctly rounded s / M (a power-of-two
// scaling commutes with rounding), so staging a tile takes no division.
// Zero-mass entries and the padding past the list hold (1e30, 1e30, 0, 0):
// |t|^2 overflows to inf, v_rcp_f32 gives 0 and the term adds exactly 0.
// Minimum resident waves per SIMD the FP32 kernel is compiled for (bounds its
// VGPRs so that more waves hide the reciprocal's latency): 5 for 4 particles
// per lane, 8 for 1 or 2 (one fewer with per-term Kahan sums, which hold more
// registers).  At 2 particles per lane the 8-wave build keeps a few per-item
// values in scratch and still runs 1.4 % faster than the 7-wave one.
template <int P, int KAHAN>
struct F32Waves {
    static constexpr int value = (P >= 4 ? 5 : 8) - (KAHAN == 2 ? 1 : 0);
};

template <int P, int LAW, int KAHAN>
__global__ __launch_bounds__(64, (F32Waves<P, KAHAN>::value)) void k_force32(ForceArgs A)
{
    constexpr int FT = 2 * BH_WAVE;            // far-field tile: two entries staged per lane
    __shared__ float4 ftile[FT];
    __shared__ int32_t ktile[FT];              // kx << 16 | ky (LAW 0 masked tiles)
    __shared__ NearE ntile[BH_WAVE];
    const int lane = threadIdx.x;
    const int nitems = A.stats[ST_NITEMS];
    const int S = A.split > 1 ? A.split : 1;
    const int nt = S > 1 ? min(A.n_tail, nitems) : 0;
    const int nu = nitems - nt;
    const int total = nu + nt * S;
    int qb, qe;
    part_range(total, A.qpart, A.nqparts, qb, qe);
    for (;;) {
        const int q = qb + next_item(A.stats, lane);
        if (q >= qe) break;
        int pos = q, part = -1;
        if (q >= nu) {   // slice-major: every tail item's slice 0, then slice 1, ...
            const int r = q - nu;
            part = r / nt;
            pos = nu + (r - part * nt);
        }
        const long long t0 = A.cost_log ? (long long)wall_clock64() : 0;
        const int itx = A.order ? uniform_int(A.order[pos]) : pos;
        if ((unsigned)itx >= (unsigned)nitems) continue;   // malformed order entry
        const Item it = A.items[itx];
        const SysDev& d = A.sys[it.sys];
        const int G = d.G;
        const int64_t coff = d.c_off;
        const int cx = it.cx;
        const int oy = (it.y0 + it.y1) >> 1;
        const int ylo = max(it.y0 - 1, 0), yhi = min(it.y1 + 1, G - 1);
        const double step = d.step;
        const double inv_step = 1.0 / step;
        const float eps2c = (float)(A.eps2 * inv_step * inv_step);
        const float r2c = (float)(A.r2cut * inv_step * inv_step);
        // mass scale s (LAW 0): 1 / (the power of two at or below 1 / M of the
        // item's own cell), i.e. M <= s < 2 M
        float msc = 1.0f;
        if (LAW == 0) {
            const int64_t own = coff + (int64_t)cx * G + it.y0;
            const float r0 = fabsf(A.l32[A.neprefix[own]].rM);
            const float p2 = __int_as_float(__float_as_int(r0) & 0x7f800000);
            if (p2 >= 1.0e-37f && p2 <= 1.0e37f) msc = 1.0f / p2;
        }
        const float inv_msc = 1.0f / msc;

        double xi[P], yi[P];
        float ux[P], uy[P], Sx[P], Sy[P], Cx[P], Cy[P], tx[P], ty[P], s2n[P];
        int cy[P], gi[P];
        bool act[P];
#pragma unroll
        for (int k = 0; k < P; ++k) {
            int slot = lane + BH_WAVE * k;
            act[k] = slot < it.np;
            gi[k] = it.p0 + slot;
            xi[k] = 0.0;
            yi[k] = 0.0;
            ux[k] = 0.0f;
            uy[k] = 0.0f;
            cy[k] = -100000;
            if (act[k]) {
                xi[k] = A.x[gi[k]];
                yi[k] = A.y[gi[k]];
                cy[k] = (int)(A.key_sorted[gi[k]] - coff) - cx * G;
                ux[k] = (float)((xi[k] - d.min_x) / step - (double)cx);
                uy[k] = (float)((yi[k] - d.min_y) / step - (double)oy);
            }
            Sx[k] = Sy[k] = Cx[k] = Cy[k] = tx[k] = ty[k] = s2n[k] = 0.0f;
        }

        // near field (separations in FP64, rounded once to FP32); part 0 of a split item.
        // s2n sums fm * t = fm^2 / (s2 + eps2c) >= |term|^2 for the recompute's estimate.
        for (int dx = -1; dx <= (part <= 0 ? 1 : -2); ++dx) {
            const int xc = cx + dx;
            if (xc < 0 || xc >= G) continue;
            const int64_t cl = coff + (int64_t)xc * G;
            const int pa = A.offsets[cl + ylo], pb = A.offsets[cl + yhi + 1];
            for (int base = pa; base < pb; base += BH_WAVE) {
                const int j = base + lane;
                NearE e;
                if (j < pb) {
                    e.x = A.x[j];
                    e.y = A.y[j];
                    e.m = A.m[j];
                    e.row = (int)(A.key_sorted[j] - cl);
                    e.idx = j;
                } else {
                    e.x = 0.0;
                    e.y = 0.0;
                    e.m = 0.0;
                    e.row = -100000;
                    e.idx = -1;
                }
                ntile[lane] = e;
                __syncthreads();
                const int cnt = min(BH_WAVE, pb - base);
                for (int jj = 0; jj < cnt; ++jj) {
                    const NearE f = ntile[jj];
                    const float fm = (float)f.m * inv_msc;
#pragma unroll
                    for (int k = 0; k < P; ++k) {
                        const bool keep = (f.idx != gi[k]) && (abs(f.row - cy[k]) <= 1);
                        const float wx = (float)((f.x - xi[k]) * inv_step);
                        const float wy = (float)((f.y - yi[k]) * inv_step);
                        const float s2 = fmaf(wx, wx, wy * wy);
                        float t = fm * __builtin_amdgcn_rcpf(s2 + eps2c);
                        if (LAW == 1) t = (s2 > r2c) ? t : 0.0f;
                        t = keep ? t : 0.0f;
                        s2n[k] = fmaf(fm, t, s2n[k]);
                        if (KAHAN == 2) {
                            kahan_add(Sx[k], Cx[k], wx * t);
                            kahan_add(Sy[k], Cy[k], wy * t);
                        } else {
                            tx[k] = fmaf(wx, t, tx[k]);
                            ty[k] = fmaf(wy, t, ty[k]);
                        }
                    }
                }
                if (KAHAN == 1) {
#pragma unroll
                    for (int k = 0; k < P; ++k) {
                        kahan_add(Sx[k], Cx[k], tx[k]);
                        kahan_add(Sy[k], Cy[k], ty[k]);
                        tx[k] = 0.0f;
                        ty[k] = 0.0f;
                    }
                }
                __syncthreads();
            }
        }

        if (A.rc && part <= 0) {
#pragma unroll
            for (int k = 0; k < P; ++k)
                if (act[k]) A.ay[A.part_n + gi[k]] = (double)s2n[k] * ((double)msc * (double)msc);
        }

        // far field (a split item's part covers one 64-aligned slice of the list)
        int ne0 = A.neprefix[coff], ne1 = A.neprefix[coff + (int64_t)G * G];
        if (part >= 0) {
            const int chunk = (((ne1 - ne0 + S - 1) / S) + BH_WAVE - 1) & ~(BH_WAVE - 1);
            ne0 = ne0 + part * chunk;
            ne1 = min(ne1, ne0 + chunk);
        }
        int mlo[3], mhi[3];
#pragma unroll
        for (int t = 0; t < 3; ++t) {
            const int xc = cx + t - 1;
            if (xc >= 0 && xc < G) {
                const int64_t cl = coff + (int64_t)xc * G;
                mlo[t] = A.neprefix[cl + ylo];
                mhi[t] = A.neprefix[cl + yhi + 1];
            } else {
                mlo[t] = 0;
                mhi[t] = 0;
            }
        }
        Com32 nxt[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int e = ne0 + lane + h * BH_WAVE;
            if (e < ne1) nxt[h] = A.l32[e];
            else { nxt[h].fx = 1e18f; nxt[h].fy = 1e18f; nxt[h].rM = __builtin_huge_valf(); nxt[h].kxy = 0x7fff7fff; }
        }
        for (int base = ne0; base < ne1; base += FT) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const Com32 en = nxt[h];
                const int kx = en.kxy >> 16, ky = en.kxy & 0xffff;
                const float ex = (float)(kx - cx) + en.fx;
                const float ey = (float)(ky - oy) + en.fy;
                float4 v;
                if (LAW == 0) {
                    if (fabsf(en.rM) != __builtin_huge_valf()) {   // M != 0 (NaN passes)
                        const float c = msc * en.rM;
                        v.x = ex * c;
                        v.y = ey * c;
                        v.z = -c;
                        v.w = (c * c) * eps2c;
                    } else {
                        v.x = 1e30f;
                        v.y = 1e30f;
                        v.z = 0.0f;
                        v.w = 0.0f;
                    }
                    ktile[lane + h * BH_WAVE] = en.kxy;
                } else {
                    v.x = ex;
                    v.y = ey;
                    v.z = en.rM;
                    v.w = __int_as_float(en.kxy);
                }
                ftile[lane + h * BH_WAVE] = v;
            }
            __syncthreads();
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int e = base + FT + lane + h * BH_WAVE;
                if (e < ne1) nxt[h] = A.l32[e];
                else { nxt[h].fx = 1e18f; nxt[h].fy = 1e18f; nxt[h].rM = __builtin_huge_valf(); nxt[h].kxy = 0x7fff7fff; }
            }
            const int b1 = base + FT;
            const bool masked = meets(base, b1, mlo[0], mhi[0]) || meets(base, b1, mlo[1], mhi[1]) ||
                                meets(base, b1, mlo[2], mhi[2]);
            if (!masked) {
#pragma unroll 8
                for (int j = 0; j < FT; ++j) {
                    const float4 f = ftile[j];
#pragma unroll
                    for (int k = 0; k < P; ++k) {
                        float wx, wy, t;
                        if (LAW == 0) {
                            wx = fmaf(ux[k], f.z, f.x);
                            wy = fmaf(uy[k], f.z, f.y);
                            t = __builtin_amdgcn_rcpf(fmaf(wx, wx, fmaf(wy, wy, f.w)));
                        } else {
                            wx = f.x - ux[k];
                            wy = f.y - uy[k];
                            t = far_t32<LAW>(wx, wy, f.z, eps2c, r2c);
                        }
                        if (KAHAN == 2) {
                            kahan_add(Sx[k], Cx[k], wx * t);
                            kahan_add(Sy[k], Cy[k], wy * t);
                        } else {
                            tx[k] = fmaf(wx, t, tx[k]);
                            ty[k] = fmaf(wy, t, ty[k]);
                        }
                    }
                }
            } else {
                for (int j = 0; j < FT; ++j) {
                    const float4 f = ftile[j];
                    const int kxy = LAW == 0 ? ktile[j] : __float_as_int(f.w);
                    const bool colnear = abs((kxy >> 16) - cx) <= 1;
                    const int ky = kxy & 0xffff;
#pragma unroll
                    for (int k = 0; k < P; ++k) {
                        const bool keep = !(colnear && abs(ky - cy[k]) <= 1);
                        float wx, wy, t;
                        if (LAW == 0) {
                            wx = fmaf(ux[k], f.z, f.x);
                            wy = fmaf(uy[k], f.z, f.y);
                            t = __builtin_amdgcn_rcpf(fmaf(wx, wx, fmaf(wy, wy, f.w)));
                        } else {
                            wx = f.x - ux[k];
                            wy = f.y - uy[k];
                            t = far_t32<LAW>(wx, wy, f.z, eps2c, r2c);
                        }
                        t = keep ? t : 0.0f;
                        if (KAHAN == 2) {
                            kahan_add(Sx[k], Cx[k], wx * t);
                            kahan_add(Sy[k], Cy[k], wy * t);
                        } else {
                            tx[k] = fmaf(wx, t, tx[k]);
                            ty[k] = fmaf(wy, t, ty[k]);
                        }
                    }
                }
            }
            if (KAHAN == 1) {
#pragma unroll
                for (int k = 0; k < P; ++k) {
                    kahan_add(Sx[k], Cx[k], tx[k]);
                    kahan_add(Sy[k], Cy[k], ty[k]);
                    tx[k] = 0.0f;
                    ty[k] = 0.0f;
                }
            }
            __syncthreads();
        }

        const double scale = A.Gc * inv_step * (double)msc;
#pragma unroll
        for (int k = 0; k < P; ++k) {
            if (act[k]) {
                double rx, ry;
                if (KAHAN == 0) {
                    rx = (double)tx[k];
                    ry = (double)ty[k];
                } else {
                    rx = (double)Sx[k] - (double)Cx[k];
                    ry = (double)Sy[k] - (double)Cy[k];
                }
                if (part < 0) {
                    A.ax[gi[k]] += scale * rx;
                    A.ay[gi[k]] += scale * ry;
                } else {
                    const int64_t o = ((int64_t)part * A.part_n + gi[k]) * 2;
                    A.part[o] = scale * rx;
                    A.part[o + 1] = scale * ry;
                }
            }
        }
        if (A.cost_log && lane == 0 && part < 0) A.cost_log[itx] = (long long)wall_clock64() - t0;
    }
}

// Adds the parts of the split tail items, in part order (deterministic).
__global__ __launch_bounds__(256) void k_force_combine(const Item* __restrict__ items,
                                                       const int32_t* __restrict__ order,
                                                       const int32_t* __restrict__ stats,
                                                       int split, int n_tail, int pslots,
                                                       const double* __restrict__ part,
                                                       int64_t part_n,
                                                       double* __restrict__ ax,
                                                       double* __restrict__ ay)
{
    const int nitems = stats[ST_NITEMS];
    const int nt = min(n_tail, nitems);
    const int nu = nitems - nt;
    const int64_t total = (int64_t)nt * pslots;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += stride) {
        const int j = (int)(g / pslots), slot = (int)(g - (int64_t)j * pslots);
        const int pos = nu + j;
        const int itx = order ? order[pos] : pos;
        if ((unsigned)itx >= (unsigned)nitems) continue;
        const Item it = items[itx];
        if (slot >= it.np) continue;
        const int64_t gi = (int64_t)it.p0 + slot;
        double sx = 0.0, sy = 0.0;
        for (int s = 0; s < split; ++s) {
            const int64_t o = ((int64_t)s * part_n + gi) * 2;
            sx += part[o];
            sy += part[o + 1];
        }
        ax[gi] += sx;
        ay[gi] += sy;
    }
}

// ---------------------------------------------------------------- FP32 + matrix cores
// Same decomposition and coordinates as k_force32, with the far-field sum
// split as  sum_k M_k r_k (a_k - u) = sum_k r_k (M_k a_k) - u sum_k r_k M_k,
// r_k = 1/(|a_k - u|^2 + eps2).  Per COM entry the tile holds
// B_k = (M a_x, M a_y, M, 0); every lane computes r_k for its particle (2 sub,
// r^2 as mul+fma, v_rcp_f32) and v_mfma_f32_4x4x1f32 adds B_k * r_k into the
// lane's own column: with A from lane 4b+c and B from lane l, lane l's
// register c receives A(lane 4*(l/4) + c) * B(lane l).  So lane 4b+c supplies
// component c of B_k as the A operand and its own r_k as the B operand.  The matrix unit accumulates one
// 64-entry tile (FP32); tiles are Kahan-summed; the final combination
// sum r M a - u sum r M is done in FP64.  Near field as in k_force32.
typedef float mf4 __attribute__((ext_vector_type(4)));

template <int P, int LAW>
__global__ __launch_bounds__(64) void k_force32m(ForceArgs A)
{
    __shared__ float2 atile[BH_WAVE];          // (a_x, a_y) relative to the item origin
    __shared__ float btile[BH_WAVE][4];        // (M a_x, M a_y, M, 0)
    __shared__ int32_t ktile[BH_WAVE];         // kx << 16 | ky
    __shared__ NearE ntile[BH_WAVE];
    const int lane = threadIdx.x;
    const int bc = lane & 3;
    const int nitems = A.stats[ST_NITEMS];
    int qb, qe;
    part_range(nitems, A.qpart, A.nqparts, qb, qe);
    for (;;) {
        const int q = qb + next_item(A.stats, lane);
        if (q >= qe) break;
        const long long t0 = A.cost_log ? (long long)wall_clock64() : 0;
        const int itx = A.order ? uniform_int(A.order[q]) : q;
        if ((unsigned)itx >= (unsigned)nitems) continue;
        const Item it = A.items[itx];
        const SysDev& d = A.sys[it.sys];
        const int G = d.G;
        const int64_t coff = d.c_off;
        const int cx = it.cx;
        const int oy = (it.y0 + it.y1) >> 1;
        const int ylo = max(it.y0 - 1, 0), yhi = min(it.y1 + 1, G - 1);
        const double step = d.step;
        const double inv_step = 1.0 / step;
        const float eps2c = (float)(A.eps2 * inv_step * inv_step);
        const float r2c = (float)(A.r2cut * inv_step * inv_step);

        double xi[P], yi[P];
        float ux[P], uy[P], Sx[P], Sy[P], Cx[P], Cy[P], tx[P], ty[P];
        float Fx[P], Fy[P], Fm[P], Gx[P], Gy[P], Gm[P];   // far sums (Kahan) of r M a_x, r M a_y, r M
        int cy[P], gi[P];
        bool act[P];
#pragma unroll
        for (int k = 0; k < P; ++k) {
            int slot = lane + BH_WAVE * k;
            act[k] = slot < it.np;
            gi[k] = it.p0 + slot;
            xi[k] = 0.0;
            yi[k] = 0.0;
            ux[k] = 0.0f;
            uy[k] = 0.0f;
            cy[k] = -100000;
            if (act[k]) {
                xi[k] = A.x[gi[k]];
                yi[k] = A.y[gi[k]];
                cy[k] = (int)(A.key_sorted[gi[k]] - coff) - cx * G;
                ux[k] = (float)((xi[k] - d.min_x) / step - (double)cx);
                uy[k] = (float)((yi[k] - d.min_y) / step - (double)oy);
            }
            Sx[k] = Sy[k] = Cx[k] = Cy[k] = tx[k] = ty[k] = 0.0f;
            Fx[k] = Fy[k] = Fm[k] = Gx[k] = Gy[k] = Gm[k] = 0.0f;
        }

        // near field (separations in FP64, rounded once to FP32)
        for (int dx = -1; dx <= 1; ++dx) {
            const int xc = cx + dx;
            if (xc < 0 || xc >= G) continue;
            const int64_t cl = coff + (int64_t)xc * G;
            const int pa = A.offsets[cl + ylo], pb = A.offsets[cl + yhi + 1];
            for (int base = pa; base < pb; base += BH_WAVE) {
                const int j = base + lane;
                NearE e;
                if (j < pb) {
                    e.x = A.x[j];
                    e.y = A.y[j];
                    e.m = A.m[j];
                    e.row = (int)(A.key_sorted[j] - cl);
                    e.idx = j;
                } else {
                    e.x = 0.0;
                    e.y = 0.0;
                    e.m = 0.0;
                    e.row = -100000;
                    e.idx = -1;
                }
                ntile[lane] = e;
                __syncthreads();
                const int cnt = min(BH_WAVE, pb - base);
                for (int jj = 0; jj < cnt; ++jj) {
                    const NearE f = ntile[jj];
                    const float fm = (float)f.m;
#pragma unroll
                    for (int k = 0; k < P; ++k) {
                        const bool keep = (f.idx != gi[k]) && (abs(f.row - cy[k]) <= 1);
                        const float wx = (float)((f.x - xi[k]) * inv_step);
                        const float wy = (float)((f.y - yi[k]) * inv_step);
                        const float s2 = fmaf(wx, wx, wy * wy);
                        float t = fm * __builtin_amdgcn_rcpf(s2 + eps2c);
                        if (LAW == 1) t = (s2 > r2c) ? t : 0.0f;
                        t = keep ? t : 0.0f;
                        tx[k] = fmaf(wx, t, tx[k]);
                        ty[k] = fmaf(wy, t, ty[k]);
                    }
                }
#pragma unroll
                for (int k = 0; k < P; ++k) {
                    kahan_add(Sx[k], Cx[k], tx[k]);
                    kahan_add(Sy[k], Cy[k], ty[k]);
                    tx[k] = 0.0f;
                    ty[k] = 0.0f;
                }
                __syncthreads();
            }
        }

        // far field
        const int ne0 = A.neprefix[coff], ne1 = A.neprefix[coff + (int64_t)G * G];
        int mlo[3], mhi[3];
#pragma unroll
        for (int t = 0; t < 3; ++t) {
            const int xc = cx + t - 1;
            if (xc >= 0 && xc < G) {
                const int64_t cl = coff + (int64_t)xc * G;
                mlo[t] = A.neprefix[cl + ylo];
                mhi[t] = A.neprefix[cl + yhi + 1];
            } else {
                mlo[t] = 0;
                mhi[t] = 0;
            }
        }
        Com32 nxt;
        {
            const int e = ne0 + lane;
            if (e < ne1) nxt = A.l32[e];
            else { nxt.fx = 1e18f; nxt.fy = 1e18f; nxt.rM = __builtin_huge_valf(); nxt.kxy = 0x7fff7fff; }
        }
        for (int base = ne0; base < ne1; base += BH_WAVE) {
            {
                const int kx = nxt.kxy >> 16, ky = nxt.kxy & 0xffff;
                const float ax_ = (float)(kx - cx) + nxt.fx;
                const float ay_ = (float)(ky - oy) + nxt.fy;
                atile[lane] = make_float2(ax_, ay_);
                const float Mv = 1.0f / nxt.rM;
                btile[lane][0] = Mv * ax_;
                btile[lane][1] = Mv * ay_;
                btile[lane][2] = Mv;
                btile[lane][3] = 0.0f;
                ktile[lane] = nxt.kxy;
            }
            __syncthreads();
            {
                const int e = base + BH_WAVE + lane;
                if (e < ne1) nxt = A.l32[e];
                else { nxt.fx = 1e18f; nxt.fy = 1e18f; nxt.rM = __builtin_huge_valf(); nxt.kxy = 0x7fff7fff; }
            }
            const int b1 = base + BH_WAVE;
            const bool masked = meets(base, b1, mlo[0], mhi[0]) || meets(base, b1, mlo[1], mhi[1]) ||
                                meets(base, b1, mlo[2], mhi[2]);
            mf4 acc[P];
#pragma unroll
            for (int k = 0; k < P; ++k) acc[k] = (mf4){0.0f, 0.0f, 0.0f, 0.0f};
            if (!masked) {
#pragma unroll
                for (int j = 0; j < BH_WAVE; ++j) {
                    const float2 a = atile[j];
                    const float bj = btile[j][bc];
#pragma unroll
                    for (int k = 0; k < P; ++k) {
                        const float wx = a.x - ux[k], wy = a.y - uy[k];
                        float r;
                        if (LAW == 0) {
                            r = __builtin_amdgcn_rcpf(fmaf(wx, wx, fmaf(wy, wy, eps2c)));
                        } else {
                            const float s2 = fmaf(wx, wx, wy * wy);
                            r = (s2 > r2c) ? __builtin_amdgcn_rcpf(s2 + eps2c) : 0.0f;
                        }
                        acc[k] = __builtin_amdgcn_mfma_f32_4x4x1f32(bj, r, acc[k], 0, 0, 0);
                    }
                }
            } else {
                for (int j = 0; j < BH_WAVE; ++j) {
                    const float2 a = atile[j];
                    const float bj = btile[j][bc];
                    const int kxy = ktile[j];
                    const bool colnear = abs((kxy >> 16) - cx) <= 1;
                    const int ky = kxy & 0xffff;
#pragma unroll
                    for (int k = 0; k < P; ++k) {
                        const bool keep = !(colnear && abs(ky - cy[k]) <= 1);
                        const float wx = a.x - ux[k], wy = a.y - uy[k];
                        float r;
                        if (LAW == 0) {
                            r = __builtin_amdgcn_rcpf(fmaf(wx, wx, fmaf(wy, wy, eps2c)));
                        } else {
                            const float s2 = fmaf(wx, wx, wy * wy);
                            r = (s2 > r2c) ? __builtin_amdgcn_rcpf(s2 + eps2c) : 0.0f;
                        }
                        r = keep ? r : 0.0f;
                        acc[k] = __builtin_amdgcn_mfma_f32_4x4x1f32(bj, r, acc[k], 0, 0, 0);
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < P; ++k) {
                kahan_add(Fx[k], Gx[k], acc[k][0]);
                kahan_add(Fy[k], Gy[k], acc[k][1]);
                kahan_add(Fm[k], Gm[k], acc[k][2]);
            }
            __syncthreads();
        }

        const double scale = A.Gc * inv_step;
#pragma unroll
        for (int k = 0; k < P; ++k) {
            if (act[k]) {
                const double sm = (double)Fm[k] - (double)Gm[k];
                const double fx = ((double)Fx[k] - (double)Gx[k]) - (double)ux[k] * sm;
                const double fy = ((double)Fy[k] - (double)Gy[k]) - (double)uy[k] * sm;
                const double rx = fx + ((double)Sx[k] - (double)Cx[k]);
                const double ry = fy + ((double)Sy[k] - (double)Cy[k]);
                A.ax[gi[k]] += scale * rx;
                A.ay[gi[k]] += scale * ry;
            }
        }
        if (A.cost_log && lane == 0) A.cost_log[itx] = (long long)wall_clock64() - t0;
    }
}

// ---------------------------------------------------------------- dispatch table
typedef void (*force_fn)(ForceArgs);

#define F64(P, L) (force_fn)k_force64<P, L, 1>
#define F64I(P, L) (force_fn)k_force64<P, L, 0>
#define F32(P, L, K) (force_fn)k_force32<P, L, K>
#define F32M(P, L) (force_fn)k_force32m<P, L>

// [precision: 0 fp64, 1..3 fp32 kahan mode 1,2,0, 4 fp32 matrix-core, 5 fp64 IEEE
// division][P index][law]
static const force_fn g_force[6][3][2] = {
    {{F64(1, 0), F64(1, 1)}, {F64(2, 0), F64(2, 1)}, {F64(4, 0), F64(4, 1)}},
    {{F32(1, 0, 1), F32(1, 1, 1)}, {F32(2, 0, 1), F32(2, 1, 1)}, {F32(4, 0, 1), F32(4, 1, 1)}},
    {{F32(1, 0, 2), F32(1, 1, 2)}, {F32(2, 0, 2), F32(2, 1, 2)}, {F32(4, 0, 2), F32(4, 1, 2)}},
    {{F32(1, 0, 0), F32(1, 1, 0)}, {F32(2, 0, 0), F32(2, 1, 0)}, {F32(4, 0, 0), F32(4, 1, 0)}},
    {{F32M(1, 0), F32M(1, 1)}, {F32M(2, 0), F32M(2, 1)}, {F32M(4, 0), F32M(4, 1)}},
    {{F64I(1, 0), F64I(1, 1)}, {F64I(2, 0), F64I(2, 1)}, {F64I(4, 0), F64I(4, 1)}},
};

const void* bh_force_kernel(int precision, int pidx, int law)
{
    if (precision < 0 || precision > 5 || pidx < 0 || pidx > 2 || law < 0 || law > 1) return nullptr;
    return (const void*)g_force[precision][pidx][law];
}
