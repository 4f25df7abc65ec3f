// FP64 recompute of the particles whose FP32 accelerations can carry a large
// relative rounding error.
//
// Each FP32 term (near or far) is rounded to a few ulp of its own magnitude,
// with signs that do not line up, so a particle's error grows like
// eps32 * sqrt(sum_j |term_j|^2), while its acceleration is |sum_j term_j|.
// Where the terms nearly cancel (cluster centres, saddles of the field) the
// relative error is large even though the normwise error over all particles
// stays ~1e-7.  After an FP32 pass, k_rc_flag bounds sqrt(sum_j |term_j|^2)
// from above for every particle and flags those whose bound exceeds kappa
// times their acceleration; k_rc_fix64 evaluates the flagged particles in
// FP64 (the law's terms with a Newton-refined reciprocal).  The FP32 kernels
// write into scratch planes (a32x, a32y); each particle's acceleration gets
// either its FP32 result (k_rc_flag) or its FP64 one (k_rc_fix64 /
// k_rc_combine) added.
//
// The bound, for a work item (particles of column cx, rows [y0, y1]; box
// [cx, cx + 1] x [y0, y1 + 1] in cell units):
//  * near terms: summed exactly by the FP32 kernel (s2n, one fma per term);
//  * far cells of the exact region (columns cx-R..cx+R, rows y0-R..y1+R,
//    R = RC_EXACT) that are far for the particle: exact (M / d)^2, per
//    particle -- computed only for particles that a cheaper bound of this
//    region (M^2 / max(1, dist(box, COM))^2 per cell, taken per item) does
//    not already clear;
//  * other cells of the exact region's hull in RC_TILE x RC_TILE tiles:
//    M^2 / max(1, dist(box, COM))^2 -- each of the item's particles is at
//    least that far from the COM;
//  * other tiles of the item's window (the RC_BLOCK x RC_BLOCK blocks within
//    one block of the item's blocks): (sum of M^2 over the tile) / dist(box,
//    tile)^2, with dist >= R;
//  * blocks outside that window: (sum of M^2 over the block) / dist(box,
//    block)^2, with dist >= RC_BLOCK; for an item inside one block this part
//    is taken per block, with the distance from the block's box (no shorter
//    than from the item's box), by k_rc_blockfar.
// |term| = M d / (d^2 + eps2) <= M / d under softening, and a cutoff only
// removes terms, so the bound holds for every law.  The bound's quotients use
// v_rcp_f64 (relative error far below the margin kappa leaves).
//
// Flagged particles of system s are appended to list[p_off_s, p_off_s +
// cnt_s) and evaluated in groups of RC_GROUP; when there are few groups, each
// group's COM list is split into slices whose partial sums are added in slice
// order.  A particle's result depends only on its own terms and on the number
// of groups, never on which particles share its group.
#include "bh_common.h"
#include "bh_kernels.h"

// Sum of M^2 over each RC_BLOCK x RC_BLOCK block of cells (one wave per block;
// fixed summation order, so the bound is the same on every run).
__global__ __launch_bounds__(64) void k_rc_blocks(const SysDev* __restrict__ sys, int S,
                                                  int64_t nblocks,
                                                  const double* __restrict__ com_dense,
                                                  double* __restrict__ blk)
{
    const int lane = threadIdx.x;
    for (int64_t b = blockIdx.x; b < nblocks; b += gridDim.x) {
        int lo = 0, hi = S - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (sys[mid].b_off <= b) lo = mid; else hi = mid - 1;
        }
        const SysDev& d = sys[lo];
        const int64_t lb = b - d.b_off;
        const int bx = (int)(lb / d.Gb), by = (int)(lb - (int64_t)bx * d.Gb);
        double acc = 0.0;
#pragma unroll
        for (int i = 0; i < RC_BLOCK * RC_BLOCK / BH_WAVE; ++i) {
            const int t = i * BH_WAVE + lane;
            const int gx = bx * RC_BLOCK + t / RC_BLOCK, gy = by * RC_BLOCK + t % RC_BLOCK;
            if (gx < d.G && gy < d.G) {
                const double M = com_dense[3 * (d.c_off + (int64_t)gx * d.G + gy) + 2];
                acc = fma(M, M, acc);
            }
        }
        acc = wave_sum(acc);
        if (lane == 0) blk[b] = acc;
    }
}

// Sum of M^2 over each RC_TILE x RC_TILE tile of cells (one thread per tile,
// fixed order); system s's tiles start at b_off * (RC_BLOCK / RC_TILE)^2, in
// a Gb * RC_BLOCK / RC_TILE square.
__global__ __launch_bounds__(256) void k_rc_tiles(const SysDev* __restrict__ sys, int S,
                                                  int64_t ntiles,
                                                  const double* __restrict__ com_dense,
                                                  double* __restrict__ tl)
{
    constexpr int TPB = RC_BLOCK / RC_TILE;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < ntiles; q += stride) {
        int lo = 0, hi = S - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (sys[mid].b_off * TPB * TPB <= q) lo = mid; else hi = mid - 1;
        }
        const SysDev& d = sys[lo];
        const int Gt = d.Gb * TPB;
        const int64_t lt = q - d.b_off * TPB * TPB;
        const int tx = (int)(lt / Gt), ty = (int)(lt - (int64_t)tx * Gt);
        double acc = 0.0;
        for (int i = 0; i < RC_TILE; ++i) {
            const int gx = tx * RC_TILE + i;
            for (int j = 0; j < RC_TILE; ++j) {
                const int gy = ty * RC_TILE + j;
                if (gx < d.G && gy < d.G) {
                    const double M = com_dense[3 * (d.c_off + (int64_t)gx * d.G + gy) + 2];
                    acc = fma(M, M, acc);
                }
            }
        }
        tl[q] = acc;
    }
}

// Per block b: sum over the blocks outside b's 3 x 3 block neighbourhood of
// (sum of M^2) / dist(b, b')^2 (one wave per block, fixed order).
__global__ __launch_bounds__(64) void k_rc_blockfar(const SysDev* __restrict__ sys, int S,
                                                    int64_t nblocks,
                                                    const double* __restrict__ blk,
                                                    double* __restrict__ bfar)
{
    const int lane = threadIdx.x;
    for (int64_t b = blockIdx.x; b < nblocks; b += gridDim.x) {
        int lo = 0, hi = S - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (sys[mid].b_off <= b) lo = mid; else hi = mid - 1;
        }
        const SysDev& d = sys[lo];
        const int Gb = d.Gb;
        const int64_t lb = b - d.b_off;
        const int bx = (int)(lb / Gb), by = (int)(lb - (int64_t)bx * Gb);
        const double x0 = (double)bx * RC_BLOCK, y0 = (double)by * RC_BLOCK;
        double acc = 0.0;
        for (int t = lane; t < Gb * Gb; t += BH_WAVE) {
            const int tx = t / Gb, ty = t - (t / Gb) * Gb;
            if (abs(tx - bx) <= 1 && abs(ty - by) <= 1) continue;
            const double B2 = blk[d.b_off + t];
            if (B2 == 0.0) continue;
            const double lx = (double)tx * RC_BLOCK, ly = (double)ty * RC_BLOCK;
            const double ex = fmax(fmax(lx - (x0 + RC_BLOCK), x0 - (lx + RC_BLOCK)), 0.0);
            const double ey = fmax(fmax(ly - (y0 + RC_BLOCK), y0 - (ly + RC_BLOCK)), 0.0);
            acc = fma(B2, __builtin_amdgcn_rcp(fma(ex, ex, ey * ey)), acc);
        }
        acc = wave_sum(acc);
        if (lane == 0) bfar[b] = acc;
    }
}

// One wave per work item (grid-stride over the items); a wave that meets the
// same box again (items of one dense cell) reuses the item part of the bound.
__global__ __launch_bounds__(64) void k_rc_flag(const Item* __restrict__ items,
                                                const SysDev* __restrict__ sys,
                                                const int32_t* __restrict__ key_sorted,
                                                const int32_t* __restrict__ neprefix,
                                                const Com64* __restrict__ l64,
                                                const double* __restrict__ x,
                                                const double* __restrict__ y,
                                                const double* __restrict__ com,
                                                const double* __restrict__ tl,
                                                const double* __restrict__ blk,
                                                const double* __restrict__ bfar,
                                                const double* __restrict__ a32x,
                                                const double* __restrict__ a32y,
                                                const double* __restrict__ s2n, double Gc,
                                                double kappa2, const int32_t* __restrict__ stats,
                                                int32_t* rc_cnt, int32_t* __restrict__ list,
                                                double* __restrict__ ax, double* __restrict__ ay)
{
    __shared__ Com64 tile[BH_WAVE];
    const int lane = threadIdx.x;
    const int nitems = stats[ST_NITEMS];
    int psys = -1, pcx = -1, py0 = -1, py1 = -1;   // box of the last bound (wave-uniform)
    double psi = 0.0, preg = 0.0;   // item bound outside / over the exact region
    for (int itx = blockIdx.x; itx < nitems; itx += gridDim.x) {
        const Item it = items[itx];
        const SysDev& d = sys[it.sys];
        const int G = d.G, Gb = d.Gb;
        const int64_t coff = d.c_off;
        const double inv_step = 1.0 / d.step;
        const int cx = it.cx, y0 = it.y0, y1 = it.y1;
        const double bx0 = cx, bx1 = cx + 1, by0 = y0, by1 = y1 + 1;
        const int wbx0 = max(cx / RC_BLOCK - 1, 0), wbx1 = min(cx / RC_BLOCK + 1, Gb - 1);
        const int wby0 = max(y0 / RC_BLOCK - 1, 0), wby1 = min(y1 / RC_BLOCK + 1, Gb - 1);
        const bool same = it.sys == psys && cx == pcx && y0 == py0 && y1 == py1;
        psys = it.sys;
        pcx = cx;
        py0 = y0;
        py1 = y1;
        if (!same) {
            // cells of the exact region's hull in RC_TILE x RC_TILE tiles, outside
            // the exact region (dense walk)
            const int htx0 = max(cx - RC_EXACT, 0) / RC_TILE, htx1 = min(cx + RC_EXACT, G - 1) / RC_TILE;
            const int hty0 = max(y0 - RC_EXACT, 0) / RC_TILE, hty1 = min(y1 + RC_EXACT, G - 1) / RC_TILE;
            const int hx0 = htx0 * RC_TILE, hx1 = min(G, (htx1 + 1) * RC_TILE);
            const int hy0 = hty0 * RC_TILE, hy1 = min(G, (hty1 + 1) * RC_TILE);
            const int hrows = hy1 - hy0, nh = (hx1 - hx0) * hrows;
            psi = 0.0;
            for (int t = lane; t < nh; t += BH_WAVE) {
                const int col = t / hrows;
                const int gx = hx0 + col, gy = hy0 + (t - col * hrows);
                if (abs(gx - cx) <= RC_EXACT && gy >= y0 - RC_EXACT && gy <= y1 + RC_EXACT) continue;
                const double* cm = com + 3 * (coff + (int64_t)gx * G + gy);
                const double M = cm[2];
                if (M == 0.0) continue;
                const double X = (cm[0] - d.min_x) * inv_step, Y = (cm[1] - d.min_y) * inv_step;
                const double ex = fmax(fmax(bx0 - X, X - bx1), 0.0);
                const double ey = fmax(fmax(by0 - Y, Y - by1), 0.0);
                psi = fma(M * M, __builtin_amdgcn_rcp(fmax(fma(ex, ex, ey * ey), 1.0)), psi);
            }
            // tiles of the window outside the hull (at least RC_EXACT cells away)
            constexpr int TPB = RC_BLOCK / RC_TILE;
            const int Gt = Gb * TPB;
            const int tx0 = wbx0 * TPB, ntx = (wbx1 - wbx0 + 1) * TPB;
            const int ty0 = wby0 * TPB, nty = (wby1 - wby0 + 1) * TPB;
            const int64_t toff = d.b_off * TPB * TPB;
            for (int t = lane; t < ntx * nty; t += BH_WAVE) {
                const int a = t / nty;
                const int tx = tx0 + a, ty = ty0 + (t - a * nty);
                if (tx >= htx0 && tx <= htx1 && ty >= hty0 && ty <= hty1) continue;
                const double T2 = tl[toff + (int64_t)tx * Gt + ty];
                if (T2 == 0.0) continue;
                const double lx = (double)tx * RC_TILE, ly = (double)ty * RC_TILE;
                const double ex = fmax(fmax(lx - bx1, bx0 - (lx + RC_TILE)), 0.0);
                const double ey = fmax(fmax(ly - by1, by0 - (ly + RC_TILE)), 0.0);
                psi = fma(T2, __builtin_amdgcn_rcp(fma(ex, ex, ey * ey)), psi);
            }
            psi = wave_sum(psi);
            // blocks outside the window
            if (y0 / RC_BLOCK == y1 / RC_BLOCK) {
                psi += bfar[d.b_off + (int64_t)(cx / RC_BLOCK) * Gb + y0 / RC_BLOCK];
            } else {
                double pb = 0.0;
                const int nblk = Gb * Gb;
#pragma unroll 4
                for (int b = lane; b < nblk; b += BH_WAVE) {
                    const int bx = b / Gb, by = b - (b / Gb) * Gb;
                    if (bx >= wbx0 && bx <= wbx1 && by >= wby0 && by <= wby1) continue;
                    const double B2 = blk[d.b_off + b];
                    if (B2 == 0.0) continue;
                    const double lx = (double)bx * RC_BLOCK, ly = (double)by * RC_BLOCK;
                    const double ex = fmax(fmax(lx - bx1, bx0 - (lx + RC_BLOCK)), 0.0);
                    const double ey = fmax(fmax(ly - by1, by0 - (ly + RC_BLOCK)), 0.0);
                    pb = fma(B2, __builtin_amdgcn_rcp(fma(ex, ex, ey * ey)), pb);
                }
                psi += wave_sum(pb);
            }
            // the exact region at item level, for the cheap first test:
            // M^2 / max(1, dist(box, COM))^2; a one-row item's 3 x 3 block is
            // near for all of its particles and left out
            const int rxa = max(cx - RC_EXACT, 0), rxb = min(cx + RC_EXACT, G - 1);
            const int rya = max(y0 - RC_EXACT, 0), ryb = min(y1 + RC_EXACT, G - 1);
            const int rrows = ryb - rya + 1, nr = (rxb - rxa + 1) * rrows;
            preg = 0.0;
            for (int t = lane; t < nr; t += BH_WAVE) {
                const int col = t / rrows;
                const int gx = rxa + col, gy = rya + (t - col * rrows);
                if (y0 == y1 && abs(gx - cx) <= 1 && abs(gy - y0) <= 1) continue;
                const double* cm = com + 3 * (coff + (int64_t)gx * G + gy);
                const double M = cm[2];
                if (M == 0.0) continue;
                const double X = (cm[0] - d.min_x) * inv_step, Y = (cm[1] - d.min_y) * inv_step;
                const double ex = fmax(fmax(bx0 - X, X - bx1), 0.0);
                const double ey = fmax(fmax(by0 - Y, Y - by1), 0.0);
                preg = fma(M * M, __builtin_amdgcn_rcp(fmax(fma(ex, ex, ey * ey), 1.0)), preg);
            }
            preg = wave_sum(preg);
        }

        // |a| in the bound's units (mass / cells): a * step / G
        const double g2 = (d.step / Gc) * (d.step / Gc);
        // the exact region's entries: one compact range per column (lane c <
        // 2 * RC_EXACT + 1 holds column xa + c), walked as one flattened list
        const int xa = max(cx - RC_EXACT, 0), xb = min(cx + RC_EXACT, G - 1);
        const int ya = max(y0 - RC_EXACT, 0), yb = min(y1 + RC_EXACT, G - 1);
        const int nxc = xb - xa + 1;
        int eb = 0, en = 0;
        if (lane < nxc) {
            const int64_t cl = coff + (int64_t)(xa + lane) * G;
            eb = neprefix[cl + ya];
            en = neprefix[cl + yb + 1] - eb;
        }
        int einc = en;
#pragma unroll
        for (int o = 1; o < 2 * RC_EXACT + 1; o <<= 1) {
            const int v = __shfl_up(einc, o, BH_WAVE);
            if (lane >= o) einc += v;
        }
        const int etot = __shfl(einc, nxc - 1, BH_WAVE);
        const int epre = einc - en;
        for (int base = 0; base < it.np; base += BH_WAVE) {
            const int slot = base + lane;
            const bool act = slot < it.np;
            const int64_t gi = (int64_t)it.p0 + (act ? slot : 0);
            const double xi = x[gi], yi = y[gi];
            const int cy = (int)(key_sorted[gi] - coff) - cx * G;
            const double sn = s2n[gi];
            const double axv = a32x[gi], ayv = a32y[gi];
            const double A2 = fma(axv, axv, ayv * ayv) * g2;
            // first test with the item-level bound of the exact region; the
            // exact per-particle sum only runs where some lane passes it
            const double loose = psi + preg + sn;
            // a non-finite FP32 result (overflow, coincident particles) always
            // takes the FP64 path
            const bool bad = !(A2 <= 1.7976931348623157e308);
            const bool cand = act && ((loose != 0.0 && !(loose <= kappa2 * A2)) || bad);
            if (act && !cand) {
                ax[gi] += axv;
                ay[gi] += ayv;
            }
            if (__ballot(cand) == 0ull) continue;
            double est = psi + sn;
            // far cells of the exact region, per particle (LDS tiles)
            for (int tb = 0; tb < etot; tb += BH_WAVE) {
                const int t = tb + lane;
                int c = 0;
#pragma unroll
                for (int k = 1; k < 2 * RC_EXACT + 1; ++k)
                    if (k < nxc && __shfl(epre, k, BH_WAVE) <= t) c = k;
                const int e = __shfl(eb, c, BH_WAVE) + (t - __shfl(epre, c, BH_WAVE));
                if (t < etot) tile[lane] = l64[e];
                __syncthreads();
                const int cnt = min(BH_WAVE, etot - tb);
                for (int j = 0; j < cnt; ++j) {
                    const Com64 f = tile[j];
                    if (max(abs(f.kx - cx), abs(f.ky - cy)) >= 2) {
                        const double ddx = (f.X - xi) * inv_step, ddy = (f.Y - yi) * inv_step;
                        est = fma(f.M * f.M, __builtin_amdgcn_rcp(fma(ddx, ddx, ddy * ddy)), est);
                    }
                }
                __syncthreads();
            }
            // NaN or infinite estimates flag too: the FP64 terms decide
            const bool flag = cand && ((est != 0.0 && !(est <= kappa2 * A2)) || bad);
            if (cand && !flag) {
                ax[gi] += axv;
                ay[gi] += ayv;
            }
            const unsigned long long mask = __ballot(flag);
            const int n = __popcll(mask);
            if (n == 0) continue;
            int bp = 0;
            if (lane == 0) bp = atomicAdd(&rc_cnt[it.sys], n);
            bp = __shfl(bp, 0, BH_WAVE);
            if (flag) list[d.p_off + bp + __popcll(mask & ((1ull << lane) - 1ull))] = (int32_t)gi;
        }
    }
}

// Groups per system: gofs[s] = sum_{t < s} ceil(cnt_t / RC_GROUP) (one block of
// 256 threads), and the totals into stats.
__global__ __launch_bounds__(256) void k_rc_groups(const int32_t* __restrict__ rc_cnt, int S,
                                                   int32_t* __restrict__ gofs, int32_t* stats)
{
    __shared__ int32_t sg[256], sn[256];
    const int t = threadIdx.x;
    int carry_g = 0, carry_n = 0;
    for (int base = 0; base < S; base += 256) {
        const int s = base + t;
        const int c = s < S ? rc_cnt[s] : 0;
        const int gs = (c + RC_GROUP - 1) / RC_GROUP;
        sg[t] = gs;
        sn[t] = c;
        __syncthreads();
        for (int o = 1; o < 256; o <<= 1) {
            const int vg = t >= o ? sg[t - o] : 0, vn = t >= o ? sn[t - o] : 0;
            __syncthreads();
            sg[t] += vg;
            sn[t] += vn;
            __syncthreads();
        }
        if (s < S) gofs[s] = carry_g + sg[t] - gs;
        carry_g += sg[255];
        carry_n += sn[255];
        __syncthreads();
    }
    if (t == 0) {
        gofs[S] = carry_g;
        stats[ST_NRECOMP] = carry_n;
        stats[ST_RCGROUPS] = carry_g;
    }
}

template <int LAW>
__device__ __forceinline__ double rc_w64(double wx, double wy, double m, double eps2, double r2cut)
{
    const double s2 = fma(wx, wx, wy * wy);
    if (LAW == 0) return wdiv64(m, s2 + eps2);
    return (s2 > r2cut) ? wdiv64(m, s2 + eps2) : 0.0;
}

__device__ __forceinline__ int rc_slices(int ngroups)
{
    return ngroups > 0 ? max(1, min(RC_SLICES, RC_PART_CAP / ngroups)) : 1;
}

__device__ __forceinline__ int rc_sys_of_group(const int32_t* gofs, int S, int g)
{
    int lo = 0, hi = S - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (gofs[mid] <= g) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// One block per (group, slice) (grid-stride).  The threads split the slice of
// the system's COM list and, in slice 0, the near cells; the block's sums are
// reduced in a fixed order.  With one slice the FP64 result is added to the
// acceleration here, else k_rc_combine adds it.
template <int LAW>
__global__ __launch_bounds__(RC_THREADS) void k_rc_fix64(
    const int32_t* __restrict__ stats, const int32_t* __restrict__ rc_cnt,
    const int32_t* __restrict__ gofs, int S, const int32_t* __restrict__ list,
    const SysDev* __restrict__ sys, const int32_t* __restrict__ key_sorted,
    const int32_t* __restrict__ offsets, const int32_t* __restrict__ neprefix,
    const Com64* __restrict__ l64, const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ m, double Gc, double eps2, double r2cut,
    double* __restrict__ part, double* ax, double* ay)
{
    constexpr int NW = RC_THREADS / BH_WAVE;
    __shared__ double red[NW][2 * RC_GROUP];
    const int tid = threadIdx.x, lane = tid & (BH_WAVE - 1), wv = tid / BH_WAVE;
    const int ngroups = stats[ST_RCGROUPS];
    const int se = rc_slices(ngroups);
    const int64_t total = (int64_t)ngroups * se;
    for (int64_t w = blockIdx.x; w < total; w += gridDim.x) {
        const int g = (int)(w / se), sl = (int)(w - (int64_t)g * se);
        const int s = rc_sys_of_group(gofs, S, g);
        const SysDev& d = sys[s];
        const int k0 = (g - gofs[s]) * RC_GROUP;
        const int cnt = min(RC_GROUP, rc_cnt[s] - k0);
        const int64_t lb = d.p_off + k0;
        const int G = d.G;
        const int64_t coff = d.c_off;
        int gk[RC_GROUP], cxk[RC_GROUP], cyk[RC_GROUP];
        double xk[RC_GROUP], yk[RC_GROUP], sx[RC_GROUP], sy[RC_GROUP];
#pragma unroll
        for (int k = 0; k < RC_GROUP; ++k) {
            gk[k] = list[lb + min(k, cnt - 1)];   // short groups repeat their last particle
            xk[k] = x[gk[k]];
            yk[k] = y[gk[k]];
            const int cl = (int)(key_sorted[gk[k]] - coff);
            cxk[k] = cl / G;
            cyk[k] = cl - cxk[k] * G;
            sx[k] = 0.0;
            sy[k] = 0.0;
        }
        // far field: this slice of the compact COM list, minus each particle's 3x3 block
        const int ne0 = neprefix[coff], ne1 = neprefix[coff + (int64_t)G * G];
        const int chunk = (ne1 - ne0 + se - 1) / se;
        const int e0 = ne0 + sl * chunk, e1 = min(ne1, e0 + chunk);
        Com64 nxt;
        if (e0 + tid < e1) nxt = l64[e0 + tid];
        for (int e = e0 + tid; e < e1; e += RC_THREADS) {
            const Com64 f = nxt;
            if (e + RC_THREADS < e1) nxt = l64[e + RC_THREADS];
#pragma unroll
            for (int k = 0; k < RC_GROUP; ++k) {
                const bool keep = max((unsigned)(f.kx - cxk[k] + 1), (unsigned)(f.ky - cyk[k] + 1)) > 2u;
                const double wx = f.X - xk[k], wy = f.Y - yk[k];
                double w = rc_w64<LAW>(wx, wy, f.M, eps2, r2cut);
                w = keep ? w : 0.0;
                sx[k] = fma(wx, w, sx[k]);
                sy[k] = fma(wy, w, sy[k]);
            }
        }
        // near field (slice 0): particles of the 3x3 block, the particle itself excluded
        if (sl == 0) {
#pragma unroll
            for (int k = 0; k < RC_GROUP; ++k) {
                if (k < cnt) {
                    const int ra = max(cyk[k] - 1, 0), rb = min(cyk[k] + 1, G - 1);
                    for (int xc = max(cxk[k] - 1, 0); xc <= min(cxk[k] + 1, G - 1); ++xc) {
                        const int64_t cl = coff + (int64_t)xc * G;
                        const int pa = offsets[cl + ra], pb = offsets[cl + rb + 1];
                        for (int j = pa + tid; j < pb; j += RC_THREADS) {
                            const double wx = x[j] - xk[k], wy = y[j] - yk[k];
                            double w = rc_w64<LAW>(wx, wy, m[j], eps2, r2cut);
                            w = (j != gk[k]) ? w : 0.0;
                            sx[k] = fma(wx, w, sx[k]);
                            sy[k] = fma(wy, w, sy[k]);
                        }
                    }
                }
            }
        }
#pragma unroll
        for (int k = 0; k < RC_GROUP; ++k) {
            sx[k] = wave_sum(sx[k]);
            sy[k] = wave_sum(sy[k]);
        }
        if (lane == 0) {
#pragma unroll
            for (int k = 0; k < RC_GROUP; ++k) {
                red[wv][2 * k] = sx[k];
                red[wv][2 * k + 1] = sy[k];
            }
        }
        __syncthreads();
        if (tid < 2 * RC_GROUP) {
            double r = 0.0;
            for (int v = 0; v < NW; ++v) r += red[v][tid];
            const int k = tid >> 1, c = tid & 1;
            if (se == 1) {
                if (k < cnt) {
                    const int64_t gi = list[lb + k];
                    double* a = c ? ay : ax;
                    a[gi] += Gc * r;
                }
            } else {
                part[w * (2 * RC_GROUP) + tid] = r;
            }
        }
        __syncthreads();
    }
}

// Adds the slices' partial sums in slice order and the FP64 result to each
// flagged particle's acceleration (only when slices were used).
__global__ __launch_bounds__(256) void k_rc_combine(const int32_t* __restrict__ stats,
                                                    const int32_t* __restrict__ rc_cnt,
                                                    const int32_t* __restrict__ gofs, int S,
                                                    const int32_t* __restrict__ list,
                                                    const SysDev* __restrict__ sys,
                                                    const double* __restrict__ part, double Gc,
                                                    double* ax, double* ay)
{
    const int ngroups = stats[ST_RCGROUPS];
    const int se = rc_slices(ngroups);
    if (se == 1) return;
    const int64_t total = (int64_t)ngroups * RC_GROUP;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < total;
         q += (int64_t)gridDim.x * blockDim.x) {
        const int g = (int)(q / RC_GROUP), k = (int)(q - (int64_t)g * RC_GROUP);
        const int s = rc_sys_of_group(gofs, S, g);
        const int k0 = (g - gofs[s]) * RC_GROUP;
        if (k >= rc_cnt[s] - k0) continue;
        const int64_t gi = list[sys[s].p_off + k0 + k];
        double rx = 0.0, ry = 0.0;
        for (int sl = 0; sl < se; ++sl) {
            const double* p = part + ((int64_t)g * se + sl) * (2 * RC_GROUP);
            rx += p[2 * k];
            ry += p[2 * k + 1];
        }
        ax[gi] += Gc * rx;
        ay[gi] += Gc * ry;
    }
}

#define RC_FIX64_INST(L)                                                                      \
    template __global__ void k_rc_fix64<L>(const int32_t*, const int32_t*, const int32_t*, int, \
                                           const int32_t*, const SysDev*, const int32_t*,      \
                                           const int32_t*, const int32_t*, const Com64*,       \
                                           const double*, const double*, const double*,        \
                                           double, double, double, double*, double*, double*);
RC_FIX64_INST(0)
RC_FIX64_INST(1)
