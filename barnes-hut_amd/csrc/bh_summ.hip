// Summarize (per-cell centre of mass), tick, and the exact interaction
// counter used for work accounting.
//   g_summarize       python-bh/barneshut/kernels/grid_decomposition/gpu/grid.py:78-114
//   calc_COM          python-bh/barneshut/internals/cloud.py:9-17
//   g_tick_particles  python-bh/barneshut/kernels/grid_decomposition/gpu/grid.py:159-174
#include "bh_common.h"
#include "bh_kernels.h"

// Compact list of non-empty cells in global cell order.
__global__ void k_ne_list(const int32_t* __restrict__ counts, const int32_t* __restrict__ neprefix,
                          int64_t C, int32_t* __restrict__ ne_list)
{
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < C; c += stride)
        if (counts[c] > 0) ne_list[neprefix[c]] = (int32_t)c;
}

// FP64 sums of m*x, m*y, m over each non-empty cell's sorted particles.
// Writes the dense G*G x 3 array of the reference layout (empty cells were
// zero-filled) and the two compact lists the force kernels stream.
__global__ __launch_bounds__(256) void k_com(const SysDev* __restrict__ sys, int S,
                                             const int32_t* __restrict__ ne_list,
                                             const int32_t* __restrict__ stats,
                                             const int32_t* __restrict__ offsets,
                                             const double* __restrict__ x,
                                             const double* __restrict__ y,
                                             const double* __restrict__ m,
                                             double* __restrict__ com_dense,
                                             Com64* __restrict__ l64, Com32* __restrict__ l32)
{
    // A wave works on four cells at a time, one per 16-lane quarter (most
    // cells hold a few dozen particles at most); the quarter's sums are
    // butterfly-reduced inside the quarter (fixed order, deterministic).
    constexpr int QW = 16;
    const int lane = threadIdx.x & 63;
    const int q = lane >> 4, ql = lane & (QW - 1);
    const int64_t wv = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
    const int nne = stats[ST_NCELLS_NE];
    // each wave takes a contiguous run of the non-empty list, so it crosses
    // few system boundaries and looks its system up only when it does
    const int64_t per = (((nne + nw - 1) / nw) + 3) & ~(int64_t)3;
    const int64_t w0 = wv * per, w1 = min((int64_t)nne, w0 + per);
    int s = -1;
    int64_t s_end = -1;
    for (int64_t wb = w0; wb < w1; wb += 4) {
        const int64_t w = wb + q;
        const bool live = w < w1;
        const int c = live ? ne_list[w] : 0;
        if (live && c >= s_end) {
            s = sys_of_cell(sys, S, c);
            s_end = (s + 1 < S) ? sys[s + 1].c_off : INT64_MAX;
        }
        const int a = live ? offsets[c] : 0, b = live ? offsets[c + 1] : 0;
        double sx = 0.0, sy = 0.0, sm = 0.0;
        for (int i = a + ql; i < b; i += QW) {
            double mi = m[i];
            sx = fma(x[i], mi, sx);
            sy = fma(y[i], mi, sy);
            sm += mi;
        }
#pragma unroll
        for (int o = QW / 2; o > 0; o >>= 1) {
            sx += __shfl_xor(sx, o, BH_WAVE);
            sy += __shfl_xor(sy, o, BH_WAVE);
            sm += __shfl_xor(sm, o, BH_WAVE);
        }
        if (live && ql == 0) {
            double X = 0.0, Y = 0.0, M = 0.0;
            if (sm != 0.0) {
                X = sx / sm;
                Y = sy / sm;
                M = sm;
            }
            com_dense[3 * (int64_t)c + 0] = X;
            com_dense[3 * (int64_t)c + 1] = Y;
            com_dense[3 * (int64_t)c + 2] = M;
            const SysDev& d = sys[s];
            int local = (int)(c - d.c_off);
            int kx = local / d.G, ky = local - kx * d.G;
            Com64 e64;
            e64.X = X;
            e64.Y = Y;
            e64.M = M;
            e64.kx = kx;
            e64.ky = ky;
            l64[w] = e64;
            Com32 e32;
            e32.fx = (float)((X - d.min_x) / d.step - (double)kx);
            e32.fy = (float)((Y - d.min_y) / d.step - (double)ky);
            e32.rM = 1.0f / (float)M;
            e32.kxy = (kx << 16) | ky;
            l32[w] = e32;
        }
    }
}

// v += a*dt; a = 0; x += v*dt  (separate multiply and add; the library is
// built with -ffp-contract=off).
__global__ __launch_bounds__(256) void k_tick(double* __restrict__ x, double* __restrict__ y,
                                              double* __restrict__ vx, double* __restrict__ vy,
                                              double* __restrict__ ax, double* __restrict__ ay,
                                              int64_t N, double dt)
{
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += stride) {
        double vxi = vx[i] + ax[i] * dt;
        double vyi = vy[i] + ay[i] * dt;
        vx[i] = vxi;
        vy[i] = vyi;
        ax[i] = 0.0;
        ay[i] = 0.0;
        x[i] = x[i] + vxi * dt;
        y[i] = y[i] + vyi * dt;
    }
}

// Exact interaction counts of one evaluation, per system and in total:
//   near_s = sum_c n_c * sum_{3x3} n_c'   (the host subtracts N_s)
//   far_s  = sum_c n_c * (K_s - #non-empty in 3x3)
__global__ __launch_bounds__(256) void k_interactions(const SysDev* __restrict__ sys, int S,
                                                      const int32_t* __restrict__ counts,
                                                      const int32_t* __restrict__ neprefix,
                                                      int64_t C, unsigned long long* out)
{
    // Each wave keeps running sums for the system its cells belong to and
    // flushes them with atomics only when that system changes and at the end
    // (cells of one system are contiguous), so the four counters see a few
    // atomics per wave instead of one per 64 cells.
    const int lane = threadIdx.x & 63;
    unsigned long long acc_n = 0, acc_f = 0;
    int acc_s = -1;
    auto flush = [&]() {
        if (lane == 0 && acc_s >= 0) {
            atomicAdd(&out[2 * acc_s], acc_n);
            atomicAdd(&out[2 * acc_s + 1], acc_f);
            atomicAdd(&out[2 * S], acc_n);
            atomicAdd(&out[2 * S + 1], acc_f);
        }
        acc_n = acc_f = 0;
    };
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t c0 = (int64_t)blockIdx.x * blockDim.x; c0 < C; c0 += stride) {
        int64_t c = c0 + threadIdx.x;
        unsigned long long nr = 0, fr = 0;
        int s = -1;
        if (c < C && counts[c] > 0) {
            s = sys_of_cell(sys, S, c);
            const SysDev& d = sys[s];
            int G = d.G;
            int local = (int)(c - d.c_off);
            int cx = local / G, cy = local - cx * G;
            long long n9 = 0, k9 = 0;
            for (int dx = -1; dx <= 1; ++dx)
                for (int dy = -1; dy <= 1; ++dy) {
                    int nx = cx + dx, ny = cy + dy;
                    if (nx < 0 || ny < 0 || nx >= G || ny >= G) continue;
                    int cc = counts[d.c_off + (int64_t)nx * G + ny];
                    n9 += cc;
                    k9 += cc > 0;
                }
            long long K = neprefix[d.c_off + (int64_t)G * G] - neprefix[d.c_off];
            long long nc = counts[c];
            nr = (unsigned long long)(nc * n9);
            fr = (unsigned long long)(nc * (K - k9));
        }
        const int sv = wave_max(s);
        const bool uni = __all(s == sv || s < 0);
        if (uni) {
            if (sv >= 0) {
                const unsigned long long a = wave_sum(nr), b = wave_sum(fr);
                if (sv != acc_s) {
                    flush();
                    acc_s = sv;
                }
                acc_n += a;
                acc_f += b;
            }
        } else if (s >= 0) {
            atomicAdd(&out[2 * s], nr);
            atomicAdd(&out[2 * s + 1], fr);
            atomicAdd(&out[2 * S], nr);
            atomicAdd(&out[2 * S + 1], fr);
        }
    }
    flush();
}

// Energy diagnostic (direct sum, FP64): for each particle i of one system's
// rows [p0, p0 + n), u[i] = m_i * sum_{j != i} m_j * log(r_ij^2 + eps2) and
// k[i] = m_i * (vx_i^2 + vy_i^2).  The host adds them in row order:
// U = G/4 * sum u (each pair twice, log r = log(r^2)/2), K = 1/2 * sum k.
// Sources are staged through LDS 256 at a time; the loop runs over the
// system's real rows only.
__global__ __launch_bounds__(256) void k_energy(const double* __restrict__ x,
                                                const double* __restrict__ y,
                                                const double* __restrict__ m,
                                                const double* __restrict__ vx,
                                                const double* __restrict__ vy, int64_t p0,
                                                int64_t n, double eps2, double* __restrict__ u,
                                                double* __restrict__ k)
{
    __shared__ double sx[256], sy[256], sm[256];
    const int64_t li = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool act = li < n;
    const int64_t i = p0 + (act ? li : 0);
    const double xi = x[i], yi = y[i];
    double acc = 0.0;
    for (int64_t base = 0; base < n; base += 256) {
        const int64_t j = base + threadIdx.x;
        __syncthreads();
        if (j < n) {
            sx[threadIdx.x] = x[p0 + j];
            sy[threadIdx.x] = y[p0 + j];
            sm[threadIdx.x] = m[p0 + j];
        }
        __syncthreads();
        const int cnt = (int)min((int64_t)256, n - base);
        for (int jj = 0; jj < cnt; ++jj) {
            if (base + jj == li) continue;
            const double dx = sx[jj] - xi, dy = sy[jj] - yi;
            acc += sm[jj] * log(dx * dx + dy * dy + eps2);
        }
    }
    if (act) {
        u[i] = m[i] * acc;
        k[i] = m[i] * (vx[i] * vx[i] + vy[i] * vy[i]);
    }
}
