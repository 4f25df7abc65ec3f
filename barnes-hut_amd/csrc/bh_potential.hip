// Grid-method potential: the energy diagnostic of runs too large for the
// direct O(N^2) pair sum.  No reference counterpart (python-bh reports no
// energy).  For particle i in cell (cx, cy)
//   phi_i = sum_{j != i in the 3x3 cells} m_j log(r_ij^2 + eps2)
//         + sum_{non-empty cells k outside the 3x3 block} M_k log(|X_k - x_i|^2 + eps2),
// the decomposition of the force kernels with log(s + eps2) in place of
// (x_j - x_i) / (s + eps2); U = G / 4 * sum_i m_i phi_i.  The work items,
// the masked-tile split and the cell-relative FP32 coordinates are those of
// k_force32: in cell units (positions relative to the item origin (cx, oy))
//   log(r^2 + eps2) = 2 ln(step) + ln 2 * log2(r_c^2 + eps2_c),
// so each term is one v_log_f32 and one fma, FP32 per 64-entry tile and FP64
// across tiles, and 2 ln(step) * (M_sys - m_i) is added once per particle in
// FP64 (the masses of the near particles and far cells add up to M_sys - m_i).
// Cost: about one FP32 force pass.
#include "bh_common.h"
#include "bh_kernels.h"

// Total mass per system: the sum of its non-empty cells' masses (one block of
// 256 threads per system, fixed reduction order).
__global__ __launch_bounds__(256) void k_sys_mass(const SysDev* __restrict__ sys,
                                                  const int32_t* __restrict__ neprefix,
                                                  const Com64* __restrict__ l64,
                                                  double* __restrict__ msys)
{
    __shared__ double red[256];
    const SysDev& d = sys[blockIdx.x];
    const int ne0 = neprefix[d.c_off], ne1 = neprefix[d.c_off + (int64_t)d.G * d.G];
    double s = 0.0;
    for (int e = ne0 + (int)threadIdx.x; e < ne1; e += 256) s += l64[e].M;
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) msys[blockIdx.x] = red[0];
}

struct NearP {
    double x, y, m;
    int32_t row, idx;
};

// u[i] = m_i phi_i and k[i] = m_i |v_i|^2 for every particle of the items.
template <int P>
__global__ __launch_bounds__(64) void k_potential(PotArgs A)
{
    __shared__ float4 ftile[BH_WAVE];          // (ex, ey, M, kx << 16 | ky as bits)
    __shared__ NearP ntile[BH_WAVE];
    const int lane = threadIdx.x;
    const int nitems = A.stats[ST_NITEMS];
    const double ln2 = 0.6931471805599453;
    for (;;) {
        int q = 0;
        if (lane == 0) q = atomicAdd(&A.stats[ST_QUEUE], 1);
        q = uniform_int(__shfl(q, 0, BH_WAVE));
        if (q >= nitems) break;
        const Item it = A.items[q];
        const SysDev& d = A.sys[it.sys];
        const int G = d.G;
        const int64_t coff = d.c_off;
        const int cx = it.cx;
        const int oy = (it.y0 + it.y1) >> 1;
        const int ylo = max(it.y0 - 1, 0), yhi = min(it.y1 + 1, G - 1);
        const double step = d.step;
        const double inv_step = 1.0 / step;
        const float eps2c = (float)(A.eps2 * inv_step * inv_step);

        double xi[P], yi[P], acc[P];
        float ux[P], uy[P];
        int cy[P], gi[P];
        bool act[P];
#pragma unroll
        for (int k = 0; k < P; ++k) {
            const int slot = lane + BH_WAVE * k;
            act[k] = slot < it.np;
            gi[k] = it.p0 + slot;
            xi[k] = yi[k] = 0.0;
            ux[k] = uy[k] = 0.0f;
            cy[k] = -100000;
            acc[k] = 0.0;
            if (act[k]) {
                xi[k] = A.x[gi[k]];
                yi[k] = A.y[gi[k]];
                cy[k] = (int)(A.key_sorted[gi[k]] - coff) - cx * G;
                ux[k] = (float)((xi[k] - d.min_x) / step - (double)cx);
                uy[k] = (float)((yi[k] - d.min_y) / step - (double)oy);
            }
        }

        // near field: the particles of columns cx-1..cx+1, rows ylo..yhi
        for (int dx = -1; dx <= 1; ++dx) {
            const int xc = cx + dx;
            if (xc < 0 || xc >= G) continue;
            const int64_t cl = coff + (int64_t)xc * G;
            const int pa = A.offsets[cl + ylo], pb = A.offsets[cl + yhi + 1];
            for (int base = pa; base < pb; base += BH_WAVE) {
                const int j = base + lane;
                NearP e;
                if (j < pb) {
                    e.x = A.x[j];
                    e.y = A.y[j];
                    e.m = A.m[j];
                    e.row = (int)(A.key_sorted[j] - cl);
                    e.idx = j;
                } else {
                    e.x = e.y = e.m = 0.0;
                    e.row = -100000;
                    e.idx = -1;
                }
                ntile[lane] = e;
                __syncthreads();
                const int cnt = min(BH_WAVE, pb - base);
                float t[P];
#pragma unroll
                for (int k = 0; k < P; ++k) t[k] = 0.0f;
                for (int jj = 0; jj < cnt; ++jj) {
                    const NearP f = ntile[jj];
                    const float fm = (float)f.m;
#pragma unroll
                    for (int k = 0; k < P; ++k) {
                        const bool keep = (f.idx != gi[k]) && (abs(f.row - cy[k]) <= 1);
                        const float wx = (float)((f.x - xi[k]) * inv_step);
                        const float wy = (float)((f.y - yi[k]) * inv_step);
                        const float s2 = fmaf(wx, wx, fmaf(wy, wy, eps2c));
                        const float v = fm * __builtin_amdgcn_logf(s2);
                        t[k] += keep ? v : 0.0f;
                    }
                }
#pragma unroll
                for (int k = 0; k < P; ++k) acc[k] += (double)t[k];
                __syncthreads();
            }
        }

        // far field: the system's compact COM list minus each particle's 3x3 block
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
                mlo[t] = mhi[t] = 0;
            }
        }
        auto stage = [&](int e) {
            float4 v;
            if (e < ne1) {
                const Com32 c = A.l32[e];
                const int kx = c.kxy >> 16, ky = c.kxy & 0xffff;
                v.x = (float)(kx - cx) + c.fx;
                v.y = (float)(ky - oy) + c.fy;
                v.z = (float)A.l64[e].M;
                v.w = __int_as_float(c.kxy);
            } else {   // zero mass far away: the term is exactly 0
                v.x = 1e18f;
                v.y = 1e18f;
                v.z = 0.0f;
                v.w = __int_as_float(0x7fff7fff);
            }
            return v;
        };
        float4 nxt = stage(ne0 + lane);
        for (int base = ne0; base < ne1; base += BH_WAVE) {
            ftile[lane] = nxt;
            __syncthreads();
            nxt = stage(base + BH_WAVE + lane);
            const int b1 = base + BH_WAVE;
            const bool masked = (mlo[0] < mhi[0] && base < mhi[0] && b1 > mlo[0]) ||
                                (mlo[1] < mhi[1] && base < mhi[1] && b1 > mlo[1]) ||
                                (mlo[2] < mhi[2] && base < mhi[2] && b1 > mlo[2]);
            float t[P];
#pragma unroll
            for (int k = 0; k < P; ++k) t[k] = 0.0f;
            if (!masked) {
#pragma unroll 8
                for (int j = 0; j < BH_WAVE; ++j) {
                    const float4 f = ftile[j];
#pragma unroll
                    for (int k = 0; k < P; ++k) {
                        const float wx = f.x - ux[k], wy = f.y - uy[k];
                        const float s2 = fmaf(wx, wx, fmaf(wy, wy, eps2c));
                        t[k] = fmaf(f.z, __builtin_amdgcn_logf(s2), t[k]);
                    }
                }
            } else {
                for (int j = 0; j < BH_WAVE; ++j) {
                    const float4 f = ftile[j];
                    const int kxy = __float_as_int(f.w);
                    const bool colnear = abs((kxy >> 16) - cx) <= 1;
                    const int ky = kxy & 0xffff;
#pragma unroll
                    for (int k = 0; k < P; ++k) {
                        const bool keep = !(colnear && abs(ky - cy[k]) <= 1);
                        const float wx = f.x - ux[k], wy = f.y - uy[k];
                        const float s2 = fmaf(wx, wx, fmaf(wy, wy, eps2c));
                        const float v = f.z * __builtin_amdgcn_logf(s2);
                        t[k] += keep ? v : 0.0f;
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < P; ++k) acc[k] += (double)t[k];
            __syncthreads();
        }

        const double shift = 2.0 * log(step);
        const double msys = A.msys[it.sys];
#pragma unroll
        for (int k = 0; k < P; ++k) {
            if (act[k]) {
                const double mi = A.m[gi[k]];
                const double phi = ln2 * acc[k] + shift * (msys - mi);
                A.u[gi[k]] = mi * phi;
                const double vx = A.vx[gi[k]], vy = A.vy[gi[k]];
                A.k[gi[k]] = mi * (vx * vx + vy * vy);
            }
        }
    }
}

template __global__ void k_potential<1>(PotArgs);
template __global__ void k_potential<2>(PotArgs);
template __global__ void k_potential<4>(PotArgs);
