// Cloud-to-cloud direct gravity: the reference's gravity plugin
//   fn(self_cloud, other_cloud, G, update_other)
// (python-bh/barneshut/kernels/gravity/cpu_numba.py:8-92, selected by
// gravity/__init__.py:6-41 and called from internals/cloud.py:113-114).
//
// A cloud is n rows of ld doubles in the reference's particle row layout
// (px, py, mass, vx, vy, ax, ay, ...; internals/particle.py:3-15), in device
// memory.  For every target row i the kernel adds
//     G * sum_j m_j (x_j - x_i) w(|x_j - x_i|^2)
// over the source rows j (j != i when targets and sources are one cloud) to
// the row's ax, ay, with the law's w (bh_law: cutoff r2_cut > 0 drops pairs
// with r^2 <= r2_cut, as the reference's CPU kernel does at 1e-5).
//
// One 256-thread block owns P * G' consecutive targets (G' = 256 / L groups of
// L lanes, P = 1 or 2 targets per lane).  Sources are staged through LDS in tiles of
// 256 rows; lane l of a group takes the tile rows l, l + L, l + 2L, ... for
// its P targets (with one group per block, L = 256, the lanes read their rows
// straight from memory instead), and the L partial sums of a target are added
// by a fixed tree.  The shape (P, L) is a function of the two cloud sizes only, so a
// result never depends on timing.  FP64 throughout; each weight m / q by the
// refined hardware reciprocal (wdiv64, within about one ulp of the quotient).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>

#include "bh_common.h"
#include "bh_kernels.h"
#include "../../include/bh_hip.h"

int bh_ctx_fail(int code, const char* msg);

#define CL_THREADS 256
#define CL_MIN_BLOCKS 2048   // blocks a launch aims for before lanes are widened
#define CL_MIN_PER_LANE 8    // fewest source rows a lane is given when widening

namespace {

// Reference row columns (internals/particle.py:3-15).
constexpr int COL_X = 0, COL_Y = 1, COL_M = 2, COL_AX = 5, COL_AY = 6;

struct CloudArgs {
    const double* src;   // source rows (px, py, mass read)
    int64_t ns, lds;
    double* tgt;         // target rows (px, py read; ax, ay added to)
    int64_t nt, ldt;
    int32_t mask_self;   // 1: row i of the sources is target i, skip it by index
    int32_t llog2;       // log2 of the lanes per target
    double G, eps2, r2cut;
};

// m w(r^2) of the law: LAW 1 drops r^2 <= r2cut; SOFT 0 has eps2 = 0 (the
// reference laws), so r^2 + eps2 is not formed.
template <int LAW, int SOFT>
__device__ __forceinline__ double cloud_w(double wx, double wy, double m, double eps2,
                                          double r2cut)
{
    const double s2 = fma(wx, wx, wy * wy);
    const double w = wdiv64(m, SOFT ? s2 + eps2 : s2);
    return LAW ? ((s2 > r2cut) ? w : 0.0) : w;
}

template <int LAW, int SOFT, int P, int MASK>
__device__ __forceinline__ void cloud_tile(const double* sx, const double* sy, const double* sm,
                                           int cnt, int lane, int L, int64_t t0,
                                           const int64_t* ti, const double* xi, const double* yi,
                                           double* ax, double* ay, double eps2, double r2cut)
{
#pragma unroll 2
    for (int k = lane; k < cnt; k += L) {
        const double x = sx[k], y = sy[k], m = sm[k];
#pragma unroll
        for (int p = 0; p < P; ++p) {
            const double wx = x - xi[p], wy = y - yi[p];
            double w = cloud_w<LAW, SOFT>(wx, wy, m, eps2, r2cut);
            if (MASK) w = (t0 + k == ti[p]) ? 0.0 : w;
            ax[p] = fma(wx, w, ax[p]);
            ay[p] = fma(wy, w, ay[p]);
        }
    }
}

// Source rows lane, lane + 256, ... read straight from memory (L = 256).
template <int LAW, int SOFT, int P, int MASK>
__device__ __forceinline__ void cloud_direct(const CloudArgs& A, int lane, const int64_t* ti,
                                             const double* xi, const double* yi, double* ax,
                                             double* ay)
{
#pragma unroll 4
    for (int64_t j = lane; j < A.ns; j += CL_THREADS) {
        const double* r = A.src + j * A.lds;
        const double x = r[COL_X], y = r[COL_Y], m = r[COL_M];
#pragma unroll
        for (int p = 0; p < P; ++p) {
            const double wx = x - xi[p], wy = y - yi[p];
            double w = cloud_w<LAW, SOFT>(wx, wy, m, A.eps2, A.r2cut);
            if (MASK) w = (j == ti[p]) ? 0.0 : w;
            ax[p] = fma(wx, w, ax[p]);
            ay[p] = fma(wy, w, ay[p]);
        }
    }
}

template <int LAW, int SOFT, int P>
__global__ __launch_bounds__(CL_THREADS) void k_cloud(CloudArgs A)
{
    // tiles use 3 x 256 doubles, the final tree 2 P x 256
    __shared__ double buf[4 * CL_THREADS];
    double* sx = buf;
    double* sy = buf + CL_THREADS;
    double* sm = buf + 2 * CL_THREADS;
    const int tid = threadIdx.x;
    const int L = 1 << A.llog2;
    const int lane = tid & (L - 1);
    const int groups = CL_THREADS >> A.llog2;
    const int grp = tid >> A.llog2;
    const int64_t base = (int64_t)blockIdx.x * (P * groups);

    int64_t ti[P];
    double xi[P], yi[P], ax[P], ay[P];
#pragma unroll
    for (int p = 0; p < P; ++p) {
        ti[p] = base + (int64_t)p * groups + grp;
        const bool live = ti[p] < A.nt;
        xi[p] = live ? A.tgt[ti[p] * A.ldt + COL_X] : 0.0;
        yi[p] = live ? A.tgt[ti[p] * A.ldt + COL_Y] : 0.0;
        ax[p] = 0.0;
        ay[p] = 0.0;
    }

    if (groups == 1) {
        // one target group per block: a tile would serve only its P targets,
        // so each lane reads its own source rows (unrolled, loads in flight)
        if (A.mask_self)
            cloud_direct<LAW, SOFT, P, 1>(A, lane, ti, xi, yi, ax, ay);
        else
            cloud_direct<LAW, SOFT, P, 0>(A, lane, ti, xi, yi, ax, ay);
    }
    for (int64_t t0 = 0; groups > 1 && t0 < A.ns; t0 += CL_THREADS) {
        __syncthreads();   // every lane is done with the previous tile
        const int64_t j = t0 + tid;
        if (j < A.ns) {
            const double* r = A.src + j * A.lds;
            sx[tid] = r[COL_X];
            sy[tid] = r[COL_Y];
            sm[tid] = r[COL_M];
        }
        __syncthreads();
        const int cnt = (int)std::min<int64_t>(CL_THREADS, A.ns - t0);
        if (A.mask_self)
            cloud_tile<LAW, SOFT, P, 1>(sx, sy, sm, cnt, lane, L, t0, ti, xi, yi, ax, ay,
                                        A.eps2, A.r2cut);
        else
            cloud_tile<LAW, SOFT, P, 0>(sx, sy, sm, cnt, lane, L, t0, ti, xi, yi, ax, ay,
                                        A.eps2, A.r2cut);
    }

    __syncthreads();
#pragma unroll
    for (int p = 0; p < P; ++p) {
        buf[(2 * p) * CL_THREADS + tid] = ax[p];
        buf[(2 * p + 1) * CL_THREADS + tid] = ay[p];
    }
    __syncthreads();
    for (int s = L >> 1; s > 0; s >>= 1) {
        if (lane < s) {
#pragma unroll
            for (int q = 0; q < 2 * P; ++q)
                buf[q * CL_THREADS + tid] += buf[q * CL_THREADS + tid + s];
        }
        __syncthreads();
    }
    if (lane == 0) {
#pragma unroll
        for (int p = 0; p < P; ++p) {
            if (ti[p] < A.nt) {
                double* r = A.tgt + ti[p] * A.ldt;
                r[COL_AX] += A.G * buf[(2 * p) * CL_THREADS + tid];
                r[COL_AY] += A.G * buf[(2 * p + 1) * CL_THREADS + tid];
            }
        }
    }
}

inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Shape of one launch from the two sizes: 2 targets per lane (one LDS read
// and the loop's bookkeeping serve both) once there are 2 per thread of a
// block, and more lanes per target while blocks are fewer than CL_MIN_BLOCKS
// and every lane keeps at least CL_MIN_PER_LANE source rows.
void cloud_shape(int64_t nt, int64_t ns, int* P, int* llog2)
{
    *P = nt >= 2LL * CL_THREADS ? 2 : 1;
    int l = 0;
    while (l < 8 && cdiv(nt, (int64_t)*P * (CL_THREADS >> l)) < CL_MIN_BLOCKS &&
           ((int64_t)CL_MIN_PER_LANE << (l + 1)) <= ns)
        ++l;
    *llog2 = l;
}

int launch_cloud(const bh_cloud* tgt, const bh_cloud* src, int mask_self, const bh_law* law,
                 hipStream_t st)
{
    CloudArgs A;
    A.src = src->rows;
    A.ns = src->n;
    A.lds = src->ld;
    A.tgt = tgt->rows;
    A.nt = tgt->n;
    A.ldt = tgt->ld;
    A.mask_self = mask_self;
    A.G = law->G;
    A.eps2 = law->eps2;
    A.r2cut = law->r2_cut;
    int P, l;
    cloud_shape(A.nt, A.ns, &P, &l);
    A.llog2 = l;
    const int64_t blocks = cdiv(A.nt, (int64_t)P * (CL_THREADS >> l));
    if (blocks > 0x7fffffffLL) return bh_ctx_fail(BH_EINVAL, "cloud too large for one launch");
    typedef void (*Kern)(CloudArgs);
    // [cutoff][softened][P == 2]
    static const Kern table[2][2][2] = {
        {{k_cloud<0, 0, 1>, k_cloud<0, 0, 2>}, {k_cloud<0, 1, 1>, k_cloud<0, 1, 2>}},
        {{k_cloud<1, 0, 1>, k_cloud<1, 0, 2>}, {k_cloud<1, 1, 1>, k_cloud<1, 1, 2>}}};
    const Kern k = table[law->r2_cut > 0.0][law->eps2 > 0.0][P == 2];
    hipLaunchKernelGGL(k, dim3((unsigned)blocks), dim3(CL_THREADS), 0, st, A);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : bh_ctx_fail(BH_EHIP, hipGetErrorString(e));
}

int check_cloud(const bh_cloud* c, const char* name)
{
    char msg[160];
    if (c->n < 0 || c->ld < 7 || (c->n > 0 && !c->rows) ||
        (c->n > 0 && c->n > INT64_MAX / c->ld)) {
        snprintf(msg, sizeof msg, "%s cloud needs n >= 0, ld >= 7 and rows (n=%lld, ld=%lld)",
                 name, (long long)c->n, (long long)c->ld);
        return bh_ctx_fail(BH_EINVAL, msg);
    }
    return 0;
}

}  // namespace

extern "C" int bh_cloud_gravity(const bh_cloud* self, const bh_cloud* other, int32_t same,
                                const bh_law* law, int32_t update_other, void* stream)
{
    if (!self || !law || (!same && !other)) return bh_ctx_fail(BH_EINVAL, "NULL argument");
    if (!(law->eps2 >= 0.0) || !(law->r2_cut >= 0.0) || !std::isfinite(law->G))
        return bh_ctx_fail(BH_EINVAL, "law needs finite G, eps2 >= 0, r2_cut >= 0");
    int rc = check_cloud(self, "self");
    if (rc) return rc;
    if (!same && (rc = check_cloud(other, "other"))) return rc;
    hipStream_t st = (hipStream_t)stream;
    if (same) {
        // the self term has r^2 = 0: the cutoff drops it and softening makes
        // it 0 * w; only the plain law (0 / 0) needs it skipped by index
        const int mask = !(law->r2_cut > 0.0) && !(law->eps2 > 0.0);
        return self->n > 0 ? launch_cloud(self, self, mask, law, st) : 0;
    }
    if (self->n == 0 || other->n == 0) return 0;
    if ((rc = launch_cloud(self, other, 0, law, st))) return rc;
    if (update_other) return launch_cloud(other, self, 0, law, st);
    return 0;
}
