// Kernel declarations and tile constants shared by the .hip translation units.
#pragma once
#include "bh_common.h"

#define PLACE_CHUNK 2048   // particles per block in bbox / keys / scatter
#define HIST_BINS 8192     // LDS histogram window (keys) per block
#define SCAN_TILE 2048     // elements per block in the scans
#define SORT_W 1024        // within-cell rank window (rows)
#define SORT_THREADS 1024  // threads of a k_sort_units block (2 rows of the 2 SORT_W window each)

struct SortFields {
    const double *px, *py, *m, *vx, *vy, *ax, *ay;
    const int32_t* pid;
};

// Arguments of the force kernels (passed by value).
struct ForceArgs {
    const Item* items;
    const int32_t* order;      // queue position -> item (NULL: identity)
    int32_t* stats;
    const SysDev* sys;
    const int32_t* offsets;    // cell -> first sorted particle (C + 1)
    const int32_t* neprefix;   // cell -> first compact non-empty index (C + 1)
    const int32_t* key_sorted; // sorted particle -> global cell
    const double *x, *y, *m;
    const Com64* l64;
    const Com32* l32;
    double *ax, *ay;
    double Gc, eps2, r2cut;
    long long* cost_log;       // per item duration (wall clock ticks) or NULL
    // tail split (FP32 kernels): the last n_tail queue positions are each run
    // as `split` far-field parts, issued slice-major so that resident waves
    // share one slice of the COM list in L2; part holds their sums, one plane
    // of part_n particles x 2 doubles per slice (element ((s * part_n) + i) * 2),
    // added in slice order by k_force_combine.  split == 1: off.
    int split, n_tail;
    double* part;
    int64_t part_n;
    // this launch covers queue positions [total * qpart / nqparts,
    // total * (qpart + 1) / nqparts) (one launch: part 0 of 1)
    int qpart, nqparts;
    // FP64 recompute on (FP32 kernels): ax / ay point at zeroed scratch
    // planes of part_n doubles, and ay[part_n + i] receives particle i's sum
    // over its near-field terms of (m_j / r_ij)^2 (an upper bound of sum
    // |term|^2 under any softening or cutoff); the recompute then adds the
    // FP32 or the FP64 result to the particles' accelerations.  A flag rather
    // than a third pointer keeps the kernel's scalar registers unchanged.
    int rc;
};

// Arguments of the grid-method potential (bh_potential.hip).
struct PotArgs {
    const Item* items;
    int32_t* stats;
    const SysDev* sys;
    const int32_t* offsets;
    const int32_t* neprefix;
    const int32_t* key_sorted;
    const double *x, *y, *m, *vx, *vy;
    const Com64* l64;
    const Com32* l32;
    const double* msys;   // total mass per system
    double eps2;
    double *u, *k;        // m_i phi_i and m_i |v_i|^2 per particle
};

// m / q to FP64 accuracy without the IEEE division sequence: v_rcp_f64 gives
// r0 with e = 1 - q r0 small (|e| < 2^-22), and 1/q = r0 (1 + e + e^2 + e^3 + ...),
// so m r0 (1 + e + e^2) differs from m / q by the relative e^3 plus the
// roundings of m r0 and of the final fma: within about one ulp.  Three VALU
// instructions and the reciprocal, against ten and the reciprocal for the
// correctly rounded quotient.  q = 0 gives NaN (m / 0 * 0 is NaN as well).
__device__ __forceinline__ double wdiv64(double m, double q)
{
    const double r0 = __builtin_amdgcn_rcp(q);
    const double e = fma(-q, r0, 1.0);
    const double ee = fma(e, e, e);
    const double w0 = m * r0;
    return fma(w0, ee, w0);
}

// One FP64 weight m * w(r2) of the law (LAW 0: no cutoff; LAW 1: r2 <= r2cut
// gives 0).  FAST 1: the quotient by wdiv64 (the FP64 kernel's near field);
// FAST 0: the IEEE division (the fp64_ieee kernel).
template <int FAST>
__device__ __forceinline__ double qdiv64(double m, double q)
{
    return FAST ? wdiv64(m, q) : m / q;
}

template <int LAW, int FAST>
__device__ __forceinline__ double far_w64(double wx, double wy, double M, double eps2, double r2cut)
{
    if (LAW == 0) return qdiv64<FAST>(M, fma(wx, wx, fma(wy, wy, eps2)));
    const double s2 = fma(wx, wx, wy * wy);
    return (s2 > r2cut) ? qdiv64<FAST>(M, s2 + eps2) : 0.0;
}

template <int LAW, int FAST>
__device__ __forceinline__ double near_w64(double wx, double wy, double m, double eps2, double r2cut)
{
    const double s2 = fma(wx, wx, wy * wy);
    if (LAW == 0) return qdiv64<FAST>(m, s2 + eps2);
    return (s2 > r2cut) ? qdiv64<FAST>(m, s2 + eps2) : 0.0;
}

// grid-method potential
__global__ void k_sys_mass(const SysDev* sys, const int32_t* neprefix, const Com64* l64,
                           double* msys);
template <int P>
__global__ void k_potential(PotArgs A);

// placement
__global__ void k_fill_i32(int32_t* p, int64_t n, int32_t v);
__global__ void k_fill_f64(double* p, int64_t n, double v);
__global__ void k_bbox_init(uint64_t* bb, int S);
__global__ void k_bbox(const double* px, const double* py, const int32_t* sys_of, int64_t N,
                       uint64_t* bb, int32_t* stats);
__global__ void k_bbox_finalize(SysDev* sys, const uint64_t* bb, int S, int32_t* stats);
__global__ void k_keys_hist(const double* px, const double* py, const int32_t* sys_of,
                            const SysDev* sys, int64_t N, int32_t* key_out, int32_t* counts);
template <bool CNT>
__global__ void k_scan_reduce(const int32_t* in, int64_t n, int32_t* bsum, int32_t* bne,
                              int32_t* stats);
template <bool CNT>
__global__ void k_scan_blocks(int32_t* bsum, int32_t* bne, int64_t nb, int32_t* stats,
                              int total_slot);
template <bool CNT>
__global__ void k_scan_final(const int32_t* in, int64_t n, const int32_t* bsum, const int32_t* bne,
                             int32_t* out_v, int32_t* out_f);
__global__ void k_scatter(const int32_t* key, int64_t N, const int32_t* offsets, int32_t* cursor,
                          int32_t* idx_out, int32_t* key_sorted, const double* px,
                          const double* py, uint64_t* pxkey_out, uint64_t* pykey_out);
__global__ void k_sort_units(SortFields F, int64_t N, const int32_t* idx_in,
                             const int32_t* key_sorted, const uint64_t* pxkey,
                             const uint64_t* pykey, const int32_t* offsets, int32_t* idx_out);
__global__ void k_sort_merge(SortFields F, int64_t N, const int32_t* idx_in,
                             const int32_t* key_sorted, const int32_t* offsets, int32_t* idx_out,
                             int L, const int32_t* stats);
__global__ void k_gather(SortFields S, BhParticles D, int64_t N, const int32_t* idx_a,
                         const int32_t* idx_b, const int32_t* key_sorted, const int32_t* offsets,
                         const int32_t* stats, int32_t* perm_out);

// summarize + tick
__global__ void k_ne_list(const int32_t* counts, const int32_t* neprefix, int64_t C,
                          int32_t* ne_list);
__global__ void k_com(const SysDev* sys, int S, const int32_t* ne_list, const int32_t* stats,
                      const int32_t* offsets, const double* x, const double* y, const double* m,
                      double* com_dense, Com64* l64, Com32* l32);
__global__ void k_tick(double* x, double* y, double* vx, double* vy, double* ax, double* ay,
                       int64_t N, double dt);
__global__ void k_interactions(const SysDev* sys, int S, const int32_t* counts,
                               const int32_t* neprefix, int64_t C,
                               unsigned long long* out /* 2*S + 2 */);
__global__ void k_energy(const double* x, const double* y, const double* m, const double* vx,
                         const double* vy, int64_t p0, int64_t n, double eps2, double* u,
                         double* k);

// force
__global__ void k_clamp_items(int32_t* stats, int64_t max_items);
__global__ void k_force_combine(const Item* items, const int32_t* order, const int32_t* stats,
                                int split, int n_tail, int pslots, const double* part,
                                int64_t part_n, double* ax, double* ay);
// FP64 recompute of the particles whose FP32 rounding-error estimate is
// large against their acceleration
__global__ void k_rc_blocks(const SysDev* sys, int S, int64_t nblocks, const double* com_dense,
                            double* blk);
__global__ void k_rc_tiles(const SysDev* sys, int S, int64_t ntiles, const double* com_dense,
                           double* tl);
__global__ void k_rc_blockfar(const SysDev* sys, int S, int64_t nblocks, const double* blk,
                              double* bfar);
__global__ void k_rc_flag(const Item* items, const SysDev* sys, const int32_t* key_sorted,
                          const int32_t* neprefix, const Com64* l64, const double* x,
                          const double* y, const double* com, const double* tl,
                          const double* blk, const double* bfar, const double* a32x,
                          const double* a32y, const double* s2n, double Gc, double kappa2,
                          const int32_t* stats, int32_t* rc_cnt, int32_t* list, double* ax,
                          double* ay);
__global__ void k_rc_groups(const int32_t* rc_cnt, int S, int32_t* gofs, int32_t* stats);
template <int LAW>
__global__ void k_rc_fix64(const int32_t* stats, const int32_t* rc_cnt, const int32_t* gofs,
                           int S, const int32_t* list, const SysDev* sys,
                           const int32_t* key_sorted, const int32_t* offsets,
                           const int32_t* neprefix, const Com64* l64, const double* x,
                           const double* y, const double* m, double Gc, double eps2,
                           double r2cut, double* part, double* ax, double* ay);
__global__ void k_rc_combine(const int32_t* stats, const int32_t* rc_cnt, const int32_t* gofs,
                             int S, const int32_t* list, const SysDev* sys, const double* part,
                             double Gc, double* ax, double* ay);

template <bool WRITE>
__global__ void k_items(const SysDev* sys, int S, int64_t ncols, const int32_t* counts,
                        const int32_t* offsets, int pitem, int yspan, int32_t* col_items,
                        Item* items, int64_t max_items, int32_t* stats);
