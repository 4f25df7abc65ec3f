// Shared device-side definitions for the grid-method kernels (gfx950).
//
// Data layout in HBM (all structure-of-arrays, one slot per particle, the
// particles of system s occupying [p_off, p_off + n)):
//   px, py, mass, vx, vy, ax, ay : double
//   pid                           : int32  (original particle index)
// Cells of all systems share one global index space: system s owns global
// cells [c_off, c_off + G*G), cell (gx, gy) -> c_off + gx*G + gy (x-major, the
// order of python-bh/barneshut/implementations/sequential.py:37,39-50).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define BH_WAVE 64

// Error bits raised by kernels into ctx->stats[ST_ERR].
#define BH_ERR_NONFINITE 1      // a particle coordinate is NaN or infinite
#define BH_ERR_DEGENERATE 2     // bounding box of zero extent (step <= 0)
#define BH_ERR_ITEMS 4          // work-item capacity exceeded

// Indices into the int32 stats block.
enum {
    ST_ERR = 0,
    ST_MAXCOUNT = 1,   // largest cell occupancy (all systems)
    ST_NITEMS = 2,     // number of force work items
    ST_QUEUE = 3,      // persistent-kernel queue cursor
    ST_NCELLS_NE = 4,  // number of non-empty cells (all systems)
    ST_NPASS = 5,      // merge passes the within-cell sort needed
    ST_NRECOMP = 6,    // particles flagged for the FP64 recompute (last FP32 pass)
    ST_RCGROUPS = 7,   // their groups (up to RC_GROUP particles of one system)
    ST_COUNT = 16
};

// Per-system parameters. Host fills p_off, n, c_off, col_off, b_off, G, Gb;
// the bounding-box kernels fill the FP64 box fields every step.
struct SysDev {
    int64_t p_off;
    int64_t n;
    int64_t c_off;
    int64_t col_off;
    int64_t b_off;   // first of the system's Gb x Gb blocks of RC_BLOCK x RC_BLOCK cells
    int32_t G;
    int32_t Gb;
    double min_x, min_y, max_x, max_y, step;
};

#define RC_BLOCK 16   // cells per side of the blocks the recompute bound aggregates
#define RC_EXACT 3    // cells around a work item whose terms the bound takes exactly
#define RC_TILE 4     // cells per side of the bound's middle level (divides RC_BLOCK)
#define RC_GROUP 8        // flagged particles evaluated together by one FP64 recompute block
#define RC_THREADS 256    // threads of that block
#define RC_SLICES 8       // most COM-list slices per group (when groups are few)
#define RC_PART_CAP 16384 // (group, slice) partial sums the scratch holds

// Compact list of non-empty cells (one entry per non-empty cell, in global
// cell order), FP64 physical COM.
struct Com64 {
    double X, Y, M;
    int32_t kx, ky;
};

// FP32 variant: COM expressed in cell units relative to its own cell corner
// (X = (kx + fx) * step + min_x), and the reciprocal of the cell's mass
// rM = 1 / (float)M (IEEE division, once per cell and step; +inf for M = 0):
// the FP32 kernels need s / M for a power of two s, which is s * rM exactly.
struct Com32 {
    float fx, fy, rM;
    int32_t kxy;   // (kx << 16) | ky; kx, ky < 32768 (bh_set_systems caps grid_dim)
};

// A force work item: a run of sorted particles [p0, p0 + np) that all lie in
// column cx of system sys, in rows [y0, y1].
struct Item {
    int32_t p0, np, sys, cx;
    int32_t y0, y1, pad0, pad1;
};

__device__ __forceinline__ uint64_t bh_ord_key(double d)
{
    uint64_t u = (uint64_t)__double_as_longlong(d);
    return (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
}

__device__ __forceinline__ double bh_ord_val(uint64_t k)
{
    uint64_t u = (k & 0x8000000000000000ull) ? (k & ~0x8000000000000000ull) : ~k;
    return __longlong_as_double((long long)u);
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, BH_WAVE);
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_min(T v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        T w = __shfl_xor(v, o, BH_WAVE);
        v = w < v ? w : v;
    }
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_max(T v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        T w = __shfl_xor(v, o, BH_WAVE);
        v = w > v ? w : v;
    }
    return v;
}

__device__ __forceinline__ int uniform_int(int v)
{
    return __builtin_amdgcn_readfirstlane(v);
}

// System of global particle i / global cell c: binary search over the
// host-written offsets (S is at most a few thousand).
__device__ __forceinline__ int sys_of_cell(const SysDev* sys, int S, int64_t c)
{
    int lo = 0, hi = S - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (sys[mid].c_off <= c) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ int sys_of_col(const SysDev* sys, int S, int64_t col)
{
    int lo = 0, hi = S - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (sys[mid].col_off <= col) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Kernel entry points (defined in the .hip files, launched from bh_abi.hip).
struct BhParticles {
    double *px, *py, *m, *vx, *vy, *ax, *ay;
    int32_t* pid;
};
