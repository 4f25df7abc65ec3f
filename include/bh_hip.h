/*
 * bh_hip.h -- C ABI of the MI355X (gfx950) grid-method step library
 * (libbh_hip.so).  Plain pointers and sizes only; all particle buffers are
 * device pointers owned by the caller, scratch and per-step state are owned
 * by the context.  Every call returns BH_OK (0) or an error code; the text of
 * the last error of the calling thread is bh_last_error().  Launches are
 * asynchronous on the given stream (a hipStream_t, NULL = default stream).
 * A context is used by one host thread at a time.
 *
 * Reference interfaces replaced (python-bh/...):
 *   bh_place_particles     set_particles_bounding_box  barneshut/implementations/base.py:211-239
 *                          g_place_particles          barneshut/kernels/grid_decomposition/gpu/grid.py:15-50
 *   bh_calculate_box_cumm  g_calculate_box_cumm       .../gpu/grid.py:52-60
 *   bh_sort_particles      g_sort_particles           .../gpu/grid.py:62-76 (order of sequential.py:37)
 *   bh_summarize           g_summarize                .../gpu/grid.py:78-114
 *   bh_evaluate_boxes      g_evaluate_boxes           .../gpu/grid.py:116-156
 *   bh_tick_particles      g_tick_particles           .../gpu/grid.py:159-174
 *   bh_step / bh_step_graph  one BaseBarnesHut.run round, base.py:94-135
 *   bh_create/bh_set_systems  SingleGPUBarnesHut.__init_device_arrays, singlegpu.py:28-39
 */
#ifndef BH_HIP_H
#define BH_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    BH_OK = 0,
    BH_EINVAL = 1,   /* bad argument */
    BH_EHIP = 2,     /* HIP runtime error */
    BH_EDEVICE = 3,  /* a kernel raised an error flag (see bh_read_stats) */
    BH_ESTATE = 4    /* call made before bh_set_systems */
};

/* Device-side error bits reported by bh_read_stats()[0]. */
#define BH_DEVERR_NONFINITE 1   /* NaN/inf particle coordinate */
#define BH_DEVERR_DEGENERATE 2  /* zero-extent bounding box */
#define BH_DEVERR_ITEMS 4       /* work-item capacity exceeded */

enum {
    BH_PREC_FP64 = 0,            /* plain FP64 forces; each weight m / (r^2 + eps2) from the
                                    hardware reciprocal with one third-order refinement
                                    (within about one ulp of the quotient) */
    BH_PREC_FP32_KAHAN = 1,      /* FP32 cell-relative, Kahan-summed 64-term tiles */
    BH_PREC_FP32_KAHAN_TERM = 2, /* FP32 cell-relative, Kahan per term */
    BH_PREC_FP32_PLAIN = 3,      /* FP32 cell-relative, uncompensated */
    BH_PREC_FP32_MFMA = 4,       /* FP32 cell-relative, far-field sums on the matrix cores,
                                    Kahan-summed 64-term tiles */
    BH_PREC_FP64_IEEE = 5        /* plain FP64 forces with the IEEE (correctly rounded)
                                    division per weight */
};

/* SoA particle buffers (device), length = total particles of all systems. */
typedef struct bh_particles {
    double *px, *py, *mass, *vx, *vy, *ax, *ay;
    int32_t* pid;
} bh_particles;

/* a_i += G * sum m_j (x_j - x_i) * w(r2),  w(s) = 1/(s + eps2), and
 * w(s) = 0 for s <= r2_cut when r2_cut > 0.
 * GPU reference: eps2 = 0, r2_cut = 0; CPU reference: r2_cut = 1e-5. */
typedef struct bh_law {
    double G;
    double eps2;
    double r2_cut;
} bh_law;

/* Launch shape of the force kernel. p: particles per lane (1, 2 or 4);
 * yspan: max rows per work item (>= 1); waves_per_cu: persistent waves per
 * compute unit (>= 1); tail_split: the FP32 kernels run each of the last
 * tail_items queue entries as tail_split far-field slices (1 = off, <= 16;
 * tail_items >= the item count splits every item), so the queue drains in
 * finer pieces; the slices' sums are added in a fixed order by a second
 * kernel (scratch: particles x tail_split x 16 bytes).
 * Default: p 2, yspan 512, waves_per_cu 32, tail_split 16, tail_items 2^20. */
typedef struct bh_tuning {
    int32_t p;
    int32_t yspan;
    int32_t waves_per_cu;
    int32_t tail_split;
    int32_t tail_items;
} bh_tuning;

/* What bh_export copies out. */
enum {
    BH_EXPORT_COUNTS = 0,     /* int32 [n_cells]   particles per global cell          */
    BH_EXPORT_OFFSETS = 1,    /* int32 [n_cells+1] first sorted particle of each cell */
    BH_EXPORT_NEPREFIX = 2,   /* int32 [n_cells+1] compact index of each cell         */
    BH_EXPORT_PERM = 3,       /* int32 [n]         sorted position -> source row      */
    BH_EXPORT_KEYS = 4,       /* int32 [n]         global cell of each sorted row     */
    BH_EXPORT_COM = 5,        /* double [n_cells*3] X, Y, M per cell (0 if empty)     */
    BH_EXPORT_BBOX = 6,       /* double [5*S] min_x, min_y, max_x, max_y, step        */
    BH_EXPORT_ITEMS = 7,      /* int32 [n_items*8] p0, np, sys, cx, y0, y1, 0, 0      */
    BH_EXPORT_COST = 8,       /* int64 [n_items]   per-item wall-clock ticks (100 MHz)*/
    BH_EXPORT_KEYS_IN = 9     /* int32 [n]         global cell of each input row      */
};

const char* bh_last_error(void);
int bh_version(void);

int bh_create(int device, void** ctx_out);
int bh_destroy(void* ctx);

/* Defines the independent systems (host arrays of length nsys): system s
 * holds n[s] >= 1 particles at rows [sum_{t<s} n[t], ...) and a
 * grid_dim[s]^2 grid, 1 <= grid_dim[s] <= 32767 (the FP32 COM list packs a
 * cell's column and row in 16 bits each; such a grid already needs ~170 B of
 * per-cell state, ~180 GB). (Re)allocates the context's scratch. */
int bh_set_systems(void* ctx, int32_t nsys, const int64_t* n, const int32_t* grid_dim);
int bh_set_tuning(void* ctx, const bh_tuning* t);
int bh_get_tuning(void* ctx, bh_tuning* t);
/* 1: the force kernel logs each work item's duration (BH_EXPORT_COST). */
int bh_set_cost_log(void* ctx, int32_t enable);

int bh_place_particles(void* ctx, const bh_particles* p, void* stream);
int bh_calculate_box_cumm(void* ctx, void* stream);
int bh_sort_particles(void* ctx, const bh_particles* src, const bh_particles* dst, void* stream);
int bh_summarize(void* ctx, const bh_particles* p, void* stream);
/* order: device int32 [n_items] queue order of the work items, or NULL for
 * the static (natural) order; it must come from this step's items. */
int bh_build_items(void* ctx, void* stream);
int bh_evaluate_boxes(void* ctx, const bh_particles* p, const bh_law* law, int32_t precision,
                      const int32_t* order, void* stream);
/* The force kernel alone, on the items of the last bh_build_items (order as
 * in bh_evaluate_boxes; NULL = static order).  bh_evaluate_boxes with a NULL
 * order is bh_build_items followed by this call. */
int bh_force_pass(void* ctx, const bh_particles* p, const bh_law* law, int32_t precision,
                  const int32_t* order, void* stream);
/* bh_force_pass in nparts launches: call it for part = 0 .. nparts-1 in
 * order (same arguments otherwise); part k runs the k-th contiguous range of
 * the work queue, the last part adds the slices and runs the recompute.  The
 * result is bit-identical to bh_force_pass (a particle's sums do not depend
 * on which launch runs its pieces).  For callers that need to return to the
 * host during a long pass (progress, deadlines).  No reference counterpart. */
int bh_force_pass_part(void* ctx, const bh_particles* p, const bh_law* law, int32_t precision,
                       const int32_t* order, int32_t part, int32_t nparts, void* stream);
int bh_tick_particles(void* ctx, const bh_particles* p, double dt, void* stream);

/* FP64 recompute after an FP32 pass (precisions 1-3).  No reference
 * counterpart: the reference computes in FP64 throughout; this keeps the FP32
 * path's per-particle error small where a particle's terms nearly cancel.
 * After the FP32 kernel, every particle gets an upper bound B of
 * sqrt(sum_j |term_j|^2) (near terms summed exactly, far cells bounded by
 * their distance to the particle's work item); the FP32 rounding error grows
 * like eps32 * B, so particles with B > kappa * |a| (a in the law's units
 * divided by G / step) are evaluated again with the FP64 kernel's terms and
 * their FP32 result is replaced.  bh_read_stats()[6] is the number of
 * particles recomputed by the last pass.  kappa <= 0 turns it off (then the
 * FP32 pass writes no per-particle scratch).  deferred = 0: bh_force_pass,
 * bh_evaluate_boxes and the step calls run it; deferred = 1: bh_force_pass
 * stops after the FP32 kernels and the caller runs bh_recompute before
 * anything else touches the particles (the step calls still run it).
 * Default: kappa 1.5, deferred 0. */
int bh_set_recompute(void* ctx, double kappa, int32_t deferred);
/* The recompute stage alone, for the last FP32 bh_force_pass on p (same law);
 * BH_ESTATE if that pass did not prepare it. */
int bh_recompute(void* ctx, const bh_particles* p, const bh_law* law, void* stream);

/* One full round: a is placed and sorted into b, then summarize, evaluate
 * (order NULL: static schedule) and, if do_tick, tick -- all on b. */
int bh_step(void* ctx, const bh_particles* a, const bh_particles* b, const bh_law* law,
            int32_t precision, double dt, int32_t do_tick, void* stream);
/* Same as bh_step, replayed from a captured HIP graph (captured on first use
 * for a given (a, b, law, precision, dt, do_tick, tuning)). */
int bh_step_graph(void* ctx, const bh_particles* a, const bh_particles* b, const bh_law* law,
                  int32_t precision, double dt, int32_t do_tick, void* stream);

/* bh_step with the learned queue order: after the items are built,
 * bh_learned_order(params -> order) runs and the force pass follows that
 * order.  order: device int32 of at least the work-item capacity
 * (bh_query 4).  graph != 0 replays it from a captured HIP graph. */
int bh_step_learned(void* ctx, const bh_particles* a, const bh_particles* b, const bh_law* law,
                    int32_t precision, double dt, int32_t do_tick, const float* params,
                    int32_t* order, int32_t graph, void* stream);

/* Synchronises the stream, copies the int32 stats block (16 entries:
 * [0] error bits, [1] max cell count, [2] work items, [4] non-empty cells,
 * [6] particles recomputed in FP64 by the last FP32 pass, [7] their groups)
 * to host memory out16, and returns BH_EDEVICE if an error bit is set. */
int bh_read_stats(void* ctx, int32_t* out16, void* stream);
/* Copies state `what` into dst (device or host memory), count elements. */
int bh_export(void* ctx, int32_t what, void* dst, int64_t count, void* stream);
/* Exact interaction counts of the current placement: host out[2*S+2] =
 * (near_s, far_s) per system, then the totals. Synchronises. */
int bh_interactions(void* ctx, int64_t* out, void* stream);
/* Same counts without a host sync: device uint64 dev_out[2*S+2] receives
 * sum_c n_c * (particles in the 3x3 block) (self pairs included: subtract
 * each system's N for the near count) and the far counts. */
int bh_interactions_async(void* ctx, uint64_t* dev_out, void* stream);
/* Energy diagnostic; the reference reports no energy, so there is no
 * counterpart (it serves the energy-drift report of long runs).  Per system s:
 * out[2s] = U_s = G * sum_{i<j} m_i m_j * log(r_ij^2 + eps2) / 2, the potential
 * of the law a_i = G sum_j m_j (x_j - x_i) / (r^2 + eps2), summed directly in
 * FP64 (O(n_s^2) work: a diagnostic for n_s up to a few million), and
 * out[2s+1] = K_s = sum m v^2 / 2.  Rows of p in the context's row order.
 * Synchronises. */
int bh_energy(void* ctx, const bh_particles* p, double G, double eps2, double* out, void* stream);
/* The same energy with the potential evaluated by the grid method (no
 * reference counterpart either): per system, out[2s] = G / 4 * sum_i m_i phi_i
 * with phi_i = sum over the other particles of i's 3x3 cells of
 * m_j log(r_ij^2 + eps2) plus sum over the other non-empty cells k of
 * M_k log(|X_k - x_i|^2 + eps2) -- the force kernels' decomposition with the
 * log in place of the reciprocal, evaluated in FP32 cell-relative terms
 * (relative error ~1e-7), about one FP32 force pass of work -- and out[2s+1]
 * = K_s.  p must be placed and summarized (bh_place_particles ...
 * bh_summarize, or a step); the call rebuilds the work items.  Synchronises. */
int bh_energy_grid(void* ctx, const bh_particles* p, double G, double eps2, double* out,
                   void* stream);

/* Learned cost model (MLP 8 -> 32 -> 32 -> 1, FP32, ReLU).
 * params: device float[BH_MLP_NPARAM] = W1[32x8], b1[32], W2[32x32], b2[32],
 * W3[32], b3[1] (row-major, out x in). */
#define BH_MLP_NPARAM 1377
#define BH_MLP_NFEAT 8
/* Device features [n_items x 8] of the current work items (build_items first). */
int bh_item_features(void* ctx, float* feat, void* stream);
int bh_mlp_forward(const float* params, const float* x, int64_t n, float* y, void* stream);
/* Mean-squared-error gradient over n rows: grad[BH_MLP_NPARAM] = d/dparams
 * of mean((y - target)^2); loss (device float[1]) optional. */
int bh_mlp_backward(const float* params, const float* x, const float* target, int64_t n,
                    float* grad, float* loss, void* stream);
/* Adam update: m, v moments (device float[BH_MLP_NPARAM]), step t >= 1. */
int bh_mlp_adam(float* params, const float* grad, float* m, float* v, int32_t t, float lr,
                void* stream);
/* Queue order from predicted costs (device float[n_items]): longest first,
 * by 4096 buckets between the smallest and largest prediction (order inside
 * a bucket is unspecified). Writes device int32 order[n_items]. A non-finite
 * prediction makes the whole order fall back to the static order. */
int bh_order_by_cost(void* ctx, const float* cost, int32_t* order, void* stream);
/* Features -> MLP forward -> bh_order_by_cost for the current items, all on
 * the device (no host sync; capturable). */
int bh_learned_order(void* ctx, const float* params, int32_t* order, void* stream);

/* Context sizes: what = 0 particles, 1 cells, 2 systems, 3 columns,
 * 4 work-item capacity, 5 compute units. */
int bh_query(void* ctx, int32_t what, int64_t* out);

/* Cloud-to-cloud direct gravity: the reference's gravity plugin
 * fn(self_cloud, other_cloud, G, update_other)
 * (python-bh/barneshut/kernels/gravity/cpu_numba.py:8-92; registry
 * gravity/__init__.py:6-41; caller Cloud.apply_force, internals/cloud.py:113-114).
 * A cloud is n rows of ld >= 7 doubles in device memory, in the reference's
 * particle row layout (px, py, mass, vx, vy, ax, ay, ...:
 * internals/particle.py:3-15). */
typedef struct bh_cloud {
    double* rows;
    int64_t n;
    int64_t ld;
} bh_cloud;
/* Adds to every self row's ax, ay: G * sum_j m_j (x_j - x_i) w(r2) over the
 * other rows j, with law's w (G is law->G; the reference's CPU kernels use
 * r2_cut = 1e-5, eps2 = 0).  same != 0: self and other are one cloud (other
 * may be NULL), every j != i, and update_other is ignored
 * (cpu_numba.py:9-13,30-47).  Otherwise, update_other != 0 also adds to every
 * other row's ax, ay the sum over the self rows (cpu_numba.py:22-25,49-70);
 * 0 updates self only (:26-28,72-92).  Both sums are taken from the positions
 * the call starts with.  An empty cloud is a no-op.  Asynchronous on stream;
 * the result of a call depends only on its inputs and the two sizes. */
int bh_cloud_gravity(const bh_cloud* self, const bh_cloud* other, int32_t same,
                     const bh_law* law, int32_t update_other, void* stream);

/* Allocation-free step.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory -- code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc.  bh_step for ONE system of
 * n particles on a grid_dim^2 grid, without a context: every per-step buffer
 * is carved out of one static device array per GPU (BH_NOALLOC_SCRATCH_BYTES,
 * held by each GPU the library runs on until the process ends), so the call
 * makes no hipMalloc / hipFree.  Limits: 1 <= n <= BH_NOALLOC_MAX_N,
 * 1 <= grid_dim <= BH_NOALLOC_MAX_GRID; beyond them BH_EINVAL.  a is placed
 * and sorted into b, then summarize, evaluate (static order) and, if do_tick,
 * tick, all on b -- the same kernels as bh_step on a context configured with
 * bh_set_systems(1, &n, &grid_dim), bh_set_tuning(tuning) and
 * bh_set_recompute(kappa, 0), and the same results bit for bit.
 * tuning: NULL = {p 2, yspan 512, waves_per_cu 32, tail_split 1,
 * tail_items 0}; tail_split must be 1.  kappa: FP64 recompute threshold of
 * the FP32 precisions (<= 0: off).  stats: device int32[16] that receives
 * the step's stats block (layout of bh_read_stats; [0] error bits), or NULL.
 * Runs on the calling thread's current device, asynchronously on stream.
 * Every call on one GPU shares that GPU's array (the last call's contents
 * stay in it): calls that overlap on one GPU race. */
#define BH_NOALLOC_MAX_N 262144
#define BH_NOALLOC_MAX_GRID 256
#define BH_NOALLOC_SCRATCH_BYTES (32 << 20)
int bh_step_noalloc(const bh_particles* a, const bh_particles* b, int64_t n, int32_t grid_dim,
                    const bh_law* law, int32_t precision, double dt, int32_t do_tick,
                    const bh_tuning* tuning, double kappa, int32_t* stats, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* BH_HIP_H */
