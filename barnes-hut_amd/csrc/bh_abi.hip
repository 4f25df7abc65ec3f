// C ABI of libbh_hip.so: context, scratch ownership, phase entry points and
// HIP-graph replay.  The declarations and the reference interface each entry
// point replaces are in include/bh_hip.h.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/bh_hip.h"
#include "bh_common.h"
#include "bh_kernels.h"

const void* bh_force_kernel(int precision, int pidx, int law);

static thread_local std::string g_err;

static int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return fail(BH_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_));           \
    } while (0)

#define LAUNCH_CHECK(what)                                                                     \
    do {                                                                                       \
        hipError_t e_ = hipGetLastError();                                                     \
        if (e_ != hipSuccess)                                                                  \
            return fail(BH_EHIP, std::string("launch ") + what + ": " + hipGetErrorString(e_)); \
    } while (0)

struct GraphKey {
    bh_particles a, b;
    bh_law law;
    int precision;
    double dt;
    int do_tick;
    bh_tuning tune;
    const float* params;
    int32_t* order;
    const double* part;   // slice-plane buffer the graph was captured with
    int cost_log_on;
    double rc_kappa;
    int rc_deferred;
    bool operator==(const GraphKey& o) const { return std::memcmp(this, &o, sizeof(*this)) == 0; }
};

struct bh_ctx_s {
    int device = 0;
    int num_cu = 256;
    int S = 0;
    int64_t N = 0, C = 0, ncols = 0, maxN = 0;
    std::vector<SysDev> sys_h;
    SysDev* sys = nullptr;
    int32_t* sys_of = nullptr;
    uint64_t* bb = nullptr;
    int32_t *key = nullptr, *key_sorted = nullptr, *counts = nullptr, *cursor = nullptr;
    int32_t *offsets = nullptr, *neprefix = nullptr, *ne_list = nullptr;
    int32_t *idx_a = nullptr, *idx_b = nullptr, *perm = nullptr;
    uint64_t *pxkey = nullptr, *pykey = nullptr;   // sort keys of the scattered rows
    int32_t *bsum = nullptr, *bne = nullptr;
    int32_t* stats = nullptr;
    double* com_dense = nullptr;
    Com64* l64 = nullptr;
    Com32* l32 = nullptr;
    Item* items = nullptr;
    int64_t max_items = 0;
    int32_t* col_items = nullptr;
    int32_t* col_off = nullptr;
    long long* cost_log = nullptr;
    int cost_log_on = 0;
    unsigned long long* inter = nullptr;
    int32_t* cost_scratch = nullptr;
    float* feat = nullptr;
    float* pred = nullptr;
    bh_tuning tune = {2, 512, 32, 16, 1 << 20};
    double* part = nullptr;
    size_t part_cap = 0;
    double* energy_buf = nullptr;   // 2 x N + S, allocated by the first energy call
    // FP64 recompute: threshold, the FP32 pass's per-particle scratch, the
    // flagged list and groups, per-block sums of M^2
    double rc_kappa = 1.5;
    int rc_deferred = 0;
    int rc_ready = 0;        // the last force pass prepared a32 / s2n
    double* rc_a32 = nullptr;      // the FP32 pass: x plane, y plane, near-field sums
    int32_t* rc_list = nullptr;    // flagged particles, system s at [p_off_s, p_off_s + cnt_s)
    int32_t* rc_cnt = nullptr;     // per system
    int32_t* rc_gofs = nullptr;    // per system: first group (S + 1)
    double* rc_part = nullptr;     // (group, slice) partial sums
    double* rc_tiles = nullptr;    // per RC_TILE^2 tile: sum of M^2
    double* rc_blk = nullptr;      // per block: sum of M^2
    double* rc_bfar = nullptr;     // per block: its far-block part of the bound
    int64_t nblocks = 0;
    hipStream_t cap_stream = nullptr;   // graph capture only; replays run on the caller's stream
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    GraphKey gkey[2];
    int gnext = 0;
    std::vector<void*> allocs;
};
typedef bh_ctx_s Ctx;

static void free_all(Ctx* c)
{
    for (void* p : c->allocs) (void)hipFree(p);
    c->allocs.clear();
    for (auto& g : c->gexec)
        if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
}

template <typename T>
static int dalloc(Ctx* c, T** p, int64_t count)
{
    size_t bytes = (size_t)std::max<int64_t>(count, 1) * sizeof(T);
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess)
        return fail(BH_EHIP, "hipMalloc(" + std::to_string(bytes) + "): " + hipGetErrorString(e));
    c->allocs.push_back(q);
    *p = (T*)q;
    return BH_OK;
}

static inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }
static int ensure_part(Ctx* c);
static int scratch_address(int dev, unsigned char** out, int* num_cu);
static inline unsigned grid_for(int64_t n, int bs, int64_t cap = 1 << 20)
{
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(cdiv(n, bs), cap));
}

// Workgroups of a persistent force-type kernel: waves_per_cu per CU, but no
// more than the queue is expected to hold -- about one item per 64 P
// particles plus two per grid column, each split into `split` slices -- so a
// small system does not pay for launching thousands of waves that find the
// queue empty.  Any size is correct: every wave takes items until none is left.
static unsigned persistent_grid(const Ctx* c, int split)
{
    const int64_t want = (int64_t)c->num_cu * c->tune.waves_per_cu;
    const int64_t queue = (cdiv(c->N, 64LL * c->tune.p) + 2 * c->ncols) * split;
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(want, queue));
}

// Checks the system sizes and lays the systems out in the global particle,
// cell, column and recompute-block index spaces (host only, nothing changes
// on failure).
static int layout_systems(int32_t nsys, const int64_t* n, const int32_t* grid_dim,
                          std::vector<SysDev>& sh, int64_t& N, int64_t& C, int64_t& cols,
                          int64_t& maxN, int64_t& NB)
{
    sh.assign(nsys, SysDev{});
    N = C = cols = maxN = NB = 0;
    for (int s = 0; s < nsys; ++s) {
        if (n[s] < 1) return fail(BH_EINVAL, "system " + std::to_string(s) + " has n < 1");
        // the FP32 COM list packs column and row in 16 bits each (Com32::kxy)
        if (grid_dim[s] < 1 || grid_dim[s] > 32767)
            return fail(BH_EINVAL, "system " + std::to_string(s) + " grid_dim out of [1, 32767]");
        SysDev& d = sh[s];
        d.p_off = N;
        d.n = n[s];
        d.c_off = C;
        d.col_off = cols;
        d.G = grid_dim[s];
        d.Gb = (int32_t)cdiv(grid_dim[s], RC_BLOCK);
        d.b_off = NB;
        NB += (int64_t)d.Gb * d.Gb;
        N += n[s];
        C += (int64_t)grid_dim[s] * grid_dim[s];
        cols += grid_dim[s];
        maxN = std::max(maxN, n[s]);
        if (N >= (1ll << 31) - 1 || C >= (1ll << 31) - 1)
            return fail(BH_EINVAL, "total particles and cells must each be < 2^31");
    }
    return BH_OK;
}

static void set_sizes(Ctx* c, int32_t nsys, int64_t N, int64_t C, int64_t cols, int64_t maxN,
                      int64_t NB)
{
    c->S = nsys;
    c->N = N;
    c->C = C;
    c->ncols = cols;
    c->maxN = maxN;
    c->nblocks = NB;
    c->max_items = C + cols + 2 * cdiv(N, 64) + 64;
}

// Every device buffer of a configured context with its element count, for the
// sizes set by set_sizes.  bh_set_systems gives each buffer its own
// allocation; the allocation-free step carves them out of one static array.
template <typename F>
static int for_each_buffer(Ctx* c, F&& f)
{
    const int64_t N = c->N, C = c->C, cols = c->ncols, NB = c->nblocks, S = c->S;
    const int64_t nb = std::max(cdiv(C, SCAN_TILE), cdiv(cols, SCAN_TILE)) + 1;
    int r = BH_OK;
#define B(p, cnt) \
    if ((r = f(c->p, (int64_t)(cnt))) != BH_OK) return r
    B(sys, S);
    B(sys_of, N);
    B(bb, 4 * S);
    B(key, N);
    B(key_sorted, N);
    B(counts, C);
    B(cursor, C);
    B(offsets, C + 1);
    B(neprefix, C + 1);
    B(ne_list, C);
    B(idx_a, N);
    B(idx_b, N);
    B(pxkey, N);
    B(pykey, N);
    B(perm, N);
    B(bsum, nb);
    B(bne, nb);
    B(stats, ST_COUNT);
    B(com_dense, 3 * C);
    B(l64, C);
    B(l32, C);
    B(items, c->max_items);
    B(col_items, cols);
    B(col_off, cols + 1);
    B(cost_log, c->max_items);
    B(inter, 2 * S + 2);
    B(cost_scratch, 16 + 3 * 4096 + 1);
    B(feat, BH_MLP_NFEAT * c->max_items);
    B(pred, c->max_items);
    B(rc_a32, 3 * N);
    B(rc_list, N);
    B(rc_cnt, S);
    B(rc_gofs, S + 1);
    B(rc_part, (int64_t)RC_PART_CAP * 2 * RC_GROUP);
    B(rc_tiles, NB * (RC_BLOCK / RC_TILE) * (RC_BLOCK / RC_TILE));
    B(rc_blk, NB);
    B(rc_bfar, NB);
#undef B
    return BH_OK;
}

// Back to the state of a fresh context (S = 0: the phase calls refuse with
// BH_ESTATE).  The buffers must already be released (free_all).
static void unconfigure(Ctx* c)
{
    (void)for_each_buffer(c, [](auto*& p, int64_t) {
        p = nullptr;
        return BH_OK;
    });
    set_sizes(c, 0, 0, 0, 0, 0, 0);
    c->max_items = 0;
    c->sys_h.clear();
    c->energy_buf = nullptr;
    c->rc_ready = 0;
}

// Allocates and initialises the buffers for sizes already validated by
// layout_systems (c->sys_h holds the layout).
static int configure(Ctx* c, int32_t nsys, int64_t N, int64_t C, int64_t cols, int64_t maxN,
                     int64_t NB)
{
    set_sizes(c, nsys, N, C, cols, maxN, NB);
    int r = for_each_buffer(c, [c](auto*& p, int64_t cnt) { return dalloc(c, &p, cnt); });
    if (r != BH_OK) return r;
    HIP_TRY(hipMemcpy(c->sys, c->sys_h.data(), sizeof(SysDev) * nsys, hipMemcpyHostToDevice));
    std::vector<int32_t> so(N);
    for (int s = 0; s < nsys; ++s)
        std::fill(so.begin() + c->sys_h[s].p_off, so.begin() + c->sys_h[s].p_off + c->sys_h[s].n, s);
    HIP_TRY(hipMemcpy(c->sys_of, so.data(), sizeof(int32_t) * N, hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(c->stats, 0, sizeof(int32_t) * ST_COUNT));
    HIP_TRY(hipMemset(c->offsets, 0, sizeof(int32_t) * (C + 1)));
    HIP_TRY(hipMemset(c->neprefix, 0, sizeof(int32_t) * (C + 1)));
    HIP_TRY(hipDeviceSynchronize());
    return BH_OK;
}

extern "C" {

const char* bh_last_error(void) { return g_err.c_str(); }
int bh_version(void) { return 1; }

int bh_create(int device, void** ctx_out)
{
    if (!ctx_out) return fail(BH_EINVAL, "ctx_out is NULL");
    int ndev = 0;
    HIP_TRY(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev)
        return fail(BH_EINVAL, "device " + std::to_string(device) + " out of range");
    HIP_TRY(hipSetDevice(device));
    Ctx* c = new Ctx();
    c->device = device;
    int cu = 0;
    if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess &&
        cu > 0)
        c->num_cu = cu;
    if (hipStreamCreateWithFlags(&c->cap_stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return fail(BH_EHIP, "stream creation failed");
    }
    // the allocation-free step's static scratch: its device address on this GPU
    unsigned char* sp = nullptr;
    int r = scratch_address(device, &sp, nullptr);
    if (r) {
        const std::string msg = g_err;
        bh_destroy(c);
        return fail(r, msg);
    }
    *ctx_out = c;
    return BH_OK;
}

int bh_destroy(void* ctx)
{
    Ctx* c = (Ctx*)ctx;
    if (!c) return BH_OK;
    (void)hipSetDevice(c->device);
    (void)hipDeviceSynchronize();
    free_all(c);
    if (c->part) (void)hipFree(c->part);
    if (c->cap_stream) (void)hipStreamDestroy(c->cap_stream);
    delete c;
    return BH_OK;
}

int bh_set_systems(void* ctx, int32_t nsys, const int64_t* n, const int32_t* grid_dim)
{
    Ctx* c = (Ctx*)ctx;
    if (!c) return fail(BH_EINVAL, "ctx is NULL");
    if (nsys <= 0 || !n || !grid_dim) return fail(BH_EINVAL, "need nsys >= 1 and host arrays");
    // validate and lay out everything before the old state is released, so a
    // refused call leaves the context as it was
    std::vector<SysDev> sh;
    int64_t N = 0, C = 0, cols = 0, maxN = 0, NB = 0;
    int rl = layout_systems(nsys, n, grid_dim, sh, N, C, cols, maxN, NB);
    if (rl != BH_OK) return rl;
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipDeviceSynchronize());
    // from here on a failure leaves the context unconfigured (S = 0: every
    // phase call returns BH_ESTATE), never with pointers into freed memory
    free_all(c);
    unconfigure(c);
    c->sys_h = sh;
    int r = configure(c, nsys, N, C, cols, maxN, NB);
    if (r == BH_OK) r = ensure_part(c);
    if (r != BH_OK) {
        const std::string msg = g_err;
        (void)hipDeviceSynchronize();
        free_all(c);
        unconfigure(c);
        return fail(r, msg);
    }
    return BH_OK;
}

// Scratch of the split force pass: tail_split planes of N x 2 doubles.
static int ensure_part(Ctx* c)
{
    size_t need = (c->tune.tail_split > 1 && c->tune.tail_items > 0 && c->N > 0)
                      ? (size_t)c->N * c->tune.tail_split * 2 * sizeof(double)
                      : 0;
    if (need > c->part_cap) {
        HIP_TRY(hipDeviceSynchronize());
        // captured steps hold the old buffer's address: drop them
        for (auto& g : c->gexec)
            if (g) {
                (void)hipGraphExecDestroy(g);
                g = nullptr;
            }
        if (c->part) (void)hipFree(c->part);
        c->part = nullptr;
        c->part_cap = 0;
        HIP_TRY(hipMalloc(&c->part, need));
        c->part_cap = need;
    }
    return BH_OK;
}

int bh_set_tuning(void* ctx, const bh_tuning* t)
{
    Ctx* c = (Ctx*)ctx;
    if (!c || !t) return fail(BH_EINVAL, "NULL argument");
    if (!(t->p == 1 || t->p == 2 || t->p == 4)) return fail(BH_EINVAL, "tuning.p must be 1, 2 or 4");
    if (t->yspan < 1 || t->yspan > 1 << 20) return fail(BH_EINVAL, "tuning.yspan must be >= 1");
    if (t->waves_per_cu < 1 || t->waves_per_cu > 64)
        return fail(BH_EINVAL, "tuning.waves_per_cu must be in [1, 64]");
    if (t->tail_split < 1 || t->tail_split > 16)
        return fail(BH_EINVAL, "tuning.tail_split must be in [1, 16]");
    if (t->tail_items < 0 || t->tail_items > (1 << 20))
        return fail(BH_EINVAL, "tuning.tail_items must be in [0, 2^20]");
    const bh_tuning old = c->tune;
    c->tune = *t;
    int r = ensure_part(c);
    if (r != BH_OK) {   // keep a consistent state: the previous shape and its buffer
        std::string msg = g_err;
        c->tune = old;
        (void)ensure_part(c);
        return fail(r, msg);
    }
    return BH_OK;
}

int bh_get_tuning(void* ctx, bh_tuning* t)
{
    Ctx* c = (Ctx*)ctx;
    if (!c || !t) return fail(BH_EINVAL, "NULL argument");
    *t = c->tune;
    return BH_OK;
}

int bh_set_cost_log(void* ctx, int32_t enable)
{
    Ctx* c = (Ctx*)ctx;
    if (!c) return fail(BH_EINVAL, "ctx is NULL");
    c->cost_log_on = enable ? 1 : 0;
    return BH_OK;
}

static int check_parts(const bh_particles* p, const char* name)
{
    if (!p) return fail(BH_EINVAL, std::string(name) + " is NULL");
    if (!p->px || !p->py || !p->mass || !p->vx || !p->vy || !p->ax || !p->ay || !p->pid)
        return fail(BH_EINVAL, std::string(name) + " has a NULL field");
    return BH_OK;
}

#define CTX_READY(c)                                                        \
    do {                                                                    \
        if (!(c)) return fail(BH_EINVAL, "ctx is NULL");                    \
        if ((c)->S == 0) return fail(BH_ESTATE, "bh_set_systems not called"); \
    } while (0)

int bh_place_particles(void* ctx, const bh_particles* p, void* stream)
{
    Ctx* c = (Ctx*)ctx;
    CTX_READY(c);
    int r = check_parts(p, "particles");
    if (r) return r;
    hipStream_t st = (hipStream_t)stream;
    const int64_t N = c->N;
    unsigned nchunk = (unsigned)cdiv(N, PLACE_CHUNK);
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, st, c->stats, (int64_t)ST_COUNT, 0);
    hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(c->C, 256, 4096)), dim3(256), 0, st, c->counts,
                       c->C, 0);
    hipLaunchKernelGGL(k_bbox_init, dim3(cdiv(c->S, 256)), dim3(256), 0, st, c->bb, c->S);
    hipLaunchKernelGGL(k_bbox, dim3(nchunk), dim3(256), 0, st, (const double*)p->px,
                       (const double*)p->py, (const int32_t*)c->sys_of, N, c->bb, c->stats);
    hipLaunchKernelGGL(k_bbox_finalize, dim3(cdiv(c->S, 256)), dim3(256), 0, st, c->sys,
                       (const uint64_t*)c->bb, c->S, c->stats);
    hipLaunchKernelGGL(k_keys_hist, dim3(nchunk), dim3(256), 0, st, (const double*)p->px,
                       (const double*)p->py, (const int32_t*)c->sys_of, (const SysDev*)c->sys, N,
                       c->key, c->counts);
    LAUNCH_CHECK("place");
    return BH_OK;
}

static void run_scan(Ctx* c, bool cnt, const int32_t* in, int64_t n, int32_t* out_v,
                     int32_t* out_f, int total_slot, hipStream_t st)
{
    unsigned nb = (unsigned)cdiv(n, SCAN_TILE);
    if (cnt) {
        hipLaunchKernelGGL(k_scan_reduce<true>, dim3(nb), dim3(256), 0, st, in, n, c->bsum, c->bne,
                           c->stats);
        hipLaunchKernelGGL(k_scan_blocks<true>, dim3(1), dim3(1024), 0, st, c->bsum, c->bne,
                           (int64_t)nb, c->stats, total_slot);
        hipLaunchKernelGGL(k_scan_final<true>, dim3(nb), dim3(256), 0, st, in, n,
                           (const int32_t*)c->bsum, (const int32_t*)c->bne, out_v, out_f);
    } else {
        hipLaunchKernelGGL(k_scan_reduce<false>, dim3(nb), dim3(256), 0, st, in, n, c->bsum,
                           c->bne, c->stats);
        hipLaunchKernelGGL(k_scan_blocks<false>, dim3(1), dim3(1024), 0, st, c->bsum, c->bne,
                           (int64_t)nb, c->stats, total_slot);
        hipLaunchKernelGGL(k_scan_final<false>, dim3(nb), dim3(256), 0, st, in, n,
                           (const int32_t*)c->bsum, (const int32_t*)c->bne, out_v, out_f);
    }
}

int bh_calculate_box_cumm(void* ctx, void* stream)
{
    Ctx* c = (Ctx*)ctx;
    CTX_READY(c);
    hipStream_t st = (hipStream_t)stream;
    run_scan(c, true, c->counts, c->C, c->offsets, c->neprefix, -1, st);
    LAUNCH_CHECK("scan");
    return BH_OK;
}

static SortFields fields_of(const bh_particles* p)
{
    SortFields f;
    f.px = p->px;
    f.py = p->py;
    f.m = p->mass;
    f.vx = p->vx;
    f.vy = p->vy;
    f.ax = p->ax;
    f.ay = p->ay;
    f.pid = p->pid;
    return f;
}

static BhParticles dev_of(const bh_particles* p)
{
    BhParticles d;
    d.px = p->px;
    d.py = p->py;
    d.m = p->mass;
    d.vx = p->vx;
    d.vy = p->vy;
    d.ax = p->ax;
    d.ay = p->ay;
    d.pid = p->pid;
    return d;
}

int bh_sort_particles(void* ctx, const bh_particles* src, const bh_particles* dst, void* stream)
{
    Ctx* c = (Ctx*)ctx;
    CTX_READY(c);
    int r = check_parts(src, "src");
    if (r) return r;
    if ((r = check_parts(dst, "dst"))) return r;
    if (src->px == dst->px) return fail(BH_EINVAL, "src and dst must be different buffers");
    hipStream_t st = (hipStream_t)stream;
    const int64_t N = c->N;
    SortFields F = fields_of(src);
    hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(c->C, 256, 4096)), dim3(256), 0, st, c->cursor,
                       c->C, 0);
    hipLaunchKernelGGL(k_scatter, dim3(cdiv(N, PLACE_CHUNK)), dim3(256), 0, st,
                       (const int32_t*)c->key, N, (const int32_t*)c->offsets, c->cursor, c->idx_a,
                       c->key_sorted, (const double*)src->px, (const double*)src->py, c->pxkey,
                       c->pykey);
    hipLaunchKernelGGL(k_sort_units, dim3(cdiv(N, SORT_W)), dim3(SORT_THREADS), 0, st, F, N,
                       (const int32_t*)c->idx_a, (const int32_t*)c->key_sorted,
                       (const uint64_t*)c->pxkey, (const uint64_t*)c->pykey,
                       (const int32_t*)c->offsets, c->idx_b);
    // merge passes for cells longer than SORT_W: b -> a -> b -> ...
    int maxpass = 0;
    while (((int64_t)SORT_W << maxpass) < c->maxN) ++maxpass;
    unsigned mgrid = grid_for(N, 256, (int64_t)c->num_cu * 8);
    for (int L = 0; L < maxpass; ++L) {
        const int32_t* in = (L & 1) ? c->idx_a : c->idx_b;
        int32_t* out = (L & 1) ? c->idx_b : c->idx_a;
        hipLaunchKernelGGL(k_sort_merge, dim3(mgrid), dim3(256), 0, st, F, N, in,
                           (const int32_t*)c->key_sorted, (const int32_t*)c->offsets, out, L,
                           (const int32_t*)c->stats);
    }
    hipLaunchKernelGGL(k_gather, dim3(grid_for(N, 256, (int64_t)c->num_cu * 16)), dim3(256), 0, st,
                       F, dev_of(dst), N, (const int32_t*)c->idx_a, (const int32_t*)c->idx_b,
                       (const int32_t*)c->key_sorted, (const int32_t*)c->offsets,
                       (const int32_t*)c->stats, c->perm);
    LAUNCH_CHECK("sort");
    return BH_OK;
}

int bh_summarize(void* ctx, const bh_particles* p, void* stream)
{
    Ctx* c = (Ctx*)ctx;
    CTX_READY(c);
    int r = check_parts(p, "particles");
    if (r) return r;
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_ne_list, dim3(grid_for(c->C, 256, 8192)), dim3(256), 0, st,
                       (const int32_t*)c->counts, (const int32_t*)c->neprefix, c->C, c->ne_list);
    hipLaunchKernelGGL(k_fill_f64, dim3(grid_for(3 * c->C, 256, 8192)), dim3(256), 0, st,
                       c->com_dense, 3 * c->C, 0.0);
    // one wave per non-empty cell; the grid is sized for the worst case
    int64_t waves = std::min<int64_t>(c->C, (int64_t)c->num_cu * 64);
    hipLaunchKernelGGL(k_com, dim3(grid_for(waves * 64, 256)), dim3(256), 0, st,
                       (const SysDev*)c->sys, c->S, (const int32_t*)c->ne_list,
                       (const int32_t*)c->stats, (const int32_t*)c->offsets,
                       (const double*)p->px, (const double*)p->py, (const double*)p->mass,
                       c->com_dense, c->l64, c->l32);
    LAUNCH_CHECK("summarize");
    return BH_OK;
}

int bh_build_items(void* ctx, void* stream)
{
    Ctx* c = (Ctx*)ctx;
    CTX_READY(c);
    hipStream_t st = (hipStream_t)stream;
    const int pitem = 64 * c->tune.p;
    unsigned g = (unsigned)c->ncols;   // one wave per grid column
    hipLaunchKernelGGL(k_items<false>, dim3(g), dim3(64), 0, st, (const SysDev*)c->sys, c->S,
                       c->ncols, (const int32_t*)c->counts, (const int32_t*)c->offsets, pitem,
                       c->tune.yspan, c->col_items, c->items, c->max_items, c->stats);
    run_scan(c, false, c->col_items, c->ncols, c->col_off, nullptr, ST_NITEMS, st);
    hipLaunchKernelGGL(k_clamp_items, dim3(1), dim3(64), 0, st, c->stats, c->max_items);
    hipLaunchKernelGGL(k_items<true>, dim3(g), dim3(64), 0, st, (const SysDev*)c->sys, c->S,
                       c->ncols, (const int32_t*)c->counts, (const int32_t*)c->offsets, pitem,
                       c->tune.yspan, c->col_off, c->items, c->max_items, c->stats);
    LAUNCH_CHECK("items");
    return BH_OK;
}

int bh_evaluate_boxes(void* ctx, const bh_particles* p, const bh_law* law, int32_t precision,
                      const int32_t* order, void* stream)
{
    Ctx* c = (Ctx*)ctx;
    CTX_READY(c);
    int r = check_parts(p, "particles");
    if (r) return r;
    if (!law) return fail(BH_EINVAL, "law is NULL");
    if (precision < 0 || precision > 5) return fail(BH_EINVAL, "precision must be in [0, 5]");
    if (!(law->eps2 >= 0.0) || !(law->r2_cut >= 0.0) || !std::isfinite(law->G))
        return fail(BH_EINVAL, "law needs finite G, eps2 >= 0, r2_cut >= 0");
    if (!order) {
        if ((r = bh_build_items(ctx, stream))) return r;
    }
    return bh_force_pass(ctx, p, law, precision, order, stream);
}

int bh_force_pass(void* ctx, const bh_particles* p, const bh_law* law, int32_t precision,
                  const int32_t* order, void* stream)
{
    return bh_force_pass_part(ctx, p, law, precision, order, 0, 1, stream);
}

int bh_force_pass_part(void* ctx, const bh_particles* p, const bh_law* law, int32_t precision,
                       const int32_t* order, int32_t part, int32_t nparts, void* stream)
{
    Ctx* c = (Ctx*)ctx;
    CTX_READY(c);
    if (nparts < 1 || nparts > 4096 || part < 0 || part >= nparts)
        return fail(BH_EINVAL, "need 0 <= part < nparts <= 4096");
    int r = check_parts(p, "particles");
    if (r) return r;
    if (!law) return fail(BH_EINVAL, "law is NULL");
    if (precision < 0 || precision > 5) return fail(BH_EINVAL, "precision must be in [0, 5]");
    if (!(law->eps2 >= 0.0) || !(law->r2_cut >= 0.0) || !std::isfinite(law->G))
        return fail(BH_EINVAL, "law needs finite G, eps2 >= 0, r2_cut >= 0");
    hipStream_t st = (hipStream_t)stream;
    const int pidx = c->tune.p == 1 ? 0 : (c->tune.p == 2 ? 1 : 2);
    const int lawi = law->r2_cut > 0.0 ? 1 : 0;
    const void* fn = bh_force_kernel(precision, pidx, lawi);
    if (!fn) return fail(BH_EINVAL, "no force kernel for this configuration");
    ForceArgs fa;
    fa.items = c->items;
    fa.order = order;
    fa.stats = c->stats;
    fa.sys = c->sys;
    fa.offsets = c->offsets;
    fa.neprefix = c->neprefix;
    fa.key_sorted = c->key_sorted;
    fa.x = p->px;
    fa.y = p->py;
    fa.m = p->mass;
    fa.l64 = c->l64;
    fa.l32 = c->l32;
    fa.ax = p->ax;
    fa.ay = p->ay;
    fa.Gc = law->G;
    fa.eps2 = law->eps2;
    fa.r2cut = law->r2_cut;
    fa.cost_log = c->cost_log_on ? c->cost_log : nullptr;
    const bool fp32v = precision >= 1 && precision <= 3;
    fa.split = (fp32v && c->tune.tail_split > 1 && c->tune.tail_items > 0) ? c->tune.tail_split : 1;
    fa.n_tail = fa.split > 1 ? c->tune.tail_items : 0;
    if (fa.split > 1 && c->part_cap < (size_t)c->N * fa.split * 2 * sizeof(double))
        return fail(BH_EINVAL, "slice buffer missing for tail_split (bh_set_tuning failed?)");
    fa.part = c->part;
    fa.part_n = c->N;
    const bool rc = fp32v && c->rc_kappa > 0.0;
    fa.rc = rc ? 1 : 0;
    fa.qpart = part;
    fa.nqparts = nparts;
    if (rc) {   // the FP32 result goes to scratch; the recompute adds it (or the FP64 one)
        if (part == 0)
            hipLaunchKernelGGL(k_fill_f64, dim3(grid_for(2 * c->N, 256, 8192)), dim3(256), 0, st,
                               c->rc_a32, 2 * c->N, 0.0);
        fa.ax = c->rc_a32;
        fa.ay = c->rc_a32 + c->N;
    }
    if (part == 0) c->rc_ready = 0;   // until the last part: no recompute of a partial pass
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, st, c->stats + ST_QUEUE, (int64_t)1, 0);
    void* args[] = {&fa};
    HIP_TRY(hipLaunchKernel(fn, dim3(persistent_grid(c, fa.split)), dim3(64), args, 0, st));
    if (part < nparts - 1) {   // the slice combine and the recompute follow the last part
        LAUNCH_CHECK("evaluate");
        return BH_OK;
    }
    if (fa.split > 1) {
        const int pslots = 64 * c->tune.p;
        hipLaunchKernelGGL(k_force_combine, dim3(grid_for((int64_t)fa.n_tail * pslots, 256,
                                                          (int64_t)c->num_cu * 32)), dim3(256),
                           0, st, (const Item*)c->items, order, (const int32_t*)c->stats, fa.split,
                           fa.n_tail, pslots, (const double*)c->part, fa.part_n, fa.ax, fa.ay);
    }
    if (!rc)   // the recomputed count describes the last pass
        hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, st, c->stats + ST_NRECOMP,
                           (int64_t)2, 0);
    LAUNCH_CHECK("evaluate");
    c->rc_ready = rc ? 1 : 0;
    if (rc && !c->rc_deferred) return bh_recompute(ctx, p, law, stream);
    return BH_OK;
}

int bh_set_recompute(void* ctx, double kappa, int32_t deferred)
{
    Ctx* c = (Ctx*)ctx;
    if (!c) return fail(BH_EINVAL, "ctx is NULL");
    if (std::isnan(kappa)) return fail(BH_EINVAL, "kappa is NaN");
    c->rc_kappa = kappa > 0.0 ? kappa : 0.0;
    c->rc_deferred = deferred ? 1 : 0;
    c->rc_ready = 0;
    return BH_OK;
}

int bh_recompute(void* ctx, const bh_particles* p, const bh_law* law, void* stream)
{
    Ctx* c = (Ctx*)ctx;
    CTX_READY(c);
    int r = check_parts(p, "particles");
    if (r) return r;
    if (!law) return fail(BH_EINVAL, "law is NULL");
    if (!(law->eps2 >= 0.0) || !(law->r2_cut >= 0.0) || !std::isfinite(law->G))
        return fail(BH_EINVAL, "law needs finite G, eps2 >= 0, r2_cut >= 0");
    if (!c->rc_ready) return fail(BH_ESTATE, "no FP32 force pass prepared a recompute");
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, st, c->stats + ST_NRECOMP, (int64_t)2, 0);
    if (law->G == 0.0) return BH_OK;   // every term, and so the FP32 result, is zero
    hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(c->S, 64, 1024)), dim3(64), 0, st, c->rc_cnt,
                       (int64_t)c->S, 0);
    hipLaunchKernelGGL(k_rc_blocks, dim3(grid_for(c->nblocks, 1, (int64_t)c->num_cu * 16)),
                       dim3(64), 0, st, (const SysDev*)c->sys, c->S, c->nblocks,
                       (const double*)c->com_dense, c->rc_blk);
    const int64_t ntiles = c->nblocks * (RC_BLOCK / RC_TILE) * (RC_BLOCK / RC_TILE);
    hipLaunchKernelGGL(k_rc_tiles, dim3(grid_for(ntiles, 256, (int64_t)c->num_cu * 8)), dim3(256),
                       0, st, (const SysDev*)c->sys, c->S, ntiles, (const double*)c->com_dense,
                       c->rc_tiles);
    hipLaunchKernelGGL(k_rc_blockfar, dim3(grid_for(c->nblocks, 1, (int64_t)c->num_cu * 16)),
                       dim3(64), 0, st, (const SysDev*)c->sys, c->S, c->nblocks,
                       (const double*)c->rc_blk, c->rc_bfar);
    hipLaunchKernelGGL(k_rc_flag, dim3(grid_for(c->max_items, 1, (int64_t)c->num_cu * 32)),
                       dim3(64), 0, st, (const Item*)c->items, (const SysDev*)c->sys,
                       (const int32_t*)c->key_sorted, (const int32_t*)c->neprefix,
                       (const Com64*)c->l64, (const double*)p->px, (const double*)p->py,
                       (const double*)c->com_dense, (const double*)c->rc_tiles,
                       (const double*)c->rc_blk, (const double*)c->rc_bfar,
                       (const double*)c->rc_a32, (const double*)(c->rc_a32 + c->N),
                       (const double*)(c->rc_a32 + 2 * c->N), law->G, c->rc_kappa * c->rc_kappa,
                       (const int32_t*)c->stats, c->rc_cnt, c->rc_list, p->ax, p->ay);
    hipLaunchKernelGGL(k_rc_groups, dim3(1), dim3(256), 0, st, (const int32_t*)c->rc_cnt, c->S,
                       c->rc_gofs, c->stats);
    const unsigned g = (unsigned)((int64_t)c->num_cu * 4);
    if (law->r2_cut > 0.0)
        hipLaunchKernelGGL(k_rc_fix64<1>, dim3(g), dim3(RC_THREADS), 0, st,
                           (const int32_t*)c->stats, (const int32_t*)c->rc_cnt,
                           (const int32_t*)c->rc_gofs, c->S, (const int32_t*)c->rc_list,
                           (const SysDev*)c->sys, (const int32_t*)c->key_sorted,
                           (const int32_t*)c->offsets, (const int32_t*)c->neprefix,
                           (const Com64*)c->l64, (const double*)p->px, (const double*)p->py,
                           (const double*)p->mass, law->G, law->eps2, law->r2_cut, c->rc_part,
                           p->ax, p->ay);
    else
        hipLaunchKernelGGL(k_rc_fix64<0>, dim3(g), dim3(RC_THREADS), 0, st,
                           (const int32_t*)c->stats, (const int32_t*)c->rc_cnt,
                           (const int32_t*)c->rc_gofs, c->S, (const int32_t*)c->rc_list,
                           (const SysDev*)c->sys, (const int32_t*)c->key_sorted,
                           (const int32_t*)c->offsets, (const int32_t*)c->neprefix,
                           (const Com64*)c->l64, (const double*)p->px, (const double*)p->py,
                           (const double*)p->mass, law->G, law->eps2, law->r2_cut, c->rc_part,
                           p->ax, p->ay);
    hipLaunchKernelGGL(k_rc_combine, dim3(grid_for((int64_t)RC_PART_CAP * RC_GROUP, 256, 512)),
                       dim3(256), 0, st, (const int32_t*)c->stats, (const int32_t*)c->rc_cnt,
                       (const int32_t*)c->rc_gofs, c->S, (const int32_t*)c->rc_list,
                       (const SysDev*)c->sys, (const double*)c->rc_part, law->G, p->ax, p->ay);
    LAUNCH_CHECK("recompute");
    return BH_OK;
}

int bh_tick_particles(void* ctx, const bh_particles* p, double dt, void* stream)
{
    Ctx* c = (Ctx*)ctx;
    CTX_READY(c);
    int r = check_parts(p, "particles");
    if (r) return r;
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_tick, dim3(grid_for(c->N, 256, (int64_t)c->num_cu * 16)), dim3(256), 0, st,
                       p->px, p->py, p->vx, p->vy, p->ax, p->ay, c->N, dt);
    LAUNCH_CHECK("tick");
    return BH_OK;
}

extern "C" int bh_learned_order(void* ctx, const float* params, int32_t* order, void* stream);

static int step_impl(void* ctx, const bh_particles* a, const bh_particles* b, const bh_law* law,
                     int32_t precision, double dt, int32_t do_tick, const float* params,
                     int32_t* order, void* stream)
{
    int r;
    if ((r = bh_place_particles(ctx, a, stream))) return r;
    if ((r = bh_calculate_box_cumm(ctx, stream))) return r;
    if ((r = bh_sort_particles(ctx, a, b, stream))) return r;
    if ((r = bh_summarize(ctx, b, stream))) return r;
    if ((r = bh_build_items(ctx, stream))) return r;
    const int32_t* ord = nullptr;
    if (params) {
        if ((r = bh_learned_order(ctx, params, order, stream))) return r;
        ord = order;
    }
    if ((r = bh_force_pass(ctx, b, law, precision, ord, stream))) return r;
    Ctx* c = (Ctx*)ctx;
    if (c->rc_ready && c->rc_deferred && (r = bh_recompute(ctx, b, law, stream))) return r;
    if (do_tick && (r = bh_tick_particles(ctx, b, dt, stream))) return r;
    return BH_OK;
}

int bh_step(void* ctx, const bh_particles* a, const bh_particles* b, const bh_law* law,
            int32_t precision, double dt, int32_t do_tick, void* stream)
{
    return step_impl(ctx, a, b, law, precision, dt, do_tick, nullptr, nullptr, stream);
}

static int step_graph_impl(void* ctx, const bh_particles* a, const bh_particles* b,
                           const bh_law* law, int32_t precision, double dt, int32_t do_tick,
                           const float* params, int32_t* order, void* stream)
{
    Ctx* c = (Ctx*)ctx;
    CTX_READY(c);
    int r = check_parts(a, "a");
    if (r) return r;
    if ((r = check_parts(b, "b"))) return r;
    if (!law) return fail(BH_EINVAL, "law is NULL");
    GraphKey k;
    std::memset(&k, 0, sizeof(k));
    k.a = *a;
    k.b = *b;
    k.law = *law;
    k.precision = precision;
    k.dt = dt;
    k.do_tick = do_tick;
    k.tune = c->tune;
    k.params = params;
    k.order = order;
    k.part = c->part;
    k.cost_log_on = c->cost_log_on;
    k.rc_kappa = c->rc_kappa;
    k.rc_deferred = c->rc_deferred;
    int slot = -1;
    for (int i = 0; i < 2; ++i)
        if (c->gexec[i] && c->gkey[i] == k) slot = i;
    if (slot < 0) {
        slot = c->gnext;
        c->gnext ^= 1;
        if (c->gexec[slot]) {
            HIP_TRY(hipGraphExecDestroy(c->gexec[slot]));
            c->gexec[slot] = nullptr;
        }
        HIP_TRY(hipStreamBeginCapture(c->cap_stream, hipStreamCaptureModeThreadLocal));
        int rr = step_impl(ctx, a, b, law, precision, dt, do_tick, params, order, c->cap_stream);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(c->cap_stream, &g);
        if (rr) {
            if (g) (void)hipGraphDestroy(g);
            return rr;
        }
        if (e != hipSuccess) return fail(BH_EHIP, std::string("capture: ") + hipGetErrorString(e));
        e = hipGraphInstantiate(&c->gexec[slot], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return fail(BH_EHIP, std::string("instantiate: ") + hipGetErrorString(e));
        c->gkey[slot] = k;
    }
    hipStream_t st = (hipStream_t)stream;
    // the graph was captured on the context's own stream and replays on the
    // caller's: ordered after the caller's earlier work, before its later work
    HIP_TRY(hipGraphLaunch(c->gexec[slot], st));
    return BH_OK;
}

int bh_step_graph(void* ctx, const bh_particles* a, const bh_particles* b, const bh_law* law,
                  int32_t precision, double dt, int32_t do_tick, void* stream)
{
    return step_graph_impl(ctx, a, b, law, precision, dt, do_tick, nullptr, nullptr, stream);
}

int bh_step_learned(void* ctx, const bh_particles* a, const bh_particles* b, const bh_law* law,
                    int32_t precision, double dt, int32_t do_tick, const float* params,
                    int32_t* order, int32_t graph, void* stream)
{
    if (!params || !order) return fail(BH_EINVAL, "params and order are required");
    if (graph) return step_graph_impl(ctx, a, b, law, precision, dt, do_tick, params, order, stream);
    return step_impl(ctx, a, b, law, precision, dt, do_tick, params, order, stream);
}

int bh_read_stats(void* ctx, int32_t* out16, void* stream)
{
    Ctx* c = (Ctx*)ctx;
    CTX_READY(c);
    if (!out16) return fail(BH_EINVAL, "out16 is NULL");
    HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    HIP_TRY(hipMemcpy(out16, c->stats, sizeof(int32_t) * ST_COUNT, hipMemcpyDeviceToHost));
    if (out16[ST_ERR]) {
        std::string m = "device error bits " + std::to_string(out16[ST_ERR]) + ":";
        if (out16[ST_ERR] & BH_ERR_NONFINITE) m += " non-finite particle coordinate;";
        if (out16[ST_ERR] & BH_ERR_DEGENERATE) m += " zero-extent bounding box;";
        if (out16[ST_ERR] & BH_ERR_ITEMS) m += " work-item capacity exceeded;";
        return fail(BH_EDEVICE, m);
    }
    return BH_OK;
}

int bh_export(void* ctx, int32_t what, void* dst, int64_t count, void* stream)
{
    Ctx* c = (Ctx*)ctx;
    CTX_READY(c);
    if (!dst || count < 0) return fail(BH_EINVAL, "bad dst/count");
    const void* src = nullptr;
    int64_t cap = 0;
    size_t esz = 4;
    std::vector<double> bbox;
    switch (what) {
    case BH_EXPORT_COUNTS: src = c->counts; cap = c->C; break;
    case BH_EXPORT_OFFSETS: src = c->offsets; cap = c->C + 1; break;
    case BH_EXPORT_NEPREFIX: src = c->neprefix; cap = c->C + 1; break;
    case BH_EXPORT_PERM: src = c->perm; cap = c->N; break;
    case BH_EXPORT_KEYS: src = c->key_sorted; cap = c->N; break;
    case BH_EXPORT_KEYS_IN: src = c->key; cap = c->N; break;
    case BH_EXPORT_COM: src = c->com_dense; cap = 3 * c->C; esz = 8; break;
    case BH_EXPORT_ITEMS: src = c->items; cap = 8 * c->max_items; break;
    case BH_EXPORT_COST: src = c->cost_log; cap = c->max_items; esz = 8; break;
    case BH_EXPORT_BBOX: {
        if (count > 5 * c->S) return fail(BH_EINVAL, "count exceeds 5*S");
        HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
        std::vector<SysDev> h(c->S);
        HIP_TRY(hipMemcpy(h.data(), c->sys, sizeof(SysDev) * c->S, hipMemcpyDeviceToHost));
        bbox.resize(5 * c->S);
        for (int s = 0; s < c->S; ++s) {
            bbox[5 * s + 0] = h[s].min_x;
            bbox[5 * s + 1] = h[s].min_y;
            bbox[5 * s + 2] = h[s].max_x;
            bbox[5 * s + 3] = h[s].max_y;
            bbox[5 * s + 4] = h[s].step;
        }
        HIP_TRY(hipMemcpy(dst, bbox.data(), sizeof(double) * count, hipMemcpyDefault));
        return BH_OK;
    }
    default: return fail(BH_EINVAL, "unknown export " + std::to_string(what));
    }
    if (count > cap) return fail(BH_EINVAL, "count " + std::to_string(count) + " exceeds " +
                                                std::to_string(cap));
    HIP_TRY(hipMemcpyAsync(dst, src, esz * (size_t)count, hipMemcpyDefault, (hipStream_t)stream));
    HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    return BH_OK;
}

int bh_query(void* ctx, int32_t what, int64_t* out)
{
    Ctx* c = (Ctx*)ctx;
    if (!c || !out) return fail(BH_EINVAL, "NULL argument");
    switch (what) {
    case 0: *out = c->N; break;
    case 1: *out = c->C; break;
    case 2: *out = c->S; break;
    case 3: *out = c->ncols; break;
    case 4: *out = c->max_items; break;
    case 5: *out = c->num_cu; break;
    default: return fail(BH_EINVAL, "unknown query " + std::to_string(what));
    }
    return BH_OK;
}

int bh_interactions_async(void* ctx, uint64_t* dev_out, void* stream)
{
    Ctx* c = (Ctx*)ctx;
    CTX_READY(c);
    if (!dev_out) return fail(BH_EINVAL, "dev_out is NULL");
    hipStream_t st = (hipStream_t)stream;
    const int64_t n2 = 2 * (int64_t)c->S + 2;
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(256), 0, st, (int32_t*)dev_out, 2 * n2, 0);
    hipLaunchKernelGGL(k_interactions, dim3(grid_for(c->C, 256, 1024)), dim3(256), 0, st,
                       (const SysDev*)c->sys, c->S, (const int32_t*)c->counts,
                       (const int32_t*)c->neprefix, c->C, (unsigned long long*)dev_out);
    LAUNCH_CHECK("interactions");
    return BH_OK;
}

}  // extern "C"

// Per system: out[2s] = G / 4 * sum u_i and out[2s+1] = sum k_i / 2 over the
// system's rows of energy_buf (u: N values, then k: N values), Neumaier sums in
// row order on the host.  Synchronises.
static int energy_sums(Ctx* c, double G, double* out, hipStream_t st)
{
    HIP_TRY(hipStreamSynchronize(st));
    std::vector<double> h(2 * c->N);
    HIP_TRY(hipMemcpy(h.data(), c->energy_buf, sizeof(double) * 2 * c->N, hipMemcpyDeviceToHost));
    for (int s = 0; s < c->S; ++s) {
        const SysDev& d = c->sys_h[s];
        double su = 0.0, cu = 0.0, sk = 0.0, ck = 0.0;
        for (int64_t i = d.p_off; i < d.p_off + d.n; ++i) {
            double v = h[i], t = su + v;
            cu += (std::fabs(su) >= std::fabs(v)) ? (su - t) + v : (v - t) + su;
            su = t;
            v = h[c->N + i];
            t = sk + v;
            ck += (std::fabs(sk) >= std::fabs(v)) ? (sk - t) + v : (v - t) + sk;
            sk = t;
        }
        out[2 * s] = 0.25 * G * (su + cu);
        out[2 * s + 1] = 0.5 * (sk + ck);
    }
    return BH_OK;
}

extern "C" {

int bh_energy(void* ctx, const bh_particles* p, double G, double eps2, double* out, void* stream)
{
    Ctx* c = (Ctx*)ctx;
    CTX_READY(c);
    if (!out) return fail(BH_EINVAL, "out is NULL");
    if (int e = check_parts(p, "energy")) return e;
    if (!c->energy_buf)
        if (int e = dalloc(c, &c->energy_buf, 2 * c->N + c->S)) return e;
    hipStream_t st = (hipStream_t)stream;
    double* u = c->energy_buf;
    double* k = c->energy_buf + c->N;
    for (int s = 0; s < c->S; ++s) {
        const SysDev& d = c->sys_h[s];
        if (d.n == 0) continue;
        hipLaunchKernelGGL(k_energy, dim3((unsigned)cdiv(d.n, 256)), dim3(256), 0, st, p->px, p->py,
                           p->mass, p->vx, p->vy, d.p_off, d.n, eps2, u, k);
    }
    LAUNCH_CHECK("energy");
    return energy_sums(c, G, out, st);
}

int bh_energy_grid(void* ctx, const bh_particles* p, double G, double eps2, double* out,
                   void* stream)
{
    Ctx* c = (Ctx*)ctx;
    CTX_READY(c);
    if (!out) return fail(BH_EINVAL, "out is NULL");
    if (int e = check_parts(p, "energy")) return e;
    if (!(eps2 >= 0.0) || !std::isfinite(G)) return fail(BH_EINVAL, "need finite G and eps2 >= 0");
    if (!c->energy_buf)
        if (int e = dalloc(c, &c->energy_buf, 2 * c->N + c->S)) return e;
    hipStream_t st = (hipStream_t)stream;
    if (int e = bh_build_items(ctx, stream)) return e;
    double* u = c->energy_buf;
    double* msys = c->energy_buf + 2 * c->N;
    hipLaunchKernelGGL(k_sys_mass, dim3((unsigned)c->S), dim3(256), 0, st, (const SysDev*)c->sys,
                       (const int32_t*)c->neprefix, (const Com64*)c->l64, msys);
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, st, c->stats + ST_QUEUE, (int64_t)1, 0);
    PotArgs pa;
    pa.items = c->items;
    pa.stats = c->stats;
    pa.sys = c->sys;
    pa.offsets = c->offsets;
    pa.neprefix = c->neprefix;
    pa.key_sorted = c->key_sorted;
    pa.x = p->px;
    pa.y = p->py;
    pa.m = p->mass;
    pa.vx = p->vx;
    pa.vy = p->vy;
    pa.l64 = c->l64;
    pa.l32 = c->l32;
    pa.msys = msys;
    pa.eps2 = eps2;
    pa.u = u;
    pa.k = u + c->N;
    const dim3 grid(persistent_grid(c, 1));
    if (c->tune.p == 1) hipLaunchKernelGGL(k_potential<1>, grid, dim3(64), 0, st, pa);
    else if (c->tune.p == 2) hipLaunchKernelGGL(k_potential<2>, grid, dim3(64), 0, st, pa);
    else hipLaunchKernelGGL(k_potential<4>, grid, dim3(64), 0, st, pa);
    LAUNCH_CHECK("energy_grid");
    return energy_sums(c, G, out, st);
}

int bh_interactions(void* ctx, int64_t* out, void* stream)
{
    Ctx* c = (Ctx*)ctx;
    CTX_READY(c);
    if (!out) return fail(BH_EINVAL, "out is NULL");
    hipStream_t st = (hipStream_t)stream;
    const int64_t n2 = 2 * (int64_t)c->S + 2;
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(256), 0, st, (int32_t*)c->inter, 2 * n2, 0);
    hipLaunchKernelGGL(k_interactions, dim3(grid_for(c->C, 256, 1024)), dim3(256), 0, st,
                       (const SysDev*)c->sys, c->S, (const int32_t*)c->counts,
                       (const int32_t*)c->neprefix, c->C, c->inter);
    LAUNCH_CHECK("interactions");
    HIP_TRY(hipStreamSynchronize(st));
    std::vector<unsigned long long> h(n2);
    HIP_TRY(hipMemcpy(h.data(), c->inter, sizeof(unsigned long long) * n2, hipMemcpyDeviceToHost));
    int64_t ntot = 0;
    for (int s = 0; s < c->S; ++s) {
        out[2 * s] = (int64_t)h[2 * s] - c->sys_h[s].n;
        out[2 * s + 1] = (int64_t)h[2 * s + 1];
        ntot += c->sys_h[s].n;
    }
    out[2 * c->S] = (int64_t)h[2 * c->S] - ntot;
    out[2 * c->S + 1] = (int64_t)h[2 * c->S + 1];
    return BH_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- allocation-free step
// bh_step_noalloc: bh_step for one system with every per-step buffer carved
// out of one static device array per GPU, `scratch`, so the call makes no
// hipMalloc / hipFree (nor any other allocating call).  The array keeps the
// last call's contents; the setup kernel clears what must start at zero.
// Calls that overlap on one GPU share the array and race.
static __device__ __align__(256) unsigned char scratch[BH_NOALLOC_SCRATCH_BYTES];

#define NOALLOC_MAX_DEVICES 64
static std::mutex g_scratch_mu;
static unsigned char* g_scratch_addr[NOALLOC_MAX_DEVICES];   // device address of scratch
static int g_scratch_cu[NOALLOC_MAX_DEVICES];

// Device address of `scratch` on device dev (which must be current): looked up
// once per device with hipGetSymbolAddress (each GPU holds its own copy of the
// array at its own address; the array's name is only a host-side stand-in).
static int scratch_address(int dev, unsigned char** out, int* num_cu)
{
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES)
        return fail(BH_EINVAL, "device " + std::to_string(dev) + " out of range for the static scratch");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_addr[dev]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(BH_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        int cu = 0;
        if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            cu <= 0)
            cu = 256;
        g_scratch_cu[dev] = cu;
        g_scratch_addr[dev] = (unsigned char*)p;
    }
    *out = g_scratch_addr[dev];
    if (num_cu) *num_cu = g_scratch_cu[dev];
    return BH_OK;
}

// The one system's descriptor, every row in system 0, and the state a step
// reads before writing (stats, offsets, non-empty prefix) cleared.
__global__ void k_noalloc_setup(SysDev* sys, int32_t* sys_of, int64_t n, int32_t G, int32_t Gb,
                                int32_t* stats, int32_t* offsets, int32_t* neprefix, int64_t C)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    if (t == 0) {
        SysDev d{};
        d.n = n;
        d.G = G;
        d.Gb = Gb;
        sys[0] = d;
    }
    for (int64_t i = t; i < n; i += stride) sys_of[i] = 0;
    for (int64_t i = t; i <= C; i += stride) {
        offsets[i] = 0;
        neprefix[i] = 0;
    }
    for (int64_t i = t; i < ST_COUNT; i += stride) stats[i] = 0;
}

__global__ void k_copy_i32(const int32_t* src, int32_t* dst, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

extern "C" int bh_step_noalloc(const bh_particles* a, const bh_particles* b, int64_t n,
                               int32_t grid_dim, const bh_law* law, int32_t precision, double dt,
                               int32_t do_tick, const bh_tuning* tuning, double kappa,
                               int32_t* stats, void* stream)
{
    if (n < 1 || n > BH_NOALLOC_MAX_N)
        return fail(BH_EINVAL, "bh_step_noalloc: n = " + std::to_string(n) +
                                   " outside [1, " + std::to_string(BH_NOALLOC_MAX_N) +
                                   "], the static scratch's limit");
    if (grid_dim < 1 || grid_dim > BH_NOALLOC_MAX_GRID)
        return fail(BH_EINVAL, "bh_step_noalloc: grid_dim = " + std::to_string(grid_dim) +
                                   " outside [1, " + std::to_string(BH_NOALLOC_MAX_GRID) +
                                   "], the static scratch's limit");
    bh_tuning t = {2, 512, 32, 1, 0};
    if (tuning) t = *tuning;
    if (!(t.p == 1 || t.p == 2 || t.p == 4) || t.yspan < 1 || t.yspan > (1 << 20) ||
        t.waves_per_cu < 1 || t.waves_per_cu > 64 || t.tail_items < 0)
        return fail(BH_EINVAL, "bh_step_noalloc: invalid tuning");
    if (t.tail_split != 1)
        return fail(BH_EINVAL, "bh_step_noalloc: tail_split must be 1 (the slice planes do not fit "
                               "the static scratch)");
    if (std::isnan(kappa)) return fail(BH_EINVAL, "bh_step_noalloc: kappa is NaN");
    int r = check_parts(a, "a");
    if (r) return r;
    if ((r = check_parts(b, "b"))) return r;
    if (a->px == b->px) return fail(BH_EINVAL, "a and b must be different buffers");
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    unsigned char* base = nullptr;
    int num_cu = 256;
    if ((r = scratch_address(dev, &base, &num_cu))) return r;

    Ctx c;   // host-side view: every device pointer points into scratch
    c.device = dev;
    c.num_cu = num_cu;
    const int64_t G = grid_dim, Gb = cdiv(G, RC_BLOCK);
    set_sizes(&c, 1, n, G * G, G, n, Gb * Gb);
    size_t off = 0;
    r = for_each_buffer(&c, [&](auto*& p, int64_t cnt) {
        using T = std::remove_reference_t<decltype(*p)>;
        off = (off + 255) & ~(size_t)255;
        const size_t bytes = (size_t)std::max<int64_t>(cnt, 1) * sizeof(T);
        if (off + bytes > (size_t)BH_NOALLOC_SCRATCH_BYTES)
            return fail(BH_EINVAL, "bh_step_noalloc: static scratch too small for this size");
        p = (T*)(base + off);
        off += bytes;
        return (int)BH_OK;
    });
    if (r) return r;
    c.tune = t;
    c.rc_kappa = kappa > 0.0 ? kappa : 0.0;
    c.rc_deferred = 0;
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_noalloc_setup, dim3(grid_for(std::max<int64_t>(n, c.C + 1), 256, 1024)),
                       dim3(256), 0, st, c.sys, c.sys_of, n, (int32_t)G, (int32_t)Gb, c.stats,
                       c.offsets, c.neprefix, c.C);
    LAUNCH_CHECK("noalloc setup");
    r = step_impl(&c, a, b, law, precision, dt, do_tick, nullptr, nullptr, stream);
    if (r) return r;
    if (stats) {
        hipLaunchKernelGGL(k_copy_i32, dim3(1), dim3(64), 0, st, (const int32_t*)c.stats, stats,
                           (int)ST_COUNT);
        LAUNCH_CHECK("noalloc stats");
    }
    return BH_OK;
}

// Accessors for the cost-model translation unit.
Item* bh_ctx_items(void* ctx) { return ((Ctx*)ctx)->items; }
int32_t* bh_ctx_stats(void* ctx) { return ((Ctx*)ctx)->stats; }
int64_t bh_ctx_max_items(void* ctx) { return ((Ctx*)ctx)->max_items; }
int bh_ctx_num_cu(void* ctx) { return ((Ctx*)ctx)->num_cu; }
int bh_ctx_p(void* ctx) { return ((Ctx*)ctx)->tune.p; }
const SysDev* bh_ctx_sys(void* ctx) { return ((Ctx*)ctx)->sys; }
int bh_ctx_S(void* ctx) { return ((Ctx*)ctx)->S; }
const int32_t* bh_ctx_offsets(void* ctx) { return ((Ctx*)ctx)->offsets; }
const int32_t* bh_ctx_neprefix(void* ctx) { return ((Ctx*)ctx)->neprefix; }
const Com64* bh_ctx_l64(void* ctx) { return ((Ctx*)ctx)->l64; }
int32_t* bh_ctx_scratch_i32(void* ctx, int which)
{
    Ctx* c = (Ctx*)ctx;
    return which == 0 ? c->bsum : (which == 1 ? c->bne : c->cursor);
}
int32_t* bh_ctx_cost_scratch(void* ctx) { return ((Ctx*)ctx)->cost_scratch; }
float* bh_ctx_feat(void* ctx) { return ((Ctx*)ctx)->feat; }
float* bh_ctx_pred(void* ctx) { return ((Ctx*)ctx)->pred; }
int bh_ctx_fail(int code, const char* msg) { return fail(code, msg); }
