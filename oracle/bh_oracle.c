/*
 * bh_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C, FP64 restatement of python-bh's uniform-grid "Barnes-Hut"
 * step, used as the checker for the HIP path and as the timed CPU baseline.
 * Nothing in the product (barnes-hut_amd/) links or calls this file; only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg do.
 *
 * Each function cites the reference lines it restates (paths relative to
 * the reference repository root, python-bh/...).
 *
 * Law switches (the reference has three variants of the same 2-D law):
 *   a_i = Gc * sum_j m_j (x_j - x_i) * w(r2),  w(s) = 1/(s + eps2),
 *   and w(s) = 0 when r2_cut > 0 and s <= r2_cut.
 *   GPU reference:  eps2 = 0, r2_cut = 0   (kernels/grid_decomposition/gpu/grid.py:224-230)
 *   CPU reference:  eps2 = 0, r2_cut = 1e-5 (kernels/gravity/cpu_numba.py:36,44,59,67,81,89)
 *
 * Build: see oracle/Makefile (gcc -O3 -fopenmp, -ffp-contract=off so every
 * multiply and add rounds separately, like numpy).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ */
/* Grid sizing.                                                        */
/* python-bh/barneshut/implementations/base.py:221-231 (LEAF_OCCUPANCY  */
/* = 0.7 at :13) and kernels/helpers/grid.py:5-8 (next_perfect_square). */
/* ------------------------------------------------------------------ */
int bho_grid_dim(int64_t n, int64_t ppb)
{
    double nleaves;
    if (ppb == 1) {
        nleaves = 1.0;                      /* base.py:223-224 (ppb==1 => 1 leaf) */
    } else {
        nleaves = (double)n / (0.7 * (double)ppb);   /* base.py:226-227 */
    }
    /* next_perfect_square: if n % sqrt(n) != 0: ceil(sqrt(n))**2 */
    double s = sqrt(nleaves);
    if (fmod(nleaves, s) != 0.0) {
        double c = ceil(s);
        nleaves = c * c;
    }
    /* grid_dim = int(nleaves ** 0.5) */
    return (int)sqrt(nleaves);
}

/* Bounding box, helpers/grid.py:10-26; step, base.py:237.
 * out = {min_x, min_y, max_x, max_y, step} with the shorter edge's max
 * extended so the box is square (grid.py:19-22). */
void bho_bbox(int64_t n, const double *px, const double *py, int grid_dim, double *out)
{
    double mnx = px[0], mxx = px[0], mny = py[0], mxy = py[0];
    for (int64_t i = 1; i < n; ++i) {
        if (px[i] < mnx) mnx = px[i];
        if (px[i] > mxx) mxx = px[i];
        if (py[i] < mny) mny = py[i];
        if (py[i] > mxy) mxy = py[i];
    }
    double xe = mxx - mnx, ye = mxy - mny;
    if (xe > ye) {
        mxy += (xe - ye);
    } else if (ye > xe) {
        mxx += (ye - xe);
    }
    out[0] = mnx;
    out[1] = mny;
    out[2] = mxx;
    out[3] = mxy;
    out[4] = (mxx - mnx) / (double)grid_dim;   /* base.py:237 */
}

/* Cell ids, implementations/sequential.py:32-34:
 * g = clip(floor((p - min) / step), 0, G-1); key = gx*G + gy (x-major). */
void bho_place(int64_t n, const double *px, const double *py, double min_x, double min_y,
               double step, int G, int32_t *gx, int32_t *gy, int32_t *key)
{
    for (int64_t i = 0; i < n; ++i) {
        double fx = floor((px[i] - min_x) / step);
        double fy = floor((py[i] - min_y) / step);
        if (fx < 0.0) fx = 0.0;
        if (fx > (double)(G - 1)) fx = (double)(G - 1);
        if (fy < 0.0) fy = 0.0;
        if (fy > (double)(G - 1)) fy = (double)(G - 1);
        int32_t x = (int32_t)fx, y = (int32_t)fy;
        if (gx) gx[i] = x;
        if (gy) gy[i] = y;
        key[i] = x * G + y;
    }
}

/* Permutation of implementations/sequential.py:37,
 *   particles.view('f8,...').sort(order=['f7','f8'], axis=0, kind='stable')
 * numpy compares the named fields first, then every remaining field in
 * dtype order: px, py, mass, vx, vy, ax, ay, pid (internals/particle.py:3-9).
 * The final tie-break on the input index makes the sort stable. */
static const int32_t *s_key;
static const double *s_f[8];

static int cmp_rows(const void *pa, const void *pb)
{
    int64_t a = *(const int64_t *)pa, b = *(const int64_t *)pb;
    if (s_key[a] != s_key[b]) return s_key[a] < s_key[b] ? -1 : 1;
    for (int k = 0; k < 8; ++k) {
        double u = s_f[k][a], v = s_f[k][b];
        if (u < v) return -1;
        if (u > v) return 1;
    }
    return a < b ? -1 : (a > b ? 1 : 0);
}

/* fields: px, py, m, vx, vy, ax, ay, pid (as double). perm[q] = source row. */
void bho_sort_perm(int64_t n, const int32_t *key, const double *const *fields, int64_t *perm)
{
    s_key = key;
    for (int k = 0; k < 8; ++k) s_f[k] = fields[k];
    for (int64_t i = 0; i < n; ++i) perm[i] = i;
    qsort(perm, (size_t)n, sizeof(int64_t), cmp_rows);
}

/* Per-cell counts and exclusive offsets in x-major key order
 * (the np.unique ranges of sequential.py:39-50). offsets has ncells+1. */
void bho_counts(int64_t n, const int32_t *key, int64_t ncells, int64_t *counts, int64_t *offsets)
{
    memset(counts, 0, sizeof(int64_t) * (size_t)ncells);
    for (int64_t i = 0; i < n; ++i) counts[key[i]]++;
    offsets[0] = 0;
    for (int64_t c = 0; c < ncells; ++c) offsets[c + 1] = offsets[c] + counts[c];
}

/* Neumaier-compensated running sum. */
typedef struct { double s, c; } csum;
static inline void cs_add(csum *a, double v)
{
    double t = a->s + v;
    if (fabs(a->s) >= fabs(v)) a->c += (a->s - t) + v;
    else a->c += (v - t) + a->s;
    a->s = t;
}
static inline double cs_val(const csum *a) { return a->s + a->c; }

/* Per-cell centre of mass, internals/cloud.py:9-17 (calc_COM) and
 * gpu/grid.py:78-114 (g_summarize): M = sum m, X = sum m*x / M,
 * Y = sum m*y / M; empty cell -> (0, 0, 0). Sums are compensated so the
 * oracle is more accurate than either reference variant.
 * x, y, m are in sorted (cell) order; com is ncells x 3 (X, Y, M). */
void bho_summarize(int64_t ncells, const int64_t *offsets, const double *x, const double *y,
                   const double *m, double *com)
{
    for (int64_t c = 0; c < ncells; ++c) {
        com[3 * c + 0] = 0.0;
        com[3 * c + 1] = 0.0;
        com[3 * c + 2] = 0.0;
        int64_t a = offsets[c], b = offsets[c + 1];
        if (a == b) continue;
        csum M = {0, 0}, X = {0, 0}, Y = {0, 0};
        for (int64_t i = a; i < b; ++i) {
            cs_add(&X, x[i] * m[i]);
            cs_add(&Y, y[i] * m[i]);
            cs_add(&M, m[i]);
        }
        double mm = cs_val(&M);
        if (mm != 0.0) {
            com[3 * c + 0] = cs_val(&X) / mm;
            com[3 * c + 1] = cs_val(&Y) / mm;
            com[3 * c + 2] = mm;
        }
    }
}

/* w(s) of the law header above. */
static inline double law_w(double s, double eps2, double r2_cut)
{
    if (r2_cut > 0.0 && !(s > r2_cut)) return 0.0;
    return 1.0 / (s + eps2);
}

/* Accelerations of the grid method:
 *   near: every particle j != i of the 3x3 cell block around i's cell
 *         (gpu/grid.py:139-148, d_is_neighbor :187-192; CPU
 *          seq/evaluation.py:73-109 with helpers/grid.py:28-48);
 *   far:  the COM of every NON-EMPTY cell outside that block
 *         (seq/evaluation.py:86-95 skips empty cells; box.py:15-24).
 * Inputs are in sorted order (x-major cell key, offsets from bho_counts).
 * targets: sorted indices to evaluate (NULL = all n). Output ax/ay are
 * indexed like targets (or like particles when targets == NULL).
 * Uses OpenMP over targets when built with -fopenmp (nthreads <= 0: default). */
void bho_forces(int G, const int64_t *offsets, const int32_t *key, int64_t n,
                const double *x, const double *y, const double *m, const double *com,
                double Gc, double eps2, double r2_cut,
                const int64_t *targets, int64_t ntargets, int nthreads,
                double *ax, double *ay)
{
    int64_t ncells = (int64_t)G * G;
    /* compact list of non-empty cells in key order */
    int64_t K = 0;
    for (int64_t c = 0; c < ncells; ++c) K += (offsets[c + 1] > offsets[c]);
    size_t kb = (size_t)(K > 0 ? K : 1);
    int32_t *ne = (int32_t *)malloc(sizeof(int32_t) * kb);
    int32_t *nkx = (int32_t *)malloc(sizeof(int32_t) * kb);
    int32_t *nky = (int32_t *)malloc(sizeof(int32_t) * kb);
    double *nX = (double *)malloc(sizeof(double) * kb);
    double *nY = (double *)malloc(sizeof(double) * kb);
    double *nM = (double *)malloc(sizeof(double) * kb);
    K = 0;
    for (int64_t c = 0; c < ncells; ++c)
        if (offsets[c + 1] > offsets[c]) {
            ne[K] = (int32_t)c;
            nkx[K] = (int32_t)(c / G);
            nky[K] = (int32_t)(c % G);
            nX[K] = com[3 * c];
            nY[K] = com[3 * c + 1];
            nM[K] = com[3 * c + 2];
            ++K;
        }
    int64_t nt = targets ? ntargets : n;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64)
#endif
    for (int64_t t = 0; t < nt; ++t) {
        int64_t i = targets ? targets[t] : t;
        int32_t c = key[i];
        int cx = c / G, cy = c % G;
        double xi = x[i], yi = y[i];
        double sx = 0.0, sy = 0.0;
        /* near field */
        for (int dx = -1; dx <= 1; ++dx) {
            int nx = cx + dx;
            if (nx < 0 || nx >= G) continue;
            for (int dy = -1; dy <= 1; ++dy) {
                int ny = cy + dy;
                if (ny < 0 || ny >= G) continue;
                int64_t cc = (int64_t)nx * G + ny;
                for (int64_t j = offsets[cc]; j < offsets[cc + 1]; ++j) {
                    if (j == i) continue;
                    double wx = x[j] - xi, wy = y[j] - yi;
                    double s = wx * wx + wy * wy;
                    double w = law_w(s, eps2, r2_cut) * m[j];
                    sx += wx * w;
                    sy += wy * w;
                }
            }
        }
        /* far field */
        for (int64_t k = 0; k < K; ++k) {
            if (abs(nkx[k] - cx) <= 1 && abs(nky[k] - cy) <= 1) continue;
            double wx = nX[k] - xi, wy = nY[k] - yi;
            double s = wx * wx + wy * wy;
            double w = law_w(s, eps2, r2_cut) * nM[k];
            sx += wx * w;
            sy += wy * w;
        }
        ax[t] = Gc * sx;
        ay[t] = Gc * sy;
    }
    free(ne);
    free(nkx);
    free(nky);
    free(nX);
    free(nY);
    free(nM);
}

/* O(S*N) direct sum, implementations/base.py:16-32 (nsquared_acc_check):
 * every j != i (self skipped by index), same law switches. */
void bho_direct(int64_t n, const double *x, const double *y, const double *m,
                double Gc, double eps2, double r2_cut,
                const int64_t *targets, int64_t ntargets, int nthreads,
                double *ax, double *ay)
{
    int64_t nt = targets ? ntargets : n;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16)
#endif
    for (int64_t t = 0; t < nt; ++t) {
        int64_t i = targets ? targets[t] : t;
        double xi = x[i], yi = y[i];
        csum sx = {0, 0}, sy = {0, 0};
        for (int64_t j = 0; j < n; ++j) {
            if (j == i) continue;
            double wx = x[j] - xi, wy = y[j] - yi;
            double s = wx * wx + wy * wy;
            double w = law_w(s, eps2, r2_cut) * m[j];
            cs_add(&sx, wx * w);
            cs_add(&sy, wy * w);
        }
        ax[t] = Gc * cs_val(&sx);
        ay[t] = Gc * cs_val(&sy);
    }
}

/* Energy of a state under the law above (no reference counterpart: python-bh
 * reports no energy; this pins the energy diagnostic of the HIP library).
 * U = Gc * sum_{i<j} m_i m_j * log(r_ij^2 + eps2) / 2  (the potential whose
 * negative gradient is the eps2-softened law, r2_cut ignored), K = sum m v^2 / 2.
 * Compensated sums; the pair loop visits each unordered pair once. */
void bho_energy(int64_t n, const double *x, const double *y, const double *m,
                const double *vx, const double *vy, double Gc, double eps2, int nthreads,
                double *out2)
{
    double *row = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16)
#endif
    for (int64_t i = 0; i < n; ++i) {
        csum u = {0, 0};
        for (int64_t j = i + 1; j < n; ++j) {
            double wx = x[j] - x[i], wy = y[j] - y[i];
            cs_add(&u, m[j] * log(wx * wx + wy * wy + eps2));
        }
        row[i] = m[i] * cs_val(&u);
    }
    csum U = {0, 0}, K = {0, 0};
    for (int64_t i = 0; i < n; ++i) {
        cs_add(&U, row[i]);
        cs_add(&K, m[i] * (vx[i] * vx[i] + vy[i] * vy[i]));
    }
    free(row);
    out2[0] = 0.5 * Gc * cs_val(&U);
    out2[1] = 0.5 * cs_val(&K);
}

/* Potential of the law evaluated by the grid method (no reference
 * counterpart; it serves the energy diagnostic of large runs): for each
 * target i, phi_i = sum_{j != i in the 3x3 cells} m_j log(r_ij^2 + eps2)
 *                 + sum_{non-empty cells k outside} M_k log(|X_k - x_i|^2 + eps2),
 * the decomposition of bho_forces with log(s + eps2) in place of
 * (x_j - x_i) / (s + eps2).  U_grid = Gc / 4 * sum_i m_i phi_i.  Compensated
 * sums per target. */
void bho_potential(int G, const int64_t *offsets, const int32_t *key, int64_t n,
                   const double *x, const double *y, const double *m, const double *com,
                   double eps2, const int64_t *targets, int64_t ntargets, int nthreads,
                   double *phi)
{
    int64_t ncells = (int64_t)G * G;
    int64_t K = 0;
    for (int64_t c = 0; c < ncells; ++c) K += (offsets[c + 1] > offsets[c]);
    size_t kb = (size_t)(K > 0 ? K : 1);
    int32_t *kx = (int32_t *)malloc(sizeof(int32_t) * kb);
    int32_t *ky = (int32_t *)malloc(sizeof(int32_t) * kb);
    double *cX = (double *)malloc(sizeof(double) * kb);
    double *cY = (double *)malloc(sizeof(double) * kb);
    double *cM = (double *)malloc(sizeof(double) * kb);
    K = 0;
    for (int64_t c = 0; c < ncells; ++c)
        if (offsets[c + 1] > offsets[c]) {
            kx[K] = (int32_t)(c / G);
            ky[K] = (int32_t)(c % G);
            cX[K] = com[3 * c];
            cY[K] = com[3 * c + 1];
            cM[K] = com[3 * c + 2];
            ++K;
        }
    int64_t nt = targets ? ntargets : n;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64)
#endif
    for (int64_t t = 0; t < nt; ++t) {
        int64_t i = targets ? targets[t] : t;
        int cx = key[i] / G, cy = key[i] % G;
        double xi = x[i], yi = y[i];
        csum s = {0, 0};
        for (int dx = -1; dx <= 1; ++dx) {
            int nx = cx + dx;
            if (nx < 0 || nx >= G) continue;
            for (int dy = -1; dy <= 1; ++dy) {
                int ny = cy + dy;
                if (ny < 0 || ny >= G) continue;
                int64_t cc = (int64_t)nx * G + ny;
                for (int64_t j = offsets[cc]; j < offsets[cc + 1]; ++j) {
                    if (j == i) continue;
                    double wx = x[j] - xi, wy = y[j] - yi;
                    cs_add(&s, m[j] * log(wx * wx + wy * wy + eps2));
                }
            }
        }
        for (int64_t k = 0; k < K; ++k) {
            if (abs(kx[k] - cx) <= 1 && abs(ky[k] - cy) <= 1) continue;
            double wx = cX[k] - xi, wy = cY[k] - yi;
            cs_add(&s, cM[k] * log(wx * wx + wy * wy + eps2));
        }
        phi[t] = cs_val(&s);
    }
    free(kx);
    free(ky);
    free(cX);
    free(cY);
    free(cM);
}

/* Cloud-to-cloud gravity of the reference's default gravity plugin,
 * kernels/gravity/cpu_numba.py:8-92 (cpu_numba_kernel), in its loop order:
 *   same != 0: self_self_numba (:30-47): every pair i < j once; with
 *              d = x_j - x_i and f = Gc * d * w(|d|^2), row i gets f m_j
 *              and row j gets -f m_i;
 *   same == 0: self_other_numba_2 (:49-70) if update_other, else
 *              self_other_numba_1 (:72-92): for every other row i (outer) and
 *              self row j (inner), d = x_j - y_i and f = Gc * d * w(|d|^2);
 *              self row j gets -f M_i and, with update_other, other row i
 *              gets f m_j.
 * w = law_w above (the reference's CPU law: r2_cut = 1e-5, eps2 = 0).  Rows
 * are n x ld doubles with px, py, mass in columns 0, 1, 2.  The sums are
 * WRITTEN to accA[2 na] and accB[2 nb] (x, y interleaved; accB untouched
 * unless update_other) for the caller to add, as cpu_numba_kernel adds its
 * fields to the clouds' accelerations.  Plain FP64 sums. */
void bho_cloud_gravity(int64_t na, const double *A, int64_t lda, int64_t nb, const double *B,
                       int64_t ldb, int same, double Gc, double eps2, double r2_cut,
                       int update_other, double *accA, double *accB)
{
    memset(accA, 0, sizeof(double) * 2 * (size_t)na);
    if (same) {
        for (int64_t i = 0; i < na; ++i)
            for (int64_t j = i + 1; j < na; ++j) {
                double dx = A[j * lda] - A[i * lda], dy = A[j * lda + 1] - A[i * lda + 1];
                double w = law_w(dx * dx + dy * dy, eps2, r2_cut);
                double fx = Gc * dx * w, fy = Gc * dy * w;
                accA[2 * i] += fx * A[j * lda + 2];
                accA[2 * i + 1] += fy * A[j * lda + 2];
                accA[2 * j] -= fx * A[i * lda + 2];
                accA[2 * j + 1] -= fy * A[i * lda + 2];
            }
        return;
    }
    if (update_other) memset(accB, 0, sizeof(double) * 2 * (size_t)nb);
    for (int64_t i = 0; i < nb; ++i)
        for (int64_t j = 0; j < na; ++j) {
            double dx = A[j * lda] - B[i * ldb], dy = A[j * lda + 1] - B[i * ldb + 1];
            double w = law_w(dx * dx + dy * dy, eps2, r2_cut);
            double fx = Gc * dx * w, fy = Gc * dy * w;
            accA[2 * j] -= fx * B[i * ldb + 2];
            accA[2 * j + 1] -= fy * B[i * ldb + 2];
            if (update_other) {
                accB[2 * i] += fx * A[j * lda + 2];
                accB[2 * i + 1] += fy * A[j * lda + 2];
            }
        }
}

/* Symplectic-Euler tick, gpu/grid.py:159-174 (g_tick_particles):
 * v += a*dt; a = 0; x += v*dt. Separate multiply and add (no FMA). */
void bho_tick(int64_t n, double *x, double *y, double *vx, double *vy, double *ax, double *ay,
              double dt)
{
    for (int64_t i = 0; i < n; ++i) {
        vx[i] += ax[i] * dt;
        vy[i] += ay[i] * dt;
        ax[i] = 0.0;
        ay[i] = 0.0;
        x[i] += vx[i] * dt;
        y[i] += vy[i] * dt;
    }
}

/* Exact interaction counts of one evaluation (the work basis):
 * near = sum_c n_c * (sum_{c' in 3x3(c)} n_c') - N
 * far  = sum_c n_c * #{non-empty c' outside 3x3(c)}. out = {near, far}. */
void bho_interactions(int G, const int64_t *counts, int64_t *out)
{
    int64_t K = 0, N = 0, near = 0, far = 0;
    int64_t ncells = (int64_t)G * G;
    for (int64_t c = 0; c < ncells; ++c) {
        K += counts[c] > 0;
        N += counts[c];
    }
    for (int cx = 0; cx < G; ++cx)
        for (int cy = 0; cy < G; ++cy) {
            int64_t nc = counts[(int64_t)cx * G + cy];
            if (!nc) continue;
            int64_t n9 = 0, k9 = 0;
            for (int dx = -1; dx <= 1; ++dx)
                for (int dy = -1; dy <= 1; ++dy) {
                    int nx = cx + dx, ny = cy + dy;
                    if (nx < 0 || ny < 0 || nx >= G || ny >= G) continue;
                    int64_t cc = counts[(int64_t)nx * G + ny];
                    n9 += cc;
                    k9 += cc > 0;
                }
            near += nc * n9;
            far += nc * (K - k9);
        }
    out[0] = near - N;
    out[1] = far;
}
