/*
 * tsph_oracle.c -- CPU restatement of TensorSPH's per-step particle pipeline.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle: the GPU
 * library is checked against it in tests/, in __graft_entry__.smoke() and
 * timed as the CPU baseline in bench.py.  Nothing in the product package
 * (tensorsph_amd/) links, loads or calls it.
 *
 * Every routine restates one reference function in the reference's own pair
 * formulation (unordered pairs i < j scattered to both ends in pair order),
 * in plain C, serial.  Citations are file:line in cxs503/TensorSPH.
 *
 * Pinning (what checks each routine independently of its own formula):
 *   pairs (cell-linked, periodic) ... brute force O(N^2) + reference golden
 *                                     pair lists (tests/golden/)
 *   kernels ......................... closed-form normalisation/quadrature
 *   rhs_morris (2-D rhs, rhs_3d) .... dense all-pairs restatement (2-D, no
 *                                     walls), reference golden rho/p/acc
 *                                     with walls, Newton-3 wall reactions
 *   threed_accelerations ............ reference golden vectors, momentum
 *                                     antisymmetry, calibrated-lattice rho
 *   delta_plus_step ................. reference golden vectors, closed-form
 *                                     interior shift, affine invariance
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off, no fast-math).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_API __attribute__((visibility("default")))

/* ------------------------------------------------------------------------ */
/* Kernels                                                                   */
/* ------------------------------------------------------------------------ */

/* 2-D cubic spline, droplet_impact.py:276-288 / pair_kernels :349-371.
 * sigma = 10/(7 pi h^2); returns W and the gradient scale s with
 * grad_i W = delta * s. */
static void cubic2d(double r, double h, double* w, double* s) {
    double q = r / h;
    double sigma = 10.0 / (7.0 * M_PI * h * h);
    double first = 1.0 - 1.5 * (q * q) + 0.75 * (q * q * q);
    double t = 2.0 - q; if (t < 0.0) t = 0.0;
    double second = 0.25 * (t * t * t);
    double shape = q < 1.0 ? first : (q < 2.0 ? second : 0.0);
    double dshape = q < 1.0 ? (-3.0 * q + 2.25 * (q * q))
                            : (q < 2.0 ? -0.75 * ((2.0 - q) * (2.0 - q)) : 0.0);
    if (shape < 0.0) shape = 0.0;
    *w = sigma * shape;
    double rc = r > 1.0e-12 ? r : 1.0e-12;
    *s = sigma * dshape / (h * rc);
}

/* 3-D cubic spline, threed/kernel3.py:48-181 (sigma = 1/(pi h^3)), with the
 * radial Laplacian f'' + 2 f'/q and its q -> 0 limit -6 (kernel3.py:78-87). */
static void cubic3d(double r, double h, double* w, double* s, double* lap) {
    double q = r / h;
    double sig = 1.0 / (M_PI * h * h * h);
    double first = 1.0 - 1.5 * (q * q) + 0.75 * (q * q * q);
    double t = 2.0 - q; if (t < 0.0) t = 0.0;
    double second = 0.25 * (t * t * t);
    double shape = q < 1.0 ? first : (q < 2.0 ? second : 0.0);
    if (shape < 0.0) shape = 0.0;
    double fp = q < 1.0 ? (-3.0 * q + 2.25 * (q * q))
                        : (q < 2.0 ? -0.75 * ((2.0 - q) * (2.0 - q)) : 0.0);
    double fpp = q < 1.0 ? (-3.0 + 4.5 * q) : (q < 2.0 ? 1.5 * (2.0 - q) : 0.0);
    double qc = q > 1.0e-12 ? q : 1.0e-12;
    double twice = q > 1.0e-10 ? 2.0 * fp / qc : -6.0;
    double radial = fpp + twice;
    *w = sig * shape;
    double rc = r > 1.0e-12 ? r : 1.0e-12;
    *s = sig * fp / (h * rc);
    if (lap) *lap = q < 2.0 ? sig * radial / (h * h) : 0.0;
}

/* Wendland C2 (2-D), droplet_impact.py:320-346: sigma = 7/(4 pi h^2). */
static void wendland2d(double r, double h, double* w, double* s) {
    double q = r / h;
    double sigma = 7.0 / (4.0 * M_PI * h * h);
    double t = 1.0 - 0.5 * q; if (t < 0.0) t = 0.0;
    double t2 = t * t;
    *w = sigma * (t2 * t2) * (2.0 * q + 1.0);
    double dshape = -5.0 * q * (t2 * t);
    double rc = r > 1.0e-12 ? r : 1.0e-12;
    *s = sigma * dshape / (h * rc);
}

ORC_API void orc_kernel_eval(int kind, const double* r, int64_t n, double h,
                             double* w, double* s, double* lap) {
    for (int64_t k = 0; k < n; ++k) {
        double ww, ss, ll = 0.0;
        if (kind == 0) cubic2d(r[k], h, &ww, &ss);
        else if (kind == 1) cubic3d(r[k], h, &ww, &ss, &ll);
        else wendland2d(r[k], h, &ww, &ss);
        w[k] = ww; s[k] = ss; if (lap) lap[k] = ll;
    }
}

/* ------------------------------------------------------------------------ */
/* Sorting helpers                                                           */
/* ------------------------------------------------------------------------ */

typedef struct { int64_t key; int64_t idx; } kv_t;

static int kv_cmp(const void* a, const void* b) {
    const kv_t* x = (const kv_t*)a; const kv_t* y = (const kv_t*)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    if (x->idx != y->idx) return x->idx < y->idx ? -1 : 1;
    return 0;
}

typedef struct { int64_t i, j; } pr_t;
static int pr_cmp(const void* a, const void* b) {
    const pr_t* x = (const pr_t*)a; const pr_t* y = (const pr_t*)b;
    if (x->i != y->i) return x->i < y->i ? -1 : 1;
    if (x->j != y->j) return x->j < y->j ? -1 : 1;
    return 0;
}

typedef struct { pr_t* v; int64_t n, cap; } prvec_t;
static void prvec_push(prvec_t* p, int64_t i, int64_t j) {
    if (p->n == p->cap) {
        p->cap = p->cap ? p->cap * 2 : 1024;
        p->v = (pr_t*)realloc(p->v, (size_t)p->cap * sizeof(pr_t));
    }
    p->v[p->n].i = i; p->v[p->n].j = j; p->n++;
}

static int64_t finish_pairs(prvec_t* pv, int64_t** out) {
    qsort(pv->v, (size_t)pv->n, sizeof(pr_t), pr_cmp);
    int64_t* o = (int64_t*)malloc((size_t)(pv->n > 0 ? pv->n : 1) * 2 * sizeof(int64_t));
    for (int64_t k = 0; k < pv->n; ++k) { o[2 * k] = pv->v[k].i; o[2 * k + 1] = pv->v[k].j; }
    free(pv->v);
    *out = o;
    return pv->n;
}

ORC_API void orc_free(void* p) { free(p); }

static int64_t bsearch_key(const int64_t* keys, int64_t m, int64_t key) {
    int64_t lo = 0, hi = m;
    while (lo < hi) { int64_t mid = (lo + hi) >> 1; if (keys[mid] < key) lo = mid + 1; else hi = mid; }
    return (lo < m && keys[lo] == key) ? lo : -1;
}

/* ------------------------------------------------------------------------ */
/* Linked-cell pairs, neighbor_search.py:44-120.  Arithmetic in the
 * positions' own precision (T), as torch does: x / (T)cutoff, floor, and
 * d^2 = ((dx^2 + dy^2) + dz^2) compared with (T)(cutoff*cutoff).          */
/* ------------------------------------------------------------------------ */

#define DEFINE_CLP(T, SUF, FLOOR)                                                      \
ORC_API int64_t orc_cell_linked_pairs_##SUF(const T* pos, int64_t n, int dim,          \
        double cutoff, int64_t** out, int64_t* cells_out, int64_t* keys_out,           \
        int64_t* order_out) {                                                          \
    prvec_t pv = {0, 0, 0};                                                            \
    if (n < 2) { *out = (int64_t*)malloc(16); return 0; }                              \
    T c = (T)cutoff;                                                                   \
    int64_t* cells = (int64_t*)malloc((size_t)n * dim * sizeof(int64_t));              \
    int64_t cmin[3], cmax[3], ext[3], stride[3];                                       \
    for (int d = 0; d < dim; ++d) { cmin[d] = INT64_MAX; cmax[d] = INT64_MIN; }        \
    for (int64_t i = 0; i < n; ++i)                                                    \
        for (int d = 0; d < dim; ++d) {                                                \
            T len = pos[i * dim + d] / c;                                              \
            int64_t cc = (int64_t)FLOOR(len);                                          \
            cells[i * dim + d] = cc;                                                   \
            if (cc < cmin[d]) cmin[d] = cc;                                            \
            if (cc > cmax[d]) cmax[d] = cc;                                            \
        }                                                                              \
    for (int d = 0; d < dim; ++d) ext[d] = cmax[d] - cmin[d] + 1;                      \
    stride[dim - 1] = 1;                                                               \
    for (int d = dim - 2; d >= 0; --d) stride[d] = stride[d + 1] * ext[d + 1];         \
    kv_t* kv = (kv_t*)malloc((size_t)n * sizeof(kv_t));                                \
    for (int64_t i = 0; i < n; ++i) {                                                  \
        int64_t key = 0;                                                               \
        for (int d = 0; d < dim; ++d) {                                                \
            cells[i * dim + d] -= cmin[d];                                             \
            key += cells[i * dim + d] * stride[d];                                     \
        }                                                                              \
        kv[i].key = key; kv[i].idx = i;                                                \
        if (keys_out) keys_out[i] = key;                                               \
    }                                                                                  \
    if (cells_out) memcpy(cells_out, cells, (size_t)n * dim * sizeof(int64_t));        \
    qsort(kv, (size_t)n, sizeof(kv_t), kv_cmp);                                        \
    if (order_out) for (int64_t i = 0; i < n; ++i) order_out[i] = kv[i].idx;           \
    int64_t* ukeys = (int64_t*)malloc((size_t)n * sizeof(int64_t));                    \
    int64_t* ustart = (int64_t*)malloc((size_t)(n + 1) * sizeof(int64_t));             \
    int64_t m = 0;                                                                     \
    for (int64_t i = 0; i < n; ++i)                                                    \
        if (i == 0 || kv[i].key != kv[i - 1].key) { ukeys[m] = kv[i].key; ustart[m] = i; ++m; } \
    ustart[m] = n;                                                                     \
    int noff = 1; for (int d = 0; d < dim; ++d) noff *= 3;                             \
    T c2 = (T)(cutoff * cutoff);                                                       \
    for (int64_t i = 0; i < n; ++i) {                                                  \
        for (int o = 0; o < noff; ++o) {                                               \
            int rem = o, inside = 1; int64_t nk = 0;                                   \
            int64_t off[3];                                                            \
            for (int d = dim - 1; d >= 0; --d) { off[d] = rem % 3 - 1; rem /= 3; }     \
            for (int d = 0; d < dim; ++d) {                                            \
                int64_t nb = cells[i * dim + d] + off[d];                              \
                if (nb < 0 || nb >= ext[d]) inside = 0;                                \
                nk += (nb < 0 ? 0 : nb) * stride[d];                                   \
            }                                                                          \
            if (!inside) continue;                                                     \
            int64_t slot = bsearch_key(ukeys, m, nk);                                  \
            if (slot < 0) continue;                                                    \
            for (int64_t s = ustart[slot]; s < ustart[slot + 1]; ++s) {                \
                int64_t j = kv[s].idx;                                                 \
                if (!(i < j)) continue;                                                \
                T acc = 0;                                                             \
                for (int d = 0; d < dim; ++d) {                                        \
                    T df = pos[i * dim + d] - pos[j * dim + d];                        \
                    T sq = df * df;                                                    \
                    acc = d == 0 ? sq : acc + sq;                                      \
                }                                                                      \
                if (acc <= c2) prvec_push(&pv, i, j);                                  \
            }                                                                          \
        }                                                                              \
    }                                                                                  \
    free(cells); free(kv); free(ukeys); free(ustart);                                  \
    return finish_pairs(&pv, out);                                                     \
}

DEFINE_CLP(double, f64, floor)
DEFINE_CLP(float, f32, floorf)

/* Brute force, neighbor_search.py:29-41. */
#define DEFINE_BF(T, SUF)                                                              \
ORC_API int64_t orc_brute_force_pairs_##SUF(const T* pos, int64_t n, int dim,          \
                                            double cutoff, int64_t** out) {            \
    prvec_t pv = {0, 0, 0};                                                            \
    T c2 = (T)(cutoff * cutoff);                                                       \
    for (int64_t i = 0; i < n; ++i)                                                    \
        for (int64_t j = i + 1; j < n; ++j) {                                          \
            T acc = 0;                                                                 \
            for (int d = 0; d < dim; ++d) {                                            \
                T df = pos[i * dim + d] - pos[j * dim + d];                            \
                T sq = df * df; acc = d == 0 ? sq : acc + sq;                          \
            }                                                                          \
            if (acc <= c2) prvec_push(&pv, i, j);                                      \
        }                                                                              \
    return finish_pairs(&pv, out);                                                     \
}
DEFINE_BF(double, f64)
DEFINE_BF(float, f32)

/* Minimum image, threed/neighbors3.py:45-53 (torch.round = half to even). */
static inline double min_image(double d, double L) { return d - L * nearbyint(d / L); }

/* Periodic pairs, threed/neighbors3.py:71-142. */
ORC_API int64_t orc_periodic_pairs_f64(const double* pos, int64_t n, int dim,
                                       const double* box, double cutoff, int64_t** out) {
    prvec_t pv = {0, 0, 0};
    if (n < 2) { *out = (int64_t*)malloc(16); return 0; }
    int64_t ncell[3], stride[3];
    double width[3];
    for (int d = 0; d < dim; ++d) {
        int64_t nc = (int64_t)floor(box[d] / cutoff);
        ncell[d] = nc < 3 ? 3 : nc;
        width[d] = box[d] / (double)ncell[d];
    }
    stride[dim - 1] = 1;
    for (int d = dim - 2; d >= 0; --d) stride[d] = stride[d + 1] * ncell[d + 1];
    int64_t* cells = (int64_t*)malloc((size_t)n * dim * sizeof(int64_t));
    kv_t* kv = (kv_t*)malloc((size_t)n * sizeof(kv_t));
    for (int64_t i = 0; i < n; ++i) {
        int64_t key = 0;
        for (int d = 0; d < dim; ++d) {
            int64_t c = (int64_t)floor(pos[i * dim + d] / width[d]);
            c %= ncell[d]; if (c < 0) c += ncell[d];
            cells[i * dim + d] = c; key += c * stride[d];
        }
        kv[i].key = key; kv[i].idx = i;
    }
    qsort(kv, (size_t)n, sizeof(kv_t), kv_cmp);
    int64_t* ukeys = (int64_t*)malloc((size_t)n * sizeof(int64_t));
    int64_t* ustart = (int64_t*)malloc((size_t)(n + 1) * sizeof(int64_t));
    int64_t m = 0;
    for (int64_t i = 0; i < n; ++i)
        if (i == 0 || kv[i].key != kv[i - 1].key) { ukeys[m] = kv[i].key; ustart[m] = i; ++m; }
    ustart[m] = n;
    int noff = 1; for (int d = 0; d < dim; ++d) noff *= 3;
    double c2 = cutoff * cutoff;
    for (int64_t i = 0; i < n; ++i) {
        for (int o = 0; o < noff; ++o) {
            int rem = o; int64_t nk = 0; int64_t off[3];
            for (int d = dim - 1; d >= 0; --d) { off[d] = rem % 3 - 1; rem /= 3; }
            for (int d = 0; d < dim; ++d) {
                int64_t nb = cells[i * dim + d] + off[d];
                nb %= ncell[d]; if (nb < 0) nb += ncell[d];
                nk += nb * stride[d];
            }
            int64_t slot = bsearch_key(ukeys, m, nk);
            if (slot < 0) continue;
            for (int64_t s = ustart[slot]; s < ustart[slot + 1]; ++s) {
                int64_t j = kv[s].idx;
                if (!(i < j)) continue;
                double acc = 0;
                for (int d = 0; d < dim; ++d) {
                    double df = min_image(pos[i * dim + d] - pos[j * dim + d], box[d]);
                    double sq = df * df; acc = d == 0 ? sq : acc + sq;
                }
                if (acc <= c2) prvec_push(&pv, i, j);
            }
        }
    }
    free(cells); free(kv); free(ukeys); free(ustart);
    return finish_pairs(&pv, out);
}

/* ------------------------------------------------------------------------ */
/* Small symmetric eigen-solver (cyclic Jacobi) used for the renormalised    */
/* density gradient (density_diffusion.py:95-142) and the delta-plus first  */
/* moment (delta_plus.py:34-49).  Only the lower triangle is read, as       */
/* torch.linalg.eigh(UPLO='L') does.                                        */
/* ------------------------------------------------------------------------ */

static void sym_eig(int n, const double* a_in, double* w, double* v) {
    double a[9];
    for (int r = 0; r < n; ++r)
        for (int c = 0; c < n; ++c)
            a[r * n + c] = r >= c ? a_in[r * n + c] : a_in[c * n + r];
    for (int r = 0; r < n; ++r) for (int c = 0; c < n; ++c) v[r * n + c] = r == c ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 60; ++sweep) {
        double off = 0.0;
        for (int p = 0; p < n; ++p) for (int q = p + 1; q < n; ++q) off += a[p * n + q] * a[p * n + q];
        if (off < 1e-300) break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                double apq = a[p * n + q];
                if (apq == 0.0) continue;
                double theta = (a[q * n + q] - a[p * n + p]) / (2.0 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; ++k) {
                    double akp = a[k * n + p], akq = a[k * n + q];
                    a[k * n + p] = c * akp - s * akq; a[k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {
                    double apk = a[p * n + k], aqk = a[q * n + k];
                    a[p * n + k] = c * apk - s * aqk; a[q * n + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    double vkp = v[k * n + p], vkq = v[k * n + q];
                    v[k * n + p] = c * vkp - s * vkq; v[k * n + q] = s * vkp + c * vkq;
                }
            }
    }
    for (int k = 0; k < n; ++k) w[k] = a[k * n + k];
    /* ascending order, as eigh returns */
    for (int x = 0; x < n; ++x)
        for (int y = x + 1; y < n; ++y)
            if (w[y] < w[x]) {
                double t = w[x]; w[x] = w[y]; w[y] = t;
                for (int k = 0; k < n; ++k) { t = v[k * n + x]; v[k * n + x] = v[k * n + y]; v[k * n + y] = t; }
            }
}

ORC_API void orc_sym_eig(int n, const double* a, double* w, double* v) { sym_eig(n, a, w, v); }

/* ------------------------------------------------------------------------ */
/* Density diffusion on fluid-fluid pairs, density_diffusion.py:78-197.      */
/* pair arrays: i, j (fluid indices), delta (P x dim), grad (P x dim).       */
/* ------------------------------------------------------------------------ */

static void diffusion_rate(int dim, int64_t n, const double* rho, const double* mass,
                           int64_t P, const int64_t* pi, const int64_t* pj,
                           const double* delta, const double* grad, double scale,
                           int form, double eig_floor, double* rate) {
    for (int64_t a = 0; a < n; ++a) rate[a] = 0.0;
    double* gr = NULL;
    if (form == 1) { /* antuono: renormalised gradient */
        double* G = (double*)calloc((size_t)n * dim, sizeof(double));
        double* M = (double*)calloc((size_t)n * 9, sizeof(double));
        for (int pass = 0; pass < 2; ++pass)
            for (int64_t k = 0; k < P; ++k) {
                int64_t i = pi[k], j = pj[k];
                int64_t a = pass == 0 ? i : j;
                double vol = pass == 0 ? mass[j] / rho[j] : mass[i] / rho[i];
                double diff = rho[j] - rho[i];
                for (int d = 0; d < dim; ++d) G[a * dim + d] += (vol * diff) * grad[k * dim + d];
                for (int r = 0; r < dim; ++r)
                    for (int c = 0; c < dim; ++c)
                        M[a * 9 + r * dim + c] += vol * (-(delta[k * dim + r] * grad[k * dim + c]));
            }
        gr = (double*)calloc((size_t)n * dim, sizeof(double));
        for (int64_t a = 0; a < n; ++a) {
            double tr = 0; for (int d = 0; d < dim; ++d) tr += M[a * 9 + d * dim + d];
            if (tr <= 0.0) continue;
            double w[3], v[9];
            sym_eig(dim, &M[a * 9], w, v);
            double trace = 0; for (int d = 0; d < dim; ++d) trace += w[d];
            double mean = trace / dim; if (mean < 1e-30) mean = 1e-30;
            double fl = eig_floor * mean;
            double proj[3];
            for (int k = 0; k < dim; ++k) {
                double s = 0; for (int r = 0; r < dim; ++r) s += v[r * dim + k] * G[a * dim + r];
                proj[k] = s / (w[k] > fl ? w[k] : fl);
            }
            for (int r = 0; r < dim; ++r) { double s = 0; for (int k = 0; k < dim; ++k) s += v[r * dim + k] * proj[k]; gr[a * dim + r] = s; }
        }
        free(G); free(M);
    }
    double* coup = (double*)malloc((size_t)(P > 0 ? P : 1) * sizeof(double));
    for (int64_t k = 0; k < P; ++k) {
        int64_t i = pi[k], j = pj[k];
        double radial = 0, d2 = 0;
        for (int d = 0; d < dim; ++d) { radial += delta[k * dim + d] * grad[k * dim + d]; d2 += delta[k * dim + d] * delta[k * dim + d]; }
        radial = -radial;
        double contrast = rho[i] - rho[j];
        if (form == 1) {
            double proj = 0; for (int d = 0; d < dim; ++d) proj += (gr[i * dim + d] + gr[j * dim + d]) * delta[k * dim + d];
            contrast = contrast - 0.5 * proj;
        }
        coup[k] = contrast * radial / (d2 + 1.0e-12);
    }
    for (int64_t k = 0; k < P; ++k) rate[pi[k]] += -scale * (mass[pj[k]] / rho[pj[k]]) * coup[k];
    for (int64_t k = 0; k < P; ++k) rate[pj[k]] += scale * (mass[pi[k]] / rho[pi[k]]) * coup[k];
    free(coup); free(gr);
}

/* ------------------------------------------------------------------------ */
/* Morris-viscosity WCSPH right-hand side:                                    */
/*   2-D  droplet_impact.rhs  droplet_impact.py:810-1228  (kernel3d = 0)      */
/*   3-D  dam_break_3d.rhs_3d dam_break_3d.py:290-422     (kernel3d = 1)      */
/* ------------------------------------------------------------------------ */

typedef struct {
    int dim, kernel3d;
    double h, rho0, c0, gamma, gravity, nu, wall_friction, dt;
    double diff_scale; int diff_form; double eig_floor;
    int reaction_style;   /* 0: none, 1: 2-D split (face pressure + reaction), 2: rhs_3d */
    /* opt-in branches of droplet_impact.rhs (2-D) */
    double wrap_width;    /* > 0: x separations modulo wrap_width, :921-927 */
    int shepard;          /* density_renormalization, :954-964 */
    double lattice_sum;   /* full-lattice kernel sum for shepard */
    int bcd;              /* boundary_consistent_density, :1090-1112 */
    int closure;          /* surface_pressure_closure, :1031-1038 and :429-492 */
    double closure_threshold;
    int closure_literal;  /* donors "all" */
    int closure_layers;
    const double* density_offset;   /* n or NULL, :936-943 */
    const double* nu_field;         /* n or NULL, :1164-1178 */
} orc_morris_params;

static double tait2d(double rho, double rho0, double c0, double gamma) {
    double bulk = rho0 * (c0 * c0) / gamma;
    double ratio = rho / rho0; if (ratio < 0.01) ratio = 0.01;
    double p = bulk * (pow(ratio, gamma) - 1.0);
    return p < 0.0 ? 0.0 : p;
}

ORC_API void orc_tait2d(const double* rho, int64_t n, double rho0, double c0, double gamma, double* p) {
    for (int64_t k = 0; k < n; ++k) p[k] = tait2d(rho[k], rho0, c0, gamma);
}

/* pos_all/vel_all/mass_all hold n fluid rows then nw wall rows (wall rows of
 * vel_all are the wall velocity, zero for static walls).  wall_accel is
 * nw x dim or NULL.  Outputs: rho, p (n), acc (n x dim); wall_p, wall_rho
 * (nw) when walls exist; reaction (nw x dim) for reaction_style != 0. */
ORC_API void orc_rhs_morris(const orc_morris_params* P, int64_t n, int64_t nw,
                            const double* pos_all, const double* vel_all,
                            const double* mass_all, const double* wall_accel,
                            const int64_t* pairs, int64_t npairs,
                            double* rho_out, double* p_out, double* acc_out,
                            double* wall_p, double* wall_rho, double* reaction) {
    const int dim = P->dim;
    const double h = P->h;
    const int64_t nt = n + nw;
    double* rho_all = (double*)calloc((size_t)nt, sizeof(double));
    double w0, s0, l0;
    if (P->kernel3d) { cubic3d(0.0, h, &w0, &s0, &l0); for (int64_t a = 0; a < nt; ++a) rho_all[a] = mass_all[a] / (M_PI * h * h * h); }
    else { cubic2d(0.0, h, &w0, &s0); for (int64_t a = 0; a < n; ++a) rho_all[a] = mass_all[a] * w0; }
    double grav[3] = {0, 0, 0}; if (dim > 1) grav[1] = -P->gravity;
    if (wall_p) for (int64_t w = 0; w < nw; ++w) wall_p[w] = 0.0;
    if (wall_rho) for (int64_t w = 0; w < nw; ++w) wall_rho[w] = P->rho0;
    if (reaction) for (int64_t w = 0; w < nw * dim; ++w) reaction[w] = 0.0;
    if (npairs == 0) {
        for (int64_t a = 0; a < n; ++a) {
            double r = rho_all[a]; if (r < 0.1 * P->rho0) r = 0.1 * P->rho0;
            rho_out[a] = r; p_out[a] = tait2d(r, P->rho0, P->c0, P->gamma);
            for (int d = 0; d < dim; ++d) acc_out[a * dim + d] = grav[d];
        }
        free(rho_all); return;
    }
    double* delta = (double*)malloc((size_t)npairs * dim * sizeof(double));
    double* grad = (double*)malloc((size_t)npairs * dim * sizeof(double));
    double* wij = (double*)malloc((size_t)npairs * sizeof(double));
    double* dist = (double*)malloc((size_t)npairs * sizeof(double));
    for (int64_t k = 0; k < npairs; ++k) {
        int64_t i = pairs[2 * k], j = pairs[2 * k + 1];
        double r2 = 0;
        for (int d = 0; d < dim; ++d) {
            double df = pos_all[i * dim + d] - pos_all[j * dim + d];
            if (d == 0 && P->wrap_width > 0.0) df = df - P->wrap_width * nearbyint(df / P->wrap_width);
            delta[k * dim + d] = df; r2 += df * df;
        }
        double r = sqrt(r2), w, s;
        if (P->kernel3d) cubic3d(r, h, &w, &s, NULL); else cubic2d(r, h, &w, &s);
        dist[k] = r; wij[k] = w;
        for (int d = 0; d < dim; ++d) grad[k * dim + d] = delta[k * dim + d] * s;
    }
    for (int64_t k = 0; k < npairs; ++k) rho_all[pairs[2 * k]] += mass_all[pairs[2 * k + 1]] * wij[k];
    for (int64_t k = 0; k < npairs; ++k) rho_all[pairs[2 * k + 1]] += mass_all[pairs[2 * k]] * wij[k];
    double* rho = (double*)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
    for (int64_t a = 0; a < n; ++a) { double r = rho_all[a]; rho[a] = r < 0.1 * P->rho0 ? 0.1 * P->rho0 : r; }
    if (P->density_offset)
        for (int64_t a = 0; a < n; ++a) { double r = rho[a] + P->density_offset[a]; rho[a] = r < 0.1 * P->rho0 ? 0.1 * P->rho0 : r; }
    if (P->shepard) {
        /* Shepard factor: full-lattice kernel sum over this particle's sum_j W_ij + W(0) */
        double* kt = (double*)calloc((size_t)nt, sizeof(double));
        for (int64_t k = 0; k < npairs; ++k) kt[pairs[2 * k]] += wij[k];
        for (int64_t k = 0; k < npairs; ++k) kt[pairs[2 * k + 1]] += wij[k];
        for (int64_t a = 0; a < n; ++a) {
            double t = kt[a] + w0; if (t < 1.0e-12) t = 1.0e-12;
            double r = rho[a] * (P->lattice_sum / t); rho[a] = r < 0.1 * P->rho0 ? 0.1 * P->rho0 : r;
        }
        free(kt);
    }

    if (P->diff_scale > 0.0) {
        int64_t nf = 0;
        for (int64_t k = 0; k < npairs; ++k) if (pairs[2 * k] < n && pairs[2 * k + 1] < n) ++nf;
        int64_t* fi = (int64_t*)malloc((size_t)(nf > 0 ? nf : 1) * sizeof(int64_t));
        int64_t* fj = (int64_t*)malloc((size_t)(nf > 0 ? nf : 1) * sizeof(int64_t));
        double* fd = (double*)malloc((size_t)(nf > 0 ? nf : 1) * dim * sizeof(double));
        double* fg = (double*)malloc((size_t)(nf > 0 ? nf : 1) * dim * sizeof(double));
        int64_t c = 0;
        for (int64_t k = 0; k < npairs; ++k)
            if (pairs[2 * k] < n && pairs[2 * k + 1] < n) {
                fi[c] = pairs[2 * k]; fj[c] = pairs[2 * k + 1];
                for (int d = 0; d < dim; ++d) { fd[c * dim + d] = delta[k * dim + d]; fg[c * dim + d] = grad[k * dim + d]; }
                ++c;
            }
        double* rate = (double*)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
        diffusion_rate(dim, n, rho, mass_all, nf, fi, fj, fd, fg, P->diff_scale, P->diff_form, P->eig_floor, rate);
        for (int64_t a = 0; a < n; ++a) { double r = rho[a] + P->dt * rate[a]; rho[a] = r < 0.1 * P->rho0 ? 0.1 * P->rho0 : r; }
        free(fi); free(fj); free(fd); free(fg); free(rate);
    }
    double* p_all = (double*)calloc((size_t)nt, sizeof(double));
    double* d_all = (double*)calloc((size_t)nt, sizeof(double));
    for (int64_t a = 0; a < n; ++a) { p_all[a] = tait2d(rho[a], P->rho0, P->c0, P->gamma); d_all[a] = rho[a]; }

    if (P->closure && n > 0) {
        /* free_surface_pressure over fluid-fluid pairs: receivers are rows
         * below threshold * rho0; each pass fills receivers from donors known
         * at the pass start ("interior") or from every neighbour ("all") */
        const double thr = P->closure_threshold * P->rho0;
        int* known = (int*)malloc((size_t)n * sizeof(int));
        int* recv = (int*)malloc((size_t)n * sizeof(int));
        double* num = (double*)malloc((size_t)n * sizeof(double));
        double* den = (double*)malloc((size_t)n * sizeof(double));
        int any = 0;
        for (int64_t a = 0; a < n; ++a) { known[a] = !(rho[a] < thr); any |= !known[a]; }
        const int layers = P->closure_literal ? 1 : (P->closure_layers > 1 ? P->closure_layers : 1);
        for (int pass = 0; any && pass < layers; ++pass) {
            int left = 0;
            for (int64_t a = 0; a < n; ++a) { recv[a] = !known[a]; left |= recv[a]; num[a] = 0.0; den[a] = 0.0; }
            if (!left) break;
            for (int64_t k = 0; k < npairs; ++k) {
                int64_t i = pairs[2 * k], j = pairs[2 * k + 1];
                if (i >= n || j >= n) continue;
                double wi = P->closure_literal ? wij[k] : wij[k] * (double)known[j];
                double wj = P->closure_literal ? wij[k] : wij[k] * (double)known[i];
                double grav = P->gravity * (dim > 1 ? delta[k * dim + 1] : 0.0);
                num[i] += wi * (p_all[j] - grav * rho[j]); den[i] += wi;
                num[j] += wj * (p_all[i] + grav * rho[i]); den[j] += wj;
            }
            for (int64_t a = 0; a < n; ++a)
                if (recv[a] && den[a] > 1.0e-30) {
                    double v = num[a] / den[a]; p_all[a] = v < 0.0 ? 0.0 : v; known[a] = 1;
                }
        }
        free(known); free(recv); free(num); free(den);
    }

    if (nw) {
        double bulk = P->rho0 * (P->c0 * P->c0) / P->gamma;
        double* num = (double*)calloc((size_t)nw, sizeof(double));
        double* den = (double*)calloc((size_t)nw, sizeof(double));
        for (int bcd_pass = 0; bcd_pass < (P->bcd ? 2 : 1); ++bcd_pass) {
        if (bcd_pass == 1) {
            /* boundary-consistent density: wall masses scaled by rho_w / rho0
             * enter the fluid summation density, then Tait and a second
             * extrapolation from the corrected fluid state */
            double* corr = (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
            for (int64_t k = 0; k < npairs; ++k) {
                int64_t i = pairs[2 * k], j = pairs[2 * k + 1];
                int ifl = i < n, jfl = j < n;
                if (ifl == jfl) continue;
                int64_t f = ifl ? i : j, w = ifl ? j : i;
                double mw = mass_all[w];
                corr[f] += (mw * (d_all[w] / P->rho0) - mw) * wij[k];
            }
            for (int64_t a = 0; a < n; ++a) {
                double r = rho[a] + corr[a]; rho[a] = r < 0.1 * P->rho0 ? 0.1 * P->rho0 : r;
                p_all[a] = tait2d(rho[a], P->rho0, P->c0, P->gamma); d_all[a] = rho[a];
            }
            free(corr);
            for (int64_t w = 0; w < nw; ++w) { num[w] = 0.0; den[w] = 0.0; }
        }
        for (int64_t k = 0; k < npairs; ++k) {
            int64_t i = pairs[2 * k], j = pairs[2 * k + 1];
            int ifl = i < n, jfl = j < n;
            double weight = (ifl != jfl) ? wij[k] : 0.0;
            int64_t f = ifl ? i : j;
            int64_t w = (ifl ? j : i) - n; if (w < 0) w = 0; if (w > nw - 1) w = nw - 1;
            double dy = ifl ? -delta[k * dim + 1] : delta[k * dim + 1];
            double body = -P->gravity * dy;
            if (wall_accel) {
                double dxw = ifl ? -delta[k * dim + 0] : delta[k * dim + 0];
                body = body - (wall_accel[w * dim + 0] * dxw + wall_accel[w * dim + 1] * dy);
            }
            double src = f < n ? p_all[f] : 0.0;
            if (P->gravity != 0.0) src = src + (f < n ? rho[f] : 0.0) * body;
            num[w] += src * weight;
        }
        for (int64_t k = 0; k < npairs; ++k) {
            int64_t i = pairs[2 * k], j = pairs[2 * k + 1];
            int ifl = i < n, jfl = j < n;
            double weight = (ifl != jfl) ? wij[k] : 0.0;
            int64_t w = (ifl ? j : i) - n; if (w < 0) w = 0; if (w > nw - 1) w = nw - 1;
            den[w] += weight;
        }
        for (int64_t w = 0; w < nw; ++w) {
            double dd = den[w] > 1.0e-30 ? den[w] : 1.0e-30;
            double pw = num[w] / dd; if (pw < 0.0) pw = 0.0;
            double rw = P->rho0 * pow(pw / bulk + 1.0, 1.0 / P->gamma); if (rw < P->rho0) rw = P->rho0;
            p_all[n + w] = pw; d_all[n + w] = rw;
            if (wall_p) wall_p[w] = pw;
            if (wall_rho) wall_rho[w] = rw;
        }
        }
        free(num); free(den);
    }

    double* accp = (double*)calloc((size_t)nt * dim, sizeof(double));
    double* accv = (double*)calloc((size_t)nt * dim, sizeof(double));
    double* pterm = (double*)malloc((size_t)npairs * sizeof(double));
    for (int64_t k = 0; k < npairs; ++k) {
        int64_t i = pairs[2 * k], j = pairs[2 * k + 1];
        pterm[k] = p_all[i] / (d_all[i] * d_all[i]) + p_all[j] / (d_all[j] * d_all[j]);
    }
    for (int64_t k = 0; k < npairs; ++k) {
        int64_t i = pairs[2 * k], j = pairs[2 * k + 1];
        for (int d = 0; d < dim; ++d) accp[i * dim + d] += -mass_all[j] * (pterm[k] * grad[k * dim + d]);
    }
    for (int64_t k = 0; k < npairs; ++k) {
        int64_t i = pairs[2 * k], j = pairs[2 * k + 1];
        for (int d = 0; d < dim; ++d) accp[j * dim + d] += mass_all[i] * (pterm[k] * grad[k * dim + d]);
    }
    /* viscosity scale per pair */
    double* vi = (double*)malloc((size_t)npairs * dim * sizeof(double));
    double* vj = (double*)malloc((size_t)npairs * dim * sizeof(double));
    for (int64_t k = 0; k < npairs; ++k) {
        int64_t i = pairs[2 * k], j = pairs[2 * k + 1];
        int fluid_pair = (i < n) && (j < n);
        int crossing = (i < n) != (j < n);
        double nu = P->nu;
        if (P->nu_field) {
            double ni = i < n ? P->nu_field[i] : 0.0, nj = j < n ? P->nu_field[j] : 0.0;
            nu = fluid_pair ? 0.5 * (ni + nj) : (i < n ? ni : nj);
        }
        double scale;
        if (nw && P->wall_friction > 0.0) scale = 2.0 * nu * ((double)fluid_pair + P->wall_friction * (double)crossing);
        else if (nw) scale = 2.0 * nu * (double)fluid_pair;
        else scale = 2.0 * nu;
        double dg = 0; for (int d = 0; d < dim; ++d) dg += delta[k * dim + d] * grad[k * dim + d];
        double relu = -dg > 0.0 ? -dg : 0.0;
        double lw = relu / (dist[k] * dist[k] + 0.01 * h * h);
        for (int d = 0; d < dim; ++d) {
            double dv = vel_all[i * dim + d] - vel_all[j * dim + d];
            vi[k * dim + d] = -scale * dv * (mass_all[j] / d_all[j] * lw);
            vj[k * dim + d] = scale * dv * (mass_all[i] / d_all[i] * lw);
        }
    }
    for (int64_t k = 0; k < npairs; ++k) for (int d = 0; d < dim; ++d) accv[pairs[2 * k] * dim + d] += vi[k * dim + d];
    for (int64_t k = 0; k < npairs; ++k) for (int d = 0; d < dim; ++d) accv[pairs[2 * k + 1] * dim + d] += vj[k * dim + d];
    for (int64_t a = 0; a < n; ++a) {
        rho_out[a] = rho[a]; p_out[a] = p_all[a];
        for (int d = 0; d < dim; ++d) acc_out[a * dim + d] = accp[a * dim + d] + accv[a * dim + d] + grav[d];
    }
    if (reaction && nw) {
        for (int64_t k = 0; k < npairs; ++k) {
            int64_t i = pairs[2 * k], j = pairs[2 * k + 1];
            int ifl = i < n, jfl = j < n;
            if (ifl == jfl) continue;
            int64_t w = (ifl ? j : i) - n;
            int64_t f = ifl ? i : j;
            double pm = mass_all[i] * mass_all[j];
            for (int d = 0; d < dim; ++d) {
                double pp = pterm[k] * grad[k * dim + d];
                double visc_f = ifl ? vi[k * dim + d] : vj[k * dim + d];
                if (P->reaction_style == 1) {
                    double sign = ifl ? -1.0 : 1.0;
                    reaction[w * dim + d] += -sign * pm * pp;
                } else {
                    double sign = ifl ? 1.0 : -1.0;
                    reaction[w * dim + d] += sign * pm * pp - mass_all[f] * visc_f;
                }
            }
        }
        if (P->reaction_style == 1 && P->wall_friction > 0.0)
            for (int64_t k = 0; k < npairs; ++k) {
                int64_t i = pairs[2 * k], j = pairs[2 * k + 1];
                int ifl = i < n, jfl = j < n;
                if (ifl == jfl) continue;
                int64_t w = (ifl ? j : i) - n, f = ifl ? i : j;
                for (int d = 0; d < dim; ++d) {
                    double visc_f = ifl ? vi[k * dim + d] : vj[k * dim + d];
                    reaction[w * dim + d] += -visc_f * mass_all[f];
                }
            }
    }
    free(rho_all); free(delta); free(grad); free(wij); free(dist); free(rho);
    free(p_all); free(d_all); free(accp); free(accv); free(pterm); free(vi); free(vj);
}

/* ------------------------------------------------------------------------ */
/* threed.operators3.accelerations, threed/operators3.py:212-256, with       */
/* summation_density :70-79, tait :82-93, wall_pressure_adami :173-209,      */
/* pressure_acceleration :96-113, viscous_acceleration :116-154 (order 2,   */
/* and order 4: the iterated operator L - L(beta L) of :146-151 with        */
/* beta_ij = (I4_i + I4_j) / 240 and I4 the fourth moment of :147-149).     */
/* ------------------------------------------------------------------------ */

ORC_API void orc_threed_accelerations(int64_t n, int64_t n_fluid, const double* pos,
        const double* vel, const double* mass, double h, const int64_t* pairs, int64_t npairs,
        double rho0, double c0, double gamma, double nu, const double* gravity,
        const double* box, int renormalise, double diff_delta, double dt, int diff_form,
        double eig_floor, int viscous_order, double* acc, double* rho, double* pressure) {
    const int dim = 3;
    double* delta = (double*)malloc((size_t)(npairs > 0 ? npairs : 1) * 3 * sizeof(double));
    double* grad = (double*)malloc((size_t)(npairs > 0 ? npairs : 1) * 3 * sizeof(double));
    double* wij = (double*)malloc((size_t)(npairs > 0 ? npairs : 1) * sizeof(double));
    double* lap = (double*)malloc((size_t)(npairs > 0 ? npairs : 1) * sizeof(double));
    for (int64_t k = 0; k < npairs; ++k) {
        int64_t i = pairs[2 * k], j = pairs[2 * k + 1];
        double r2 = 0;
        for (int d = 0; d < 3; ++d) {
            double df = pos[i * 3 + d] - pos[j * 3 + d];
            if (box) df = min_image(df, box[d]);
            delta[k * 3 + d] = df; r2 += df * df;
        }
        double r = sqrt(r2), w, s, l;
        cubic3d(r, h, &w, &s, &l);
        wij[k] = w; lap[k] = l;
        for (int d = 0; d < 3; ++d) grad[k * 3 + d] = delta[k * 3 + d] * s;
    }
    double w0, s0;
    cubic3d(0.0, h, &w0, &s0, NULL);
    double* rs = (double*)calloc((size_t)n, sizeof(double));
    for (int64_t k = 0; k < npairs; ++k) rs[pairs[2 * k]] += mass[pairs[2 * k + 1]] * wij[k];
    for (int64_t k = 0; k < npairs; ++k) rs[pairs[2 * k + 1]] += mass[pairs[2 * k]] * wij[k];
    for (int64_t a = 0; a < n; ++a) rho[a] = rs[a] + mass[a] * w0;
    free(rs);
    if (diff_delta > 0.0) {
        int64_t nf = 0;
        for (int64_t k = 0; k < npairs; ++k) if (pairs[2 * k] < n_fluid && pairs[2 * k + 1] < n_fluid) ++nf;
        int64_t* fi = (int64_t*)malloc((size_t)(nf > 0 ? nf : 1) * sizeof(int64_t));
        int64_t* fj = (int64_t*)malloc((size_t)(nf > 0 ? nf : 1) * sizeof(int64_t));
        double* fd = (double*)malloc((size_t)(nf > 0 ? nf : 1) * 3 * sizeof(double));
        double* fg = (double*)malloc((size_t)(nf > 0 ? nf : 1) * 3 * sizeof(double));
        int64_t c = 0;
        for (int64_t k = 0; k < npairs; ++k)
            if (pairs[2 * k] < n_fluid && pairs[2 * k + 1] < n_fluid) {
                fi[c] = pairs[2 * k]; fj[c] = pairs[2 * k + 1];
                for (int d = 0; d < 3; ++d) { fd[c * 3 + d] = delta[k * 3 + d]; fg[c * 3 + d] = grad[k * 3 + d]; }
                ++c;
            }
        double* rate = (double*)malloc((size_t)n * sizeof(double));
        diffusion_rate(3, n, rho, mass, nf, fi, fj, fd, fg, 2.0 * diff_delta * c0 * h, diff_form, eig_floor, rate);
        for (int64_t a = 0; a < n; ++a) { double r = rho[a] + dt * rate[a]; rho[a] = r < 0.1 * rho0 ? 0.1 * rho0 : r; }
        free(fi); free(fj); free(fd); free(fg); free(rate);
    }
    double stiff = rho0 * (c0 * c0) / gamma;
    double* prs = (double*)malloc((size_t)n * sizeof(double));
    double* rho_used = (double*)malloc((size_t)n * sizeof(double));
    for (int64_t a = 0; a < n; ++a) {
        double p = stiff * (pow(rho[a] / rho0, gamma) - 1.0);
        prs[a] = p < 0.0 ? 0.0 : p; rho_used[a] = rho[a];
    }
    if (n_fluid < n) {
        double* num = (double*)calloc((size_t)n, sizeof(double));
        double* den = (double*)calloc((size_t)n, sizeof(double));
        /* fluid-first orientation: the reference concatenates (i fluid, j wall)
         * pairs then (j fluid, i wall) pairs; the sums follow that order. */
        for (int pass = 0; pass < 2; ++pass)
            for (int64_t k = 0; k < npairs; ++k) {
                int64_t i = pairs[2 * k], j = pairs[2 * k + 1];
                int64_t f, w;
                if (pass == 0) { if (!(i < n_fluid && j >= n_fluid)) continue; f = i; w = j; }
                else { if (!(j < n_fluid && i >= n_fluid)) continue; f = j; w = i; }
                double r2 = 0, gd = 0;
                for (int d = 0; d < 3; ++d) {
                    double df = pos[f * 3 + d] - pos[w * 3 + d];
                    r2 += df * df;
                    if (gravity) gd += (pos[w * 3 + d] - pos[f * 3 + d]) * gravity[d];
                }
                double wt, s;
                cubic3d(sqrt(r2), h, &wt, &s, NULL);
                double donor = gravity ? prs[f] + rho[f] * gd : prs[f];
                num[w] += donor * wt; den[w] += wt;
            }
        for (int64_t w = n_fluid; w < n; ++w) {
            double dd = den[w] > 1.0e-30 ? den[w] : 1.0e-30;
            double pw = den[w] > 0.0 ? num[w] / dd : 0.0;
            if (pw < 0.0) pw = 0.0;
            prs[w] = pw;
            double pc = pw < 0.0 ? 0.0 : pw;
            rho_used[w] = rho0 * pow(1.0 + pc / stiff, 1.0 / gamma);
        }
        free(num); free(den);
    }
    for (int64_t a = 0; a < n * 3; ++a) acc[a] = 0.0;
    for (int64_t k = 0; k < npairs; ++k) {
        int64_t i = pairs[2 * k], j = pairs[2 * k + 1];
        double sym = prs[i] / (rho_used[i] * rho_used[i]) + prs[j] / (rho_used[j] * rho_used[j]);
        for (int d = 0; d < 3; ++d) acc[i * 3 + d] += -(grad[k * 3 + d] * (mass[j] * sym));
    }
    for (int64_t k = 0; k < npairs; ++k) {
        int64_t i = pairs[2 * k], j = pairs[2 * k + 1];
        double sym = prs[i] / (rho_used[i] * rho_used[i]) + prs[j] / (rho_used[j] * rho_used[j]);
        for (int d = 0; d < 3; ++d) acc[j * 3 + d] += grad[k * 3 + d] * (mass[i] * sym);
    }
    if (nu != 0.0) {
        double* moment = (double*)calloc((size_t)n, sizeof(double));
        for (int64_t k = 0; k < npairs; ++k) {
            int64_t i = pairs[2 * k], j = pairs[2 * k + 1];
            double rij = 0.5 * (rho_used[i] + rho_used[j]);
            double r2 = 0; for (int d = 0; d < 3; ++d) r2 += delta[k * 3 + d] * delta[k * 3 + d];
            moment[i] += r2 * lap[k] * (mass[j] / rij);
        }
        for (int64_t k = 0; k < npairs; ++k) {
            int64_t i = pairs[2 * k], j = pairs[2 * k + 1];
            double rij = 0.5 * (rho_used[i] + rho_used[j]);
            double r2 = 0; for (int d = 0; d < 3; ++d) r2 += delta[k * 3 + d] * delta[k * 3 + d];
            moment[j] += r2 * lap[k] * (mass[i] / rij);
        }
        for (int64_t a = 0; a < n; ++a) moment[a] = moment[a] / 3.0;
        double* out = (double*)calloc((size_t)n * 3, sizeof(double));
        double* dvs = (double*)malloc((size_t)(npairs > 0 ? npairs : 1) * 3 * sizeof(double));
        for (int64_t k = 0; k < npairs; ++k) {
            int64_t i = pairs[2 * k], j = pairs[2 * k + 1];
            double rij = 0.5 * (rho_used[i] + rho_used[j]);
            double weight = mass[j] * lap[k] / rij;
            if (renormalise) {
                double mk = 0.5 * (moment[i] + moment[j]); if (mk < 1.0e-12) mk = 1.0e-12;
                weight = weight * (2.0 / mk);
            }
            for (int d = 0; d < 3; ++d) dvs[k * 3 + d] = (vel[j * 3 + d] - vel[i * 3 + d]) * weight;
        }
        for (int64_t k = 0; k < npairs; ++k) for (int d = 0; d < 3; ++d) out[pairs[2 * k] * 3 + d] += dvs[k * 3 + d];
        for (int64_t k = 0; k < npairs; ++k) for (int d = 0; d < 3; ++d) out[pairs[2 * k + 1] * 3 + d] += -dvs[k * 3 + d];
        if (viscous_order == 4) {
            double* fourth = (double*)calloc((size_t)n, sizeof(double));
            for (int64_t k = 0; k < npairs; ++k) {
                int64_t i = pairs[2 * k], j = pairs[2 * k + 1];
                double rij = 0.5 * (rho_used[i] + rho_used[j]);
                double r2 = 0; for (int d = 0; d < 3; ++d) r2 += delta[k * 3 + d] * delta[k * 3 + d];
                fourth[i] += r2 * r2 * lap[k] * (mass[j] / rij);
            }
            for (int64_t k = 0; k < npairs; ++k) {
                int64_t i = pairs[2 * k], j = pairs[2 * k + 1];
                double rij = 0.5 * (rho_used[i] + rho_used[j]);
                double r2 = 0; for (int d = 0; d < 3; ++d) r2 += delta[k * 3 + d] * delta[k * 3 + d];
                fourth[j] += r2 * r2 * lap[k] * (mass[i] / rij);
            }
            /* second pass of the same pair Laplacian on the first one's output,
             * with the pair weight scaled by beta_ij */
            double* out2 = (double*)calloc((size_t)n * 3, sizeof(double));
            for (int64_t k = 0; k < npairs; ++k) {
                int64_t i = pairs[2 * k], j = pairs[2 * k + 1];
                double rij = 0.5 * (rho_used[i] + rho_used[j]);
                double weight = mass[j] * lap[k] / rij;
                if (renormalise) {
                    double mk = 0.5 * (moment[i] + moment[j]); if (mk < 1.0e-12) mk = 1.0e-12;
                    weight = weight * (2.0 / mk);
                }
                double beta = 0.5 * (fourth[i] + fourth[j]) / 120.0;
                for (int d = 0; d < 3; ++d) dvs[k * 3 + d] = (out[j * 3 + d] - out[i * 3 + d]) * (weight * beta);
            }
            for (int64_t k = 0; k < npairs; ++k) for (int d = 0; d < 3; ++d) out2[pairs[2 * k] * 3 + d] += dvs[k * 3 + d];
            for (int64_t k = 0; k < npairs; ++k) for (int d = 0; d < 3; ++d) out2[pairs[2 * k + 1] * 3 + d] += -dvs[k * 3 + d];
            for (int64_t a = 0; a < n * 3; ++a) out[a] = out[a] - out2[a];
            free(fourth); free(out2);
        }
        for (int64_t a = 0; a < n * 3; ++a) acc[a] = acc[a] + out[a] * nu;
        free(moment); free(out); free(dvs);
    }
    if (gravity)
        for (int64_t a = 0; a < n_fluid; ++a)
            for (int d = 0; d < 3; ++d) acc[a * 3 + d] = acc[a * 3 + d] + gravity[d];
    for (int64_t a = 0; a < n; ++a) pressure[a] = prs[a];
    free(delta); free(grad); free(wij); free(lap); free(prs); free(rho_used);
}

/* ------------------------------------------------------------------------ */
/* Boundary safety clamp, droplet_impact.py:1293-1312 (2-D) and             */
/* dam_break_3d.py:429-452 (3-D).  lo/hi: box, tol = 0.25 dx.  Returns the   */
/* number of particles found beyond the tolerance.                           */
/* ------------------------------------------------------------------------ */
ORC_API int64_t orc_apply_boundary(int dim, int64_t n, double* pos, double* vel,
                                   const double* lo, const double* hi, double tol) {
    int64_t count = 0;
    for (int64_t a = 0; a < n; ++a) {
        int hit_any = 0;
        for (int d = 0; d < dim; ++d) {
            double x = pos[a * dim + d];
            int hit = (x < lo[d] - tol) || (x > hi[d] + tol);
            hit_any |= hit;
            pos[a * dim + d] = x < lo[d] ? lo[d] : (x > hi[d] ? hi[d] : x);
            if (hit) vel[a * dim + d] = 0.0;
        }
        count += hit_any;
    }
    return count;
}

/* ------------------------------------------------------------------------ */
/* delta-plus SPH step, delta_plus.py:105-229 (Sun et al. 2017), FP64.       */
/* pos/vel (N x 2), rho/mass (N); wall/wall_vel/wall_normal (W x 2, normals  */
/* already unit length); pairs over cat(pos, wall), i < j.                   */
/* Returns 0, or 1 for a non-positive wall density, 2 for a non-positive or  */
/* non-finite new density (the reference raises FloatingPointError).        */
/* ------------------------------------------------------------------------ */

typedef struct {
    double dx, c0, ref_speed, dt, rho0, gravity, nu, diffusion, alpha;
} orc_dp_params;

static int occupied(double rx, double ry, double nx, double ny, double h) {
    double axial = rx * nx + ry * ny;
    double tangent = fabs(rx * ny - ry * nx);
    double ox = rx - h * nx, oy = ry - h * ny;
    if (rx * rx + ry * ry >= 2.0 * h * h) return (ox * ox + oy * oy) < h * h;
    return (fabs(axial - h) + tangent) < h;
}

ORC_API int orc_delta_plus_step(const orc_dp_params* P, int64_t N, int64_t W,
        const double* pos, const double* vel, const double* rho, const double* mass,
        const double* wall, const double* wall_vel, const double* wall_n,
        const int64_t* pairs, int64_t npairs,
        double* pos_out, double* vel_out, double* rho_out, double* acc_out,
        double* reaction, double* shift_out) {
    const double dx = P->dx, h = 2.0 * dx, c0 = P->c0;
    const double rv = 8.0 * P->nu + P->alpha * h * c0;
    int64_t nk = 0;
    int64_t* ki = (int64_t*)malloc((size_t)(npairs > 0 ? npairs : 1) * sizeof(int64_t));
    int64_t* kj = (int64_t*)malloc((size_t)(npairs > 0 ? npairs : 1) * sizeof(int64_t));
    for (int64_t k = 0; k < npairs; ++k) {
        int64_t i = pairs[2 * k], j = pairs[2 * k + 1];
        if (!(i < N)) continue;
        const double* xj = j < N ? &pos[2 * j] : &wall[2 * (j - N)];
        double dxx = pos[2 * i] - xj[0], dyy = pos[2 * i + 1] - xj[1];
        if (dxx * dxx + dyy * dyy < (2.0 * h) * (2.0 * h)) { ki[nk] = i; kj[nk] = j; ++nk; }
    }
    double* sep = (double*)malloc((size_t)(nk > 0 ? nk : 1) * 2 * sizeof(double));
    double* grd = (double*)malloc((size_t)(nk > 0 ? nk : 1) * 2 * sizeof(double));
    double* wgt = (double*)malloc((size_t)(nk > 0 ? nk : 1) * sizeof(double));
    for (int64_t k = 0; k < nk; ++k) {
        int64_t i = ki[k], j = kj[k];
        const double* xj = j < N ? &pos[2 * j] : &wall[2 * (j - N)];
        sep[2 * k] = pos[2 * i] - xj[0]; sep[2 * k + 1] = pos[2 * i + 1] - xj[1];
        double r = sqrt(sep[2 * k] * sep[2 * k] + sep[2 * k + 1] * sep[2 * k + 1]);
        double w, s; wendland2d(r, h, &w, &s);
        wgt[k] = w; grd[2 * k] = sep[2 * k] * s; grd[2 * k + 1] = sep[2 * k + 1] * s;
    }
    double* vol = (double*)malloc((size_t)N * sizeof(double));
    double* p = (double*)malloc((size_t)N * sizeof(double));
    for (int64_t a = 0; a < N; ++a) { vol[a] = mass[a] / rho[a]; p[a] = (c0 * c0) * (rho[a] - P->rho0); }
    double* psum = (double*)calloc((size_t)(W > 0 ? W : 1), sizeof(double));
    double* wsum = (double*)calloc((size_t)(W > 0 ? W : 1), sizeof(double));
    for (int64_t k = 0; k < nk; ++k) {
        if (kj[k] < N) continue;
        int64_t f = ki[k], w = kj[k] - N;
        psum[w] += wgt[k] * (p[f] + rho[f] * P->gravity * sep[2 * k + 1]);
    }
    for (int64_t k = 0; k < nk; ++k) if (kj[k] >= N) wsum[kj[k] - N] += wgt[k];
    double* wall_p = (double*)malloc((size_t)(W > 0 ? W : 1) * sizeof(double));
    double* wall_rho = (double*)malloc((size_t)(W > 0 ? W : 1) * sizeof(double));
    for (int64_t w = 0; w < W; ++w) {
        double den = wsum[w] > 2.2250738585072014e-308 ? wsum[w] : 2.2250738585072014e-308;
        wall_p[w] = psum[w] / den;
        wall_rho[w] = P->rho0 + wall_p[w] / (c0 * c0);
    }
    int status = 0;
    for (int64_t k = 0; k < nk; ++k) if (kj[k] >= N && wall_rho[kj[k] - N] <= 0.0) status = 1;
    double* net = (double*)calloc((size_t)N * 2, sizeof(double));
    double* fz = (double*)malloc((size_t)(nk > 0 ? nk : 1) * 2 * sizeof(double));
    double* wf = (double*)malloc((size_t)(nk > 0 ? nk : 1) * 2 * sizeof(double));
    for (int64_t k = 0; k < nk; ++k) {
        int64_t a = ki[k], b = kj[k];
        if (b < N) {
            double c = -(vol[a] * vol[b] * (p[a] + p[b]));
            double fx = c * grd[2 * k], fy = c * grd[2 * k + 1];
            if (rv != 0.0) {
                double approach = (vel[2 * a] - vel[2 * b]) * sep[2 * k] + (vel[2 * a + 1] - vel[2 * b + 1]) * sep[2 * k + 1];
                double d2 = sep[2 * k] * sep[2 * k] + sep[2 * k + 1] * sep[2 * k + 1];
                double fl = 1.0e-12 * h * h;
                double pi = approach / (d2 > fl ? d2 : fl);
                double cv = P->rho0 * rv * vol[a] * vol[b] * pi;
                fx += cv * grd[2 * k]; fy += cv * grd[2 * k + 1];
            }
            fz[2 * k] = fx; fz[2 * k + 1] = fy;
        } else {
            int64_t f = a, w = b - N;
            double nx = wall_n[2 * w], ny = wall_n[2 * w + 1];
            double gn = grd[2 * k] * nx + grd[2 * k + 1] * ny;
            double sg = gn > 0 ? 1.0 : (gn < 0 ? -1.0 : 0.0);
            double tx = nx * sg, ty = ny * sg;
            double closing = (vel[2 * f] - wall_vel[2 * w]) * tx + (vel[2 * f + 1] - wall_vel[2 * w + 1]) * ty;
            double gt = grd[2 * k] * tx + grd[2 * k + 1] * ty;
            double ngx = gt * tx, ngy = gt * ty;
            double twp = p[f] + wall_p[w] + 2.0 * rho[f] * c0 * closing;
            double c = -(vol[f] * (dx * dx) * twp);
            wf[2 * k] = c * ngx; wf[2 * k + 1] = c * ngy;
        }
    }
    /* _both(pos, a, b, force, -force) then index_add_(f, wall_force) */
    for (int64_t k = 0; k < nk; ++k) if (kj[k] < N) { net[2 * ki[k]] += fz[2 * k]; net[2 * ki[k] + 1] += fz[2 * k + 1]; }
    for (int64_t k = 0; k < nk; ++k) if (kj[k] < N) { net[2 * kj[k]] += -fz[2 * k]; net[2 * kj[k] + 1] += -fz[2 * k + 1]; }
    for (int64_t k = 0; k < nk; ++k) if (kj[k] >= N) { net[2 * ki[k]] += wf[2 * k]; net[2 * ki[k] + 1] += wf[2 * k + 1]; }
    for (int64_t w = 0; w < W * 2; ++w) reaction[w] = 0.0;
    for (int64_t k = 0; k < nk; ++k) if (kj[k] >= N) { int64_t w = kj[k] - N; reaction[2 * w] += -wf[2 * k]; reaction[2 * w + 1] += -wf[2 * k + 1]; }
    double* vnew = (double*)malloc((size_t)N * 2 * sizeof(double));
    for (int64_t a = 0; a < N; ++a) {
        acc_out[2 * a] = net[2 * a] / mass[a] + 0.0;
        acc_out[2 * a + 1] = net[2 * a + 1] / mass[a] + (-P->gravity);
        vnew[2 * a] = vel[2 * a] + P->dt * acc_out[2 * a];
        vnew[2 * a + 1] = vel[2 * a + 1] + P->dt * acc_out[2 * a + 1];
    }
    /* first moment and its regularised inverse (fluid pairs only) */
    double* mom = (double*)calloc((size_t)N * 4, sizeof(double));
    for (int pass = 0; pass < 2; ++pass)
        for (int64_t k = 0; k < nk; ++k) {
            if (kj[k] >= N) continue;
            int64_t a = ki[k], b = kj[k];
            int64_t t = pass == 0 ? a : b;
            double vv = pass == 0 ? vol[b] : vol[a];
            for (int r = 0; r < 2; ++r) for (int c = 0; c < 2; ++c)
                mom[t * 4 + r * 2 + c] += vv * (-sep[2 * k + r] * grd[2 * k + c]);
        }
    double* inv = (double*)malloc((size_t)N * 4 * sizeof(double));
    double* lam = (double*)malloc((size_t)N * sizeof(double));
    for (int64_t a = 0; a < N; ++a) {
        double w[2], v[4];
        sym_eig(2, &mom[a * 4], w, v);
        double fl = 0.1 * (0.5 * (w[0] + w[1])); if (fl < 1.0e-12) fl = 1.0e-12;
        double l0 = w[0] > fl ? w[0] : fl, l1 = w[1] > fl ? w[1] : fl;
        for (int r = 0; r < 2; ++r) for (int c = 0; c < 2; ++c)
            inv[a * 4 + r * 2 + c] = v[r * 2 + 0] / l0 * v[c * 2 + 0] + v[r * 2 + 1] / l1 * v[c * 2 + 1];
        lam[a] = w[0];
    }
    /* gradient(values) = inverse @ sum_b V (values_b - values_a) grad */
#define DP_GRADIENT(VALS, OUT) do {                                                  \
        double* sm_ = (double*)calloc((size_t)N * 2, sizeof(double));                  \
        for (int pass_ = 0; pass_ < 2; ++pass_)                                         \
            for (int64_t k_ = 0; k_ < nk; ++k_) {                                       \
                if (kj[k_] >= N) continue;                                              \
                int64_t a_ = ki[k_], b_ = kj[k_];                                       \
                double df_ = (VALS)[b_] - (VALS)[a_];                                   \
                int64_t t_ = pass_ == 0 ? a_ : b_;                                      \
                double vv_ = pass_ == 0 ? vol[b_] : vol[a_];                            \
                sm_[2 * t_] += vv_ * (df_ * grd[2 * k_]);                               \
                sm_[2 * t_ + 1] += vv_ * (df_ * grd[2 * k_ + 1]);                       \
            }                                                                           \
        for (int64_t a_ = 0; a_ < N; ++a_) {                                            \
            (OUT)[2 * a_] = inv[a_ * 4] * sm_[2 * a_] + inv[a_ * 4 + 1] * sm_[2 * a_ + 1]; \
            (OUT)[2 * a_ + 1] = inv[a_ * 4 + 2] * sm_[2 * a_] + inv[a_ * 4 + 3] * sm_[2 * a_ + 1]; \
        }                                                                               \
        free(sm_);                                                                      \
    } while (0)
    double* div = (double*)calloc((size_t)N, sizeof(double));
    for (int pass = 0; pass < 2; ++pass)
        for (int64_t k = 0; k < nk; ++k) {
            if (kj[k] >= N) continue;
            int64_t a = ki[k], b = kj[k];
            double dp = (vnew[2 * b] - vnew[2 * a]) * grd[2 * k] + (vnew[2 * b + 1] - vnew[2 * a + 1]) * grd[2 * k + 1];
            if (pass == 0) div[a] += vol[b] * dp; else div[b] += vol[a] * dp;
        }
    for (int64_t k = 0; k < nk; ++k) {
        if (kj[k] < N) continue;
        int64_t f = ki[k], w = kj[k] - N;
        double nx = wall_n[2 * w], ny = wall_n[2 * w + 1];
        double ns = (wall_vel[2 * w] - vnew[2 * f]) * nx + (wall_vel[2 * w + 1] - vnew[2 * f + 1]) * ny;
        div[f] += 2.0 * (dx * dx) * ns * (nx * grd[2 * k] + ny * grd[2 * k + 1]);
    }
    double* rate = (double*)malloc((size_t)N * sizeof(double));
    for (int64_t a = 0; a < N; ++a) rate[a] = -rho[a] * div[a];
    if (P->diffusion != 0.0) {
        double* gr = (double*)malloc((size_t)N * 2 * sizeof(double));
        DP_GRADIENT(rho, gr);
        double scale = 2.0 * P->diffusion * h * c0;
        double* cp = (double*)malloc((size_t)(nk > 0 ? nk : 1) * sizeof(double));
        double* hr = (double*)calloc((size_t)N, sizeof(double));
        for (int64_t k = 0; k < nk; ++k) {
            if (kj[k] >= N) continue;
            int64_t a = ki[k], b = kj[k];
            double proj = 0.5 * ((gr[2 * a] + gr[2 * b]) * sep[2 * k] + (gr[2 * a + 1] + gr[2 * b + 1]) * sep[2 * k + 1]);
            double contrast = (rho[a] - rho[b]) - proj;
            double radial = -(sep[2 * k] * grd[2 * k] + sep[2 * k + 1] * grd[2 * k + 1]);
            double d2 = sep[2 * k] * sep[2 * k] + sep[2 * k + 1] * sep[2 * k + 1];
            cp[k] = contrast * radial / (d2 + 1.0e-12);
        }
        for (int64_t k = 0; k < nk; ++k) if (kj[k] < N) hr[ki[k]] += -scale * (mass[kj[k]] / rho[kj[k]]) * cp[k];
        for (int64_t k = 0; k < nk; ++k) if (kj[k] < N) hr[kj[k]] += scale * (mass[ki[k]] / rho[ki[k]]) * cp[k];
        for (int64_t a = 0; a < N; ++a) rate[a] += hr[a];
        free(gr); free(cp); free(hr);
    }
    for (int64_t a = 0; a < N; ++a) {
        rho_out[a] = rho[a] + P->dt * rate[a];
        if (!(isfinite(rho_out[a]) && rho_out[a] > 0.0)) status = status ? status : 2;
    }
    /* interface band */
    double* glam = (double*)malloc((size_t)N * 2 * sizeof(double));
    DP_GRADIENT(lam, glam);
    double* nrm = (double*)malloc((size_t)N * 2 * sizeof(double));
    double* strength = (double*)malloc((size_t)N * sizeof(double));
    for (int64_t a = 0; a < N; ++a) {
        double mg = sqrt(glam[2 * a] * glam[2 * a] + glam[2 * a + 1] * glam[2 * a + 1]);
        double mc = mg > 1.0e-12 ? mg : 1.0e-12;
        nrm[2 * a] = -glam[2 * a] / mc; nrm[2 * a + 1] = -glam[2 * a + 1] / mc;
        strength[a] = mg * h;
    }
    double* covered = (double*)calloc((size_t)N, sizeof(double));
    for (int pass = 0; pass < 2; ++pass)
        for (int64_t k = 0; k < nk; ++k) {
            if (kj[k] >= N) continue;
            int64_t a = ki[k], b = kj[k];
            if (pass == 0) covered[a] += (double)occupied(-sep[2 * k], -sep[2 * k + 1], nrm[2 * a], nrm[2 * a + 1], h);
            else covered[b] += (double)occupied(sep[2 * k], sep[2 * k + 1], nrm[2 * b], nrm[2 * b + 1], h);
        }
    unsigned char* surface = (unsigned char*)malloc((size_t)N);
    for (int64_t a = 0; a < N; ++a) surface[a] = (lam[a] < 0.2) || ((lam[a] < 0.75) && (covered[a] == 0.0));
    double* nbs = (double*)calloc((size_t)N, sizeof(double));
    for (int pass = 0; pass < 2; ++pass)
        for (int64_t k = 0; k < nk; ++k) {
            if (kj[k] >= N) continue;
            int64_t a = ki[k], b = kj[k];
            if (pass == 0) nbs[a] += (double)surface[b]; else nbs[b] += (double)surface[a];
        }
    /* shifting (Sun et al. eq. 5) */
    double wref, sref; wendland2d(dx, h, &wref, &sref);
    double* raw = (double*)calloc((size_t)N * 2, sizeof(double));
    double* rep = (double*)malloc((size_t)(nk > 0 ? nk : 1) * 2 * sizeof(double));
    for (int64_t k = 0; k < nk; ++k) {
        double ratio = wgt[k] / wref; double r2 = ratio * ratio;
        double f = 1.0 + 0.2 * (r2 * r2);
        rep[2 * k] = f * grd[2 * k]; rep[2 * k + 1] = f * grd[2 * k + 1];
    }
    for (int pass = 0; pass < 2; ++pass)
        for (int64_t k = 0; k < nk; ++k) {
            if (kj[k] >= N) continue;
            int64_t a = ki[k], b = kj[k];
            double den = rho[a] + rho[b];
            double qx = rep[2 * k] / den, qy = rep[2 * k + 1] / den;
            if (pass == 0) { raw[2 * a] += mass[b] * qx; raw[2 * a + 1] += mass[b] * qy; }
            else { raw[2 * b] += -mass[a] * qx; raw[2 * b + 1] += -mass[a] * qy; }
        }
    for (int64_t k = 0; k < nk; ++k) {
        if (kj[k] < N) continue;
        int64_t f = ki[k], w = kj[k] - N;
        double ratio = wall_rho[w] * (dx * dx) / (rho[f] + wall_rho[w]);
        raw[2 * f] += ratio * rep[2 * k]; raw[2 * f + 1] += ratio * rep[2 * k + 1];
    }
    double fac = -(P->dt * c0 / h) * (P->ref_speed / c0) * ((2.0 * h) * (2.0 * h));
    for (int64_t a = 0; a < N * 2; ++a) raw[a] *= fac;
    double* sh = (double*)malloc((size_t)N * 2 * sizeof(double));
    unsigned char* band = (unsigned char*)malloc((size_t)N);
    for (int64_t a = 0; a < N; ++a) {
        band[a] = surface[a] || (nbs[a] > 0.0);
        double rn = raw[2 * a] * nrm[2 * a] + raw[2 * a + 1] * nrm[2 * a + 1];
        double tx = raw[2 * a] - rn * nrm[2 * a], ty = raw[2 * a + 1] - rn * nrm[2 * a + 1];
        int resolved = (lam[a] >= 0.55) && (strength[a] > 1.0e-6);
        if (band[a]) { sh[2 * a] = resolved ? tx : 0.0; sh[2 * a + 1] = resolved ? ty : 0.0; }
        else { sh[2 * a] = raw[2 * a]; sh[2 * a + 1] = raw[2 * a + 1]; }
    }
    /* _wall_tangent */
    int64_t nwp = 0; for (int64_t k = 0; k < nk; ++k) if (kj[k] >= N) ++nwp;
    if (nwp) {
        double* nearest = (double*)malloc((size_t)N * sizeof(double));
        int64_t* pick = (int64_t*)malloc((size_t)N * sizeof(int64_t));
        for (int64_t a = 0; a < N; ++a) { nearest[a] = INFINITY; pick[a] = -1; }
        for (int64_t k = 0; k < nk; ++k) {
            if (kj[k] < N) continue;
            double d2 = sep[2 * k] * sep[2 * k] + sep[2 * k + 1] * sep[2 * k + 1];
            if (d2 < nearest[ki[k]]) nearest[ki[k]] = d2;
        }
        for (int64_t k = 0; k < nk; ++k) {
            if (kj[k] < N) continue;
            double d2 = sep[2 * k] * sep[2 * k] + sep[2 * k + 1] * sep[2 * k + 1];
            if (d2 == nearest[ki[k]] && pick[ki[k]] < 0) pick[ki[k]] = k;
        }
        double* tg = (double*)malloc((size_t)N * 2 * sizeof(double));
        unsigned char* conflict = (unsigned char*)malloc((size_t)N);
        double* tol = (double*)malloc((size_t)N * sizeof(double));
        for (int64_t a = 0; a < N; ++a) {
            double sx = sh[2 * a], sy = sh[2 * a + 1];
            tol[a] = 1.0e-6 * sqrt(sx * sx + sy * sy);
            if (pick[a] >= 0) {
                int64_t w = kj[pick[a]] - N;
                double nx = wall_n[2 * w], ny = wall_n[2 * w + 1];
                double sn = sx * nx + sy * ny;
                tg[2 * a] = sx - sn * nx; tg[2 * a + 1] = sy - sn * ny;
            } else { tg[2 * a] = sx; tg[2 * a + 1] = sy; }
            conflict[a] = band[a] && (fabs(tg[2 * a] * nrm[2 * a] + tg[2 * a + 1] * nrm[2 * a + 1]) > tol[a]);
        }
        for (int64_t k = 0; k < nk; ++k) {
            if (kj[k] < N) continue;
            int64_t a = ki[k];
            double d2 = sep[2 * k] * sep[2 * k] + sep[2 * k + 1] * sep[2 * k + 1];
            if (d2 != nearest[a]) continue;
            int64_t w = kj[k] - N;
            if (fabs(tg[2 * a] * wall_n[2 * w] + tg[2 * a + 1] * wall_n[2 * w + 1]) > tol[a]) conflict[a] = 1;
        }
        for (int64_t a = 0; a < N; ++a)
            if (isfinite(nearest[a])) {
                sh[2 * a] = conflict[a] ? 0.0 : tg[2 * a];
                sh[2 * a + 1] = conflict[a] ? 0.0 : tg[2 * a + 1];
            }
        free(nearest); free(pick); free(tg); free(conflict); free(tol);
    }
    for (int64_t a = 0; a < N; ++a) {
        vel_out[2 * a] = vnew[2 * a]; vel_out[2 * a + 1] = vnew[2 * a + 1];
        shift_out[2 * a] = sh[2 * a]; shift_out[2 * a + 1] = sh[2 * a + 1];
        pos_out[2 * a] = (pos[2 * a] + P->dt * vnew[2 * a]) + sh[2 * a];
        pos_out[2 * a + 1] = (pos[2 * a + 1] + P->dt * vnew[2 * a + 1]) + sh[2 * a + 1];
    }
#undef DP_GRADIENT
    free(ki); free(kj); free(sep); free(grd); free(wgt); free(vol); free(p); free(psum); free(wsum);
    free(wall_p); free(wall_rho); free(net); free(fz); free(wf); free(vnew); free(mom); free(inv);
    free(lam); free(div); free(rate); free(glam); free(nrm); free(strength); free(covered);
    free(surface); free(nbs); free(raw); free(rep); free(sh); free(band);
    return status;
}
