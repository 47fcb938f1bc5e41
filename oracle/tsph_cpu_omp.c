/* Multi-threaded CPU port of the 2-D WCSPH dam-break step (OpenMP), used only
 * as the CPU baseline of the benchmark and checked against the oracle's
 * Stepper2D in the CPU tests.  Test and measurement infrastructure: never
 * linked into the product.
 *
 * One step is the oracle's Stepper2D.step (oracle.py) with the rhs of
 * orc_rhs_morris (tsph_oracle.c) on its default branch, the reference's
 * droplet_impact.rhs (droplet_impact.py:1004-1199) without the opt-in
 * branches: summation density with the 0.1 rho0 floor, Tait EOS, Adami wall
 * pressure with gravity, the symmetric pressure force, Morris viscosity on
 * fluid pairs (0.01 h^2 regulariser), gravity, symplectic Euler and the box
 * clamp (droplet_impact.py:1293-1312), with the Verlet pair cache rebuilt
 * when a row has moved more than half the skin (neighbor_search.py:166-227).
 *
 * Formulated per row (each row gathers over its own neighbour list, built
 * with cells of edge 2h + skin), so the rows split across threads with no
 * shared writes; the oracle's pair-scatter form sums the same terms in
 * another order (FP64 differences at rounding level).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define API __attribute__((visibility("default")))

typedef struct {
    double h, rho0, c0, gamma, gravity, nu, dt, skin, wall_mass;
    double lo[2], hi[2], tol;
} omp_params;

typedef struct {
    omp_params P;
    int64_t n, nw, nt;
    double* pos;      /* nt x 2: fluid then walls */
    double* vel;      /* n x 2 */
    double* mass;     /* nt */
    double* anchor;   /* nt x 2, positions at the last list build */
    int64_t* off;     /* nt + 1, CSR list offsets */
    int32_t* idx;
    int64_t cap;
    double *rho, *p, *acc, *wp, *wr;
    int64_t rebuilds, escaped;
    int built;
} omp_sim;

static void kernel2d(double r, double h, double* w, double* s) {
    double q = r / h, sigma = 10.0 / (7.0 * M_PI * h * h);
    double first = 1.0 - 1.5 * (q * q) + 0.75 * (q * q * q);
    double t = 2.0 - q; if (t < 0.0) t = 0.0;
    double second = 0.25 * (t * t * t);
    double shape = q < 1.0 ? first : (q < 2.0 ? second : 0.0);
    double dshape = q < 1.0 ? (-3.0 * q + 2.25 * (q * q)) : (q < 2.0 ? -0.75 * ((2.0 - q) * (2.0 - q)) : 0.0);
    if (shape < 0.0) shape = 0.0;
    *w = sigma * shape;
    double rc = r > 1.0e-12 ? r : 1.0e-12;
    *s = sigma * dshape / (h * rc);
}

static double tait(double rho, const omp_params* P) {
    double bulk = P->rho0 * (P->c0 * P->c0) / P->gamma;
    double ratio = rho / P->rho0; if (ratio < 0.01) ratio = 0.01;
    double p = bulk * (pow(ratio, P->gamma) - 1.0);
    return p < 0.0 ? 0.0 : p;
}

API omp_sim* omp_sim_create(const omp_params* P, int64_t n, int64_t nw, const double* pos, const double* vel,
                            const double* mass, const double* wall) {
    omp_sim* s = (omp_sim*)calloc(1, sizeof(omp_sim));
    s->P = *P; s->n = n; s->nw = nw; s->nt = n + nw;
    s->pos = (double*)malloc(sizeof(double) * 2 * (size_t)s->nt);
    memcpy(s->pos, pos, sizeof(double) * 2 * (size_t)n);
    memcpy(s->pos + 2 * n, wall, sizeof(double) * 2 * (size_t)nw);
    s->vel = (double*)malloc(sizeof(double) * 2 * (size_t)(n > 0 ? n : 1));
    memcpy(s->vel, vel, sizeof(double) * 2 * (size_t)n);
    s->mass = (double*)malloc(sizeof(double) * (size_t)s->nt);
    memcpy(s->mass, mass, sizeof(double) * (size_t)n);
    for (int64_t w = 0; w < nw; ++w) s->mass[n + w] = P->wall_mass;
    s->anchor = (double*)malloc(sizeof(double) * 2 * (size_t)s->nt);
    s->off = (int64_t*)malloc(sizeof(int64_t) * (size_t)(s->nt + 1));
    s->rho = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    s->p = (double*)malloc(sizeof(double) * (size_t)s->nt);
    s->acc = (double*)malloc(sizeof(double) * 2 * (size_t)(n > 0 ? n : 1));
    s->wp = (double*)malloc(sizeof(double) * (size_t)(nw > 0 ? nw : 1));
    s->wr = (double*)malloc(sizeof(double) * (size_t)s->nt);
    return s;
}

API void omp_sim_destroy(omp_sim* s) {
    if (!s) return;
    free(s->pos); free(s->vel); free(s->mass); free(s->anchor); free(s->off); free(s->idx);
    free(s->rho); free(s->p); free(s->acc); free(s->wp); free(s->wr); free(s);
}

/* lists at radius rc = 2h + skin over cells of edge rc (fluid rows list
 * fluid and wall rows, wall rows list fluid rows) */
static void build_lists(omp_sim* s) {
    const double rc = 2.0 * s->P.h + s->P.skin, rc2 = rc * rc;
    const int64_t nt = s->nt, n = s->n;
    double mn[2] = {INFINITY, INFINITY}, mx[2] = {-INFINITY, -INFINITY};
    for (int64_t a = 0; a < nt; ++a)
        for (int d = 0; d < 2; ++d) {
            double v = s->pos[2 * a + d];
            mn[d] = v < mn[d] ? v : mn[d]; mx[d] = v > mx[d] ? v : mx[d];
        }
    int64_t c0[2], ext[2];
    for (int d = 0; d < 2; ++d) { c0[d] = (int64_t)floor(mn[d] / rc); ext[d] = (int64_t)floor(mx[d] / rc) - c0[d] + 1; }
    const int64_t nc = ext[0] * ext[1];
    int64_t* cell = (int64_t*)malloc(sizeof(int64_t) * (size_t)nt);
    int64_t* start = (int64_t*)calloc((size_t)nc + 1, sizeof(int64_t));
    int64_t* order = (int64_t*)malloc(sizeof(int64_t) * (size_t)nt);
    for (int64_t a = 0; a < nt; ++a) {
        int64_t cx = (int64_t)floor(s->pos[2 * a] / rc) - c0[0], cy = (int64_t)floor(s->pos[2 * a + 1] / rc) - c0[1];
        cell[a] = cx * ext[1] + cy;
        start[cell[a] + 1]++;
    }
    for (int64_t c = 0; c < nc; ++c) start[c + 1] += start[c];
    int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)nc);
    memcpy(fill, start, sizeof(int64_t) * (size_t)nc);
    for (int64_t a = 0; a < nt; ++a) order[fill[cell[a]]++] = a;
    free(fill);
    /* count, then fill */
    for (int pass = 0; pass < 2; ++pass) {
        #pragma omp parallel for schedule(dynamic, 256)
        for (int64_t a = 0; a < nt; ++a) {
            const int64_t cx = cell[a] / ext[1], cy = cell[a] % ext[1];
            int64_t k = 0, base = pass ? s->off[a] : 0;
            for (int64_t ox = -1; ox <= 1; ++ox)
                for (int64_t oy = -1; oy <= 1; ++oy) {
                    int64_t x = cx + ox, y = cy + oy;
                    if (x < 0 || x >= ext[0] || y < 0 || y >= ext[1]) continue;
                    int64_t c = x * ext[1] + y;
                    for (int64_t t = start[c]; t < start[c + 1]; ++t) {
                        int64_t b = order[t];
                        if (b == a || (a >= n && b >= n)) continue;
                        double dx = s->pos[2 * a] - s->pos[2 * b], dy = s->pos[2 * a + 1] - s->pos[2 * b + 1];
                        if (dx * dx + dy * dy <= rc2) {
                            if (pass) s->idx[base + k] = (int32_t)b;
                            ++k;
                        }
                    }
                }
            if (!pass) s->off[a + 1] = k;
        }
        if (!pass) {
            s->off[0] = 0;
            for (int64_t a = 0; a < nt; ++a) s->off[a + 1] += s->off[a];
            if (s->off[nt] > s->cap) {
                free(s->idx);
                s->cap = s->off[nt] + s->off[nt] / 8 + 1;
                s->idx = (int32_t*)malloc(sizeof(int32_t) * (size_t)s->cap);
            }
        }
    }
    free(cell); free(start); free(order);
    memcpy(s->anchor, s->pos, sizeof(double) * 2 * (size_t)nt);
    s->rebuilds++;
    s->built = 1;
}

API void omp_sim_step(omp_sim* s, int64_t steps) {
    const omp_params* P = &s->P;
    const int64_t n = s->n, nw = s->nw, nt = s->nt;
    const double h = P->h, h2 = 4.0 * h * h;
    double w0, s0;
    kernel2d(0.0, h, &w0, &s0);
    const double bulk = P->rho0 * (P->c0 * P->c0) / P->gamma;
    for (int64_t it = 0; it < steps; ++it) {
        /* Verlet cache: rebuild when a row moved more than skin / 2 */
        double disp = s->built ? 0.0 : INFINITY;
        if (s->built) {
            #pragma omp parallel for reduction(max : disp)
            for (int64_t a = 0; a < n; ++a) {
                double dx = s->pos[2 * a] - s->anchor[2 * a], dy = s->pos[2 * a + 1] - s->anchor[2 * a + 1];
                double d = sqrt(dx * dx + dy * dy);
                disp = d > disp ? d : disp;
            }
        }
        if (!s->built || disp > 0.5 * P->skin) build_lists(s);
        /* density + EOS */
        #pragma omp parallel for schedule(static)
        for (int64_t a = 0; a < n; ++a) {
            double r = s->mass[a] * w0;
            for (int64_t t = s->off[a]; t < s->off[a + 1]; ++t) {
                int64_t b = s->idx[t];
                double dx = s->pos[2 * a] - s->pos[2 * b], dy = s->pos[2 * a + 1] - s->pos[2 * b + 1];
                double d2 = dx * dx + dy * dy;
                if (d2 > h2) continue;
                double w, g; kernel2d(sqrt(d2), h, &w, &g);
                r += s->mass[b] * w;
            }
            r = r < 0.1 * P->rho0 ? 0.1 * P->rho0 : r;
            s->rho[a] = r; s->p[a] = tait(r, P); s->wr[a] = r;
        }
        /* Adami wall pressure (gravity body term), wall density from Tait */
        #pragma omp parallel for schedule(static)
        for (int64_t w = 0; w < nw; ++w) {
            const int64_t a = n + w;
            double num = 0.0, den = 0.0;
            for (int64_t t = s->off[a]; t < s->off[a + 1]; ++t) {
                int64_t f = s->idx[t];
                double dx = s->pos[2 * f] - s->pos[2 * a], dy = s->pos[2 * f + 1] - s->pos[2 * a + 1];
                double d2 = dx * dx + dy * dy;
                if (d2 > h2) continue;
                double wk, g; kernel2d(sqrt(d2), h, &wk, &g);
                double src = s->p[f];
                if (P->gravity != 0.0) src = src + s->rho[f] * (-P->gravity * (s->pos[2 * a + 1] - s->pos[2 * f + 1]));
                num += src * wk; den += wk;
            }
            double dd = den > 1.0e-30 ? den : 1.0e-30;
            double pw = num / dd; if (pw < 0.0) pw = 0.0;
            double rw = P->rho0 * pow(pw / bulk + 1.0, 1.0 / P->gamma); if (rw < P->rho0) rw = P->rho0;
            s->p[a] = pw; s->wr[a] = rw; s->wp[w] = pw;
        }
        /* pressure force, Morris viscosity on fluid pairs, gravity; integrate + clamp */
        int64_t esc = 0;
        #pragma omp parallel for schedule(static) reduction(+ : esc)
        for (int64_t a = 0; a < n; ++a) {
            double ax = 0.0, ay = 0.0;
            const double pa = s->p[a] / (s->wr[a] * s->wr[a]);
            for (int64_t t = s->off[a]; t < s->off[a + 1]; ++t) {
                int64_t b = s->idx[t];
                double dx = s->pos[2 * a] - s->pos[2 * b], dy = s->pos[2 * a + 1] - s->pos[2 * b + 1];
                double d2 = dx * dx + dy * dy;
                if (d2 > h2) continue;
                double r = sqrt(d2), w, g; kernel2d(r, h, &w, &g);
                double gx = dx * g, gy = dy * g;
                double pt = pa + s->p[b] / (s->wr[b] * s->wr[b]);
                ax += -s->mass[b] * (pt * gx); ay += -s->mass[b] * (pt * gy);
                if (b < n) {
                    double dg = dx * gx + dy * gy, relu = -dg > 0.0 ? -dg : 0.0;
                    double lw = relu / (r * r + 0.01 * h * h);
                    double sc = 2.0 * P->nu * (s->mass[b] / s->wr[b] * lw);
                    ax += -sc * (s->vel[2 * a] - s->vel[2 * b]);
                    ay += -sc * (s->vel[2 * a + 1] - s->vel[2 * b + 1]);
                }
            }
            ay += -P->gravity;
            s->acc[2 * a] = ax; s->acc[2 * a + 1] = ay;
        }
        #pragma omp parallel for schedule(static) reduction(+ : esc)
        for (int64_t a = 0; a < n; ++a) {
            int hit_any = 0;
            for (int d = 0; d < 2; ++d) {
                double v = s->vel[2 * a + d] + P->dt * s->acc[2 * a + d];
                double x = s->pos[2 * a + d] + P->dt * v;
                int hit = (x < P->lo[d] - P->tol) || (x > P->hi[d] + P->tol);
                hit_any |= hit;
                x = x < P->lo[d] ? P->lo[d] : (x > P->hi[d] ? P->hi[d] : x);
                if (hit) v = 0.0;
                s->vel[2 * a + d] = v; s->pos[2 * a + d] = x;
            }
            esc += hit_any;
        }
        s->escaped += esc;
    }
}

API void omp_sim_read(const omp_sim* s, double* pos, double* vel, double* rho, double* acc, int64_t* rebuilds) {
    if (pos) memcpy(pos, s->pos, sizeof(double) * 2 * (size_t)s->n);
    if (vel) memcpy(vel, s->vel, sizeof(double) * 2 * (size_t)s->n);
    if (rho) memcpy(rho, s->rho, sizeof(double) * (size_t)s->n);
    if (acc) memcpy(acc, s->acc, sizeof(double) * 2 * (size_t)s->n);
    if (rebuilds) *rebuilds = s->rebuilds;
}

API int omp_threads(int want) {
#ifdef _OPENMP
    if (want > 0) omp_set_num_threads(want);
    return omp_get_max_threads();
#else
    (void)want;
    return 1;
#endif
}
