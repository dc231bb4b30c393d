/*
 * CPU oracle for image TV: a plain-C restatement of fastprox's algorithms for
 * the image-TV hot path. TEST INFRASTRUCTURE ONLY: the test suite, the smoke
 * check and bench.py's CPU-baseline leg load this library as the checker;
 * the product path (fastprox_amd, libfastprox_rocm) never links or calls it.
 *
 * Every function works on one image in C order, in double, serially. Sums
 * follow the reference's fixed order (per row in index order, then the rows
 * in order: src/ops/image_util.hpp:151-161), and every expression is written
 * in the reference's operation order, so with -ffp-contract=off the iso
 * solver reproduces the reference's bits.
 *
 * What each part restates:
 *   or_tv_iso          src/ops/tv_iso.cpp:333-415 (prox_img), :424-543 (fdpg),
 *                      :180-185 (radius), :251-306 (divergence, grad_rows, gap),
 *                      src/ops/image_util.hpp:22-72 (Geom, rescale, fista_t)
 *   or_tv1d            the exact 1D TV prox the single-active-axis branch calls
 *                      (tv_iso.cpp:381-398); L. Condat, IEEE SPL 20(11), 2013
 *   or_tv_aniso_dev    the device method for anisotropic TV:
 *                      bindings/python/fastprox/_tv_torch.py:149-186 (box
 *                      projection, gap), :229-283 (schedule), :588-606 (dual
 *                      forms); per-pixel arithmetic in tv_iso.cpp's form
 *   or_zerr            src/core/driver.hpp:1186-1236 (certified_error)
 *   or_iso_gap, or_aniso_gap   the certificates on a given (z, dual)
 *
 * Parity of each function is pinned by tests/test_oracle.py (against the
 * reference library built from source, the CVXPY fixtures, closed forms and
 * invariants).
 *
 * SPDX-License-Identifier: MIT
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define OR_OK 0
#define OR_EINF 1   /* an infinite entry: the reference's FP_EINVAL */
#define OR_ENOMEM 2

typedef struct {
    int k;
    int64_t L[3], S[3], N;
} geom;

static geom make_geom(int k, const int64_t *shape) {
    geom g;
    g.k = k;
    for (int a = 0; a < 3; ++a) { g.L[a] = 1; g.S[a] = 1; }
    for (int a = 0; a < k; ++a) g.L[a] = shape[a];
    g.S[k - 1] = 1;
    for (int a = k - 2; a >= 0; --a) g.S[a] = g.S[a + 1] * g.L[a + 1];
    g.N = g.S[0] * g.L[0];
    return g;
}
static int64_t g_rows(const geom *g) { return g->N / g->L[g->k - 1]; }
static int64_t g_coord(const geom *g, int64_t o, int a) { return (o / g->S[a]) % g->L[a]; }
static int g_has_next(const geom *g, int64_t o, int a) { return g_coord(g, o, a) + 1 < g->L[a]; }

/* std::max / std::min semantics (the first argument on ties and NaN). */
static double stdmax(double a, double b) { return a < b ? b : a; }
static double stdmin(double a, double b) { return b < a ? b : a; }

static double fista_t(double t) { return 0.5 * (1.0 + sqrt(1.0 + 4.0 * t * t)); }

/* image_util.hpp:56-65 */
static int rescale_exponent(const double *d, int64_t N, double gamma, double step) {
    double m = 0.0;
    for (int64_t i = 0; i < N; ++i) m = stdmax(m, fabs(d[i]));
    if (!(m > 0.0)) return 0;
    int e = ilogb(m);
    if (e >= -64 && e <= 64) return 0;
    const double top = stdmax(gamma, step);
    if (e < 0 && top > 0.0) {
        int lo = ilogb(top) + 1 - 960; /* ilogb(kMaxScaledStep = 2^960) */
        int t = e > lo ? e : lo;
        e = t < 0 ? t : 0;
    }
    return e;
}

static void scale2(double *v, int64_t n, int e) {
    for (int64_t i = 0; i < n; ++i) v[i] = ldexp(v[i], e);
}

/* tv_iso.cpp:180-185 */
static double radius(double gamma, const double *al, int64_t p, double bound) {
    if (!al) return gamma;
    if (al[p] == 0.0) return 0.0;
    const double r = gamma * al[p];
    return r <= bound ? r : bound;
}

/* tv_iso.cpp:251-270: z = d + sum_a w_a (u_a[p] - u_a[p - e_a]). */
static void divergence(const geom *g, const double *wa, const double *d, const double *u, double *z) {
    const int k = g->k;
    const int64_t n = g->L[k - 1], N = g->N, R = g_rows(g);
    for (int64_t r = 0; r < R; ++r) {
        const int64_t b = r * n;
        for (int64_t j = 0; j < n; ++j) z[b + j] = d[b + j];
        for (int a = 0; a < k; ++a) {
            if (wa[a] == 0.0) continue;
            const double *ua = u + a * N;
            if (a == k - 1) {
                for (int64_t j = 0; j < n; ++j) z[b + j] += wa[a] * (ua[b + j] - (j > 0 ? ua[b + j - 1] : 0.0));
            } else {
                const int prev = g_coord(g, b, a) > 0;
                for (int64_t j = 0; j < n; ++j)
                    z[b + j] += wa[a] * (ua[b + j] - (prev ? ua[b + j - g->S[a]] : 0.0));
            }
        }
    }
}

/* gvec of tv_iso.cpp:274-288 at entry p = b + j. */
static void grad_at(const geom *g, const double *wa, const double *z, int64_t b, int64_t j, const int *nxt,
                    double *gv) {
    const int k = g->k;
    const int64_t n = g->L[k - 1], p = b + j;
    for (int a = 0; a < k; ++a) {
        const int has = a == k - 1 ? j + 1 < n : nxt[a];
        gv[a] = (wa[a] != 0.0 && has) ? wa[a] * (z[p + g->S[a]] - z[p]) : 0.0;
    }
}

static void row_next(const geom *g, int64_t b, int *nxt) {
    for (int a = 0; a + 1 < g->k; ++a) nxt[a] = g_has_next(g, b, a);
}

/* tv_iso.cpp:290-306, the fixed order of row_sum. */
static double iso_gap_impl(const geom *g, const double *wa, double gamma, double bound, const double *al,
                           const double *z, const double *u) {
    const int k = g->k;
    const int64_t n = g->L[k - 1], N = g->N, R = g_rows(g);
    double tot = 0.0;
    for (int64_t r = 0; r < R; ++r) {
        const int64_t b = r * n;
        int nxt[3];
        row_next(g, b, nxt);
        double s = 0.0;
        for (int64_t j = 0; j < n; ++j) {
            double gv[3];
            grad_at(g, wa, z, b, j, nxt, gv);
            double nn = 0.0, dot = 0.0;
            for (int a = 0; a < k; ++a) {
                nn += gv[a] * gv[a];
                dot += gv[a] * u[a * N + b + j];
            }
            s += radius(gamma, al, b + j, bound) * sqrt(nn) - dot;
        }
        tot += s;
    }
    return tot;
}

static double row_sum_fn(const geom *g, const double *d, int what, double mean) {
    const int64_t n = g->L[g->k - 1], R = g_rows(g);
    double tot = 0.0;
    for (int64_t r = 0; r < R; ++r) {
        double s = 0.0;
        for (int64_t j = 0; j < n; ++j) {
            const double v = d[r * n + j];
            s += what == 0 ? v : (what == 1 ? fabs(v - mean) : v * v);
        }
        tot += s;
    }
    return tot;
}

/* ---------------------------------------------------------------------------
 * Exact 1D TV: prox of lam * sum |z_{i+1} - z_i|, Condat's direct algorithm
 * (the published pseudo-code, unit weights). n >= 1, lam >= 0.
 * ------------------------------------------------------------------------- */
void or_tv1d(const double *y, double *x, int64_t n, double lam) {
    if (n <= 0) return;
    if (n == 1 || !(lam > 0.0)) {
        for (int64_t i = 0; i < n; ++i) x[i] = y[i];
        return;
    }
    int64_t k = 0, k0 = 0, kmin = 0, kplus = 0;
    double vmin = y[0] - lam, vmax = y[0] + lam;
    double umin = lam, umax = -lam;
    const double twol = 2.0 * lam;
    for (;;) {
        if (k == n - 1) {
            if (umin < 0.0) {
                for (int64_t i = k0; i <= kmin; ++i) x[i] = vmin;
                k = k0 = kmin = kmin + 1;
                vmin = y[k];
                umin = lam;
                umax = y[k] + lam - vmax;
                continue;
            } else if (umax > 0.0) {
                for (int64_t i = k0; i <= kplus; ++i) x[i] = vmax;
                k = k0 = kplus = kplus + 1;
                vmax = y[k];
                umax = -lam;
                umin = y[k] - lam - vmin;
                continue;
            } else {
                const double v = vmin + umin / (double)(k - k0 + 1);
                for (int64_t i = k0; i <= k; ++i) x[i] = v;
                return;
            }
        }
        if (y[k + 1] + umin < vmin - lam) {
            for (int64_t i = k0; i <= kmin; ++i) x[i] = vmin;
            k = k0 = kmin = kplus = kmin + 1;
            vmin = y[k];
            vmax = y[k] + twol;
            umin = lam;
            umax = -lam;
        } else if (y[k + 1] + umax > vmax + lam) {
            for (int64_t i = k0; i <= kplus; ++i) x[i] = vmax;
            k = k0 = kmin = kplus = kplus + 1;
            vmin = y[k] - twol;
            vmax = y[k];
            umin = lam;
            umax = -lam;
        } else {
            k = k + 1;
            umin = umin + y[k] - vmin;
            umax = umax + y[k] - vmax;
            if (umin >= lam) {
                vmin = vmin + (umin - lam) / (double)(k - k0 + 1);
                umin = lam;
                kmin = k;
            }
            if (umax <= -lam) {
                vmax = vmax + (umax + lam) / (double)(k - k0 + 1);
                umax = -lam;
                kplus = k;
            }
        }
    }
}

/* ---------------------------------------------------------------------------
 * Isotropic TV, dual FISTA with the reference's schedule (tv_iso.cpp:424-543).
 * ------------------------------------------------------------------------- */
static void fdpg(const geom *g, const double *w_in, const double *al, double gamma_in, double tol, int max_iter,
                 int warm, const double *d, double *z, double *u, double *v, double *dual, double *gap_out,
                 int *iters_out, int *conv_out) {
    const int k = g->k;
    const int64_t N = g->N, n = g->L[k - 1], R = g_rows(g);
    double wmax = 0.0, whmin = 1.0, wa[3] = {0.0, 0.0, 0.0};
    for (int a = 0; a < k; ++a) wmax = stdmax(wmax, w_in[a]);
    for (int a = 0; a < k; ++a)
        if (w_in[a] != 0.0) {
            wa[a] = w_in[a] / wmax;
            whmin = stdmin(whmin, wa[a]);
        }
    const double mean = row_sum_fn(g, d, 0, 0.0) / (double)N;
    const double imbalance = row_sum_fn(g, d, 1, mean);
    double gamma = gamma_in * wmax;
    const double bound = 2.0 * imbalance / whmin + DBL_MIN;
    if (!(gamma <= bound) && !al) gamma = bound;
    double L2 = 0.0;
    for (int a = 0; a < k; ++a) L2 += 4.0 * wa[a] * wa[a];
    const double tau = 1.0 / L2;
    if (warm)
        for (int64_t i = 0; i < k * N; ++i) u[i] = dual[i] * wmax;
    else
        for (int64_t i = 0; i < k * N; ++i) u[i] = 0.0;
    for (int64_t r = 0; r < R; ++r) {
        const int64_t b = r * n;
        for (int64_t j = 0; j < n; ++j) {
            const int64_t p = b + j;
            double nn = 0.0;
            for (int a = 0; a < k; ++a) {
                const int has = a == k - 1 ? j + 1 < n : g_has_next(g, b, a);
                double *ua = &u[a * N + p];
                if (wa[a] == 0.0 || !has || !isfinite(*ua)) *ua = 0.0;
                nn += *ua * *ua;
            }
            const double rp = radius(gamma, al, p, bound);
            if (sqrt(nn) > rp) {
                const double sc = rp / sqrt(nn);
                for (int a = 0; a < k; ++a) u[a * N + p] *= sc;
            }
        }
    }
    memcpy(v, u, (size_t)(k * N) * sizeof(double));
    const double target = tol * 0.5 * row_sum_fn(g, d, 2, 0.0);
    double t = 1.0;
    int it = 0, since = 0, last = 0;
    double gp = INFINITY;
    if (warm) {
        divergence(g, wa, d, u, z);
        gp = iso_gap_impl(g, wa, gamma, bound, al, z, u);
    }
    while (!(gp <= target) && it < max_iter) {
        divergence(g, wa, d, v, z);
        const double tn = fista_t(t);
        const double beta = (t - 1.0) / tn;
        t = tn;
        const int test = it % 4 == 0;
        double up = 0.0;
        for (int64_t r = 0; r < R; ++r) {
            const int64_t b = r * n;
            int nxt[3];
            row_next(g, b, nxt);
            double s = 0.0;
            for (int64_t j = 0; j < n; ++j) {
                const int64_t p = b + j;
                double gv[3], cv[3], nn = 0.0;
                grad_at(g, wa, z, b, j, nxt, gv);
                for (int a = 0; a < k; ++a) {
                    cv[a] = wa[a] != 0.0 ? v[a * N + p] + tau * gv[a] : 0.0;
                    nn += cv[a] * cv[a];
                }
                const double nr = sqrt(nn), rp = radius(gamma, al, p, bound);
                const double sc = nr > rp ? rp / nr : 1.0;
                for (int a = 0; a < k; ++a) {
                    const double un = cv[a] * sc;
                    double *ua = &u[a * N + p], *va = &v[a * N + p];
                    if (test) s += (*va - un) * (un - *ua);
                    *va = un + beta * (un - *ua);
                    *ua = un;
                }
            }
            up += s;
        }
        ++since;
        if (test && up > 0.0 && since >= last) {
            memcpy(v, u, (size_t)(k * N) * sizeof(double));
            t = 1.0;
            last = since;
            since = 0;
        }
        ++it;
        if (it % 4 == 0 || it == max_iter) {
            divergence(g, wa, d, u, z);
            gp = iso_gap_impl(g, wa, gamma, bound, al, z, u);
        }
    }
    if (!isfinite(gp)) {
        divergence(g, wa, d, u, z);
        gp = iso_gap_impl(g, wa, gamma, bound, al, z, u);
    }
    for (int64_t i = 0; i < k * N; ++i) dual[i] = u[i] / wmax;
    *gap_out = stdmax(gp, 0.0);
    *iters_out = it;
    *conv_out = gp <= target;
}

/*
 * or_tv_iso: one image x (shape[0..k-1], C order) -> z, dual (k N) with the
 * reference's TVIso::prox_img semantics (tv_iso.cpp:333-415). w: k axis
 * weights; alpha: pixel weights (N) or NULL. dual is read as the warm start
 * when warm != 0 and always written. Returns OR_EINF for an infinite entry.
 * *scale_exp receives the power-of-two rescaling exponent (may be NULL).
 */
int or_tv_iso(int k, const int64_t *shape, const double *w, const double *alpha, double gamma, double tol,
              int max_iter, int warm, const double *x, double *zout, double *dual, double *gap, int *iters,
              int *converged, int *scale_exp) {
    const geom g = make_geom(k, shape);
    const int64_t N = g.N;
    double *d = (double *)malloc((size_t)((2 + 2 * k) * N + 1) * sizeof(double));
    if (!d) return OR_ENOMEM;
    double *z = d + N, *u = z + N, *v = u + k * N;
    *gap = 0.0;
    *iters = 0;
    *converged = 1;
    if (scale_exp) *scale_exp = 0;
    int nan = 0;
    for (int64_t i = 0; i < N; ++i) {
        const double t = x[i];
        if (isinf(t)) {
            free(d);
            return OR_EINF;
        }
        nan = nan || isnan(t);
        d[i] = t;
    }
    double wa[3] = {0.0, 0.0, 0.0}, wmax = 0.0;
    for (int a = 0; a < k; ++a)
        if (g.L[a] > 1) wmax = stdmax(wmax, w[a]);
    int act[3], na = 0;
    for (int a = 0; a < k; ++a)
        if (gamma > 0.0 && w[a] > 0.0 && g.L[a] > 1 && w[a] >= 1e-100 * wmax) {
            wa[a] = w[a];
            act[na++] = a;
        }
    if (nan) {
        for (int64_t i = 0; i < N; ++i) zout[i] = NAN;
        for (int64_t i = 0; i < k * N; ++i) dual[i] = NAN;
        *gap = NAN;
        free(d);
        return OR_OK;
    }
    const double gamma0 = gamma;
    double step_scale = 0.0;
    for (int a = 0; a < k; ++a) step_scale = stdmax(step_scale, w[a]);
    if (alpha) {
        double am = 0.0;
        for (int64_t i = 0; i < N; ++i) am = stdmax(am, alpha[i]);
        step_scale *= am;
    }
    int e = na ? rescale_exponent(d, N, gamma, gamma * step_scale) : 0;
    for (int a = 0; a < k && e; ++a)
        if (wa[a] > 0.0 && !(ldexp(gamma0 * wa[a], -e) > 0.0)) e = 0;
    if (e) {
        scale2(d, N, -e);
        if (warm) scale2(dual, k * N, -e);
        gamma = ldexp(gamma0, -e);
    }
    if (na == 0) {
        memcpy(z, d, (size_t)N * sizeof(double));
        for (int64_t i = 0; i < k * N; ++i) dual[i] = 0.0;
    } else if (na == 1 && !alpha) {
        const int a = act[0];
        const double lam = gamma * w[a];
        for (int64_t i = 0; i < k * N; ++i) dual[i] = 0.0;
        double *qa = dual + a * N;
        const int64_t L = g.L[a], s = g.S[a];
        double *fin = (double *)malloc((size_t)(2 * L) * sizeof(double));
        if (!fin) {
            free(d);
            return OR_ENOMEM;
        }
        double *fout = fin + L;
        for (int64_t st = 0; st < N; ++st) {
            if (g_coord(&g, st, a) != 0) continue;
            for (int64_t i = 0; i < L; ++i) fin[i] = d[st + i * s];
            or_tv1d(fin, fout, L, lam);
            double cs = 0.0;
            for (int64_t i = 0; i < L; ++i) {
                const int64_t o = st + i * s;
                z[o] = fout[i];
                cs += d[o] - z[o];
                qa[o] = i + 1 < L ? -cs / w[a] : 0.0;
            }
        }
        free(fin);
    } else {
        fdpg(&g, wa, alpha, gamma, tol, max_iter, warm, d, z, u, v, dual, gap, iters, converged);
    }
    if (e) {
        scale2(z, N, e);
        scale2(dual, k * N, e);
        *gap = ldexp(*gap, 2 * e);
        if (scale_exp) *scale_exp = e;
    }
    for (int64_t i = 0; i < N; ++i) zout[i] = z[i];
    free(d);
    return OR_OK;
}

/* The iso certificate of a given (z, u): sum_p r_p |g_p| - <g_p, u_p>,
 * g = W D z with the raw axis weights and r_p = gamma alpha_p. */
double or_iso_gap(int k, const int64_t *shape, const double *w, const double *alpha, double gamma,
                  const double *z, const double *u) {
    const geom g = make_geom(k, shape);
    double wa[3] = {0.0, 0.0, 0.0};
    for (int a = 0; a < k; ++a) wa[a] = w[a];
    return iso_gap_impl(&g, wa, gamma, INFINITY, alpha, z, u);
}

/* ---------------------------------------------------------------------------
 * Anisotropic TV, the device method (_tv_torch.py:229-283): dual FISTA on the
 * per-edge duals u_a (|u_a[p]| <= gamma beta_a[p]), step 1/(4 sum w_a^2),
 * restart and gap check every 4th iteration (1-based), no interval rule.
 * ------------------------------------------------------------------------- */
static double aniso_gap_impl(const geom *g, const double *w, double gamma, const double *beta, const double *z,
                             const double *u) {
    const int k = g->k;
    const int64_t n = g->L[k - 1], N = g->N, R = g_rows(g);
    double tot = 0.0;
    for (int64_t r = 0; r < R; ++r) {
        const int64_t b = r * n;
        int nxt[3];
        row_next(g, b, nxt);
        double s = 0.0;
        for (int64_t j = 0; j < n; ++j) {
            double gv[3];
            grad_at(g, w, z, b, j, nxt, gv);
            for (int a = 0; a < k; ++a) {
                const double re = beta ? gamma * beta[a * N + b + j] : gamma;
                s += re * fabs(gv[a]) - gv[a] * u[a * N + b + j];
            }
        }
        tot += s;
    }
    return tot;
}

/* u (k N) from the C-form dual q: u_a = -cumsum_a(q_a) / w_a (_tv_torch.py:596-606). */
static void aniso_dual_in(const geom *g, const double *w, const double *q, double *u) {
    const int k = g->k;
    const int64_t N = g->N;
    for (int a = 0; a < k; ++a) {
        const int64_t L = g->L[a], s = g->S[a];
        for (int64_t st = 0; st < N; ++st) {
            if (g_coord(g, st, a) != 0) continue;
            double cs = 0.0;
            for (int64_t i = 0; i < L; ++i) {
                const int64_t o = st + i * s;
                cs += q[a * N + o];
                u[a * N + o] = w[a] > 0.0 ? -cs / w[a] : 0.0;
            }
        }
    }
}

/* q_a = w_a D_a^T u_a: q[i] = w (u[i-1] - u[i]), u[-1] = 0, u[L-1] taken as 0. */
static void aniso_dual_out(const geom *g, const double *w, const double *u, double *q) {
    const int k = g->k;
    const int64_t N = g->N;
    for (int a = 0; a < k; ++a) {
        const int64_t L = g->L[a], s = g->S[a];
        for (int64_t o = 0; o < N; ++o) {
            const int64_t i = g_coord(g, o, a);
            const double prev = i > 0 ? u[a * N + o - s] : 0.0;
            const double cur = i + 1 < L ? u[a * N + o] : 0.0;
            q[a * N + o] = w[a] * (prev - cur);
        }
    }
}

int or_tv_aniso_dev(int k, const int64_t *shape, const double *w, const double *beta, double gamma, double tol,
                    int max_iter, int warm, const double *x, double *zout, double *dual, double *gap,
                    int *iters, int *converged) {
    const geom g = make_geom(k, shape);
    const int64_t N = g.N, n = g.L[k - 1], R = g_rows(&g);
    double *buf = (double *)malloc((size_t)((1 + 3 * k) * N + 1) * sizeof(double));
    if (!buf) return OR_ENOMEM;
    double *z = buf, *u = z + N, *v = u + k * N, *up = v + k * N;
    int nan = 0;
    for (int64_t i = 0; i < N; ++i) {
        if (isinf(x[i])) {
            free(buf);
            return OR_EINF;
        }
        nan = nan || isnan(x[i]);
    }
    *iters = 0;
    *converged = 1;
    if (nan) {
        for (int64_t i = 0; i < N; ++i) zout[i] = NAN;
        for (int64_t i = 0; i < k * N; ++i) dual[i] = NAN;
        *gap = NAN;
        free(buf);
        return OR_OK;
    }
    double wa[3] = {0.0, 0.0, 0.0}, L2 = 0.0;
    for (int a = 0; a < k; ++a) {
        wa[a] = g.L[a] > 1 ? w[a] : 0.0;
        L2 += 4.0 * w[a] * w[a];
    }
    /* the radius of edge (a, p) */
#define RAD(a, p) (beta ? gamma * beta[(a) * N + (p)] : gamma)
    if (warm) {
        aniso_dual_in(&g, w, dual, u);
        for (int a = 0; a < k; ++a)
            for (int64_t p = 0; p < N; ++p) {
                const double r = RAD(a, p);
                double *ua = &u[a * N + p];
                *ua = r > 0.0 ? fmin(fmax(*ua, -r), r) : 0.0;
            }
    } else {
        for (int64_t i = 0; i < k * N; ++i) u[i] = 0.0;
    }
    const double target = tol * row_sum_fn(&g, x, 2, 0.0) / 2.0;
    double gp = 0.0;
    int it = 0;
    if (L2 == 0.0 || !(gamma > 0.0)) {
        for (int64_t i = 0; i < k * N; ++i) u[i] = 0.0;
        divergence(&g, wa, x, u, z);
        gp = aniso_gap_impl(&g, wa, gamma, beta, z, u);
    } else {
        const double tau = 1.0 / L2;
        memcpy(v, u, (size_t)(k * N) * sizeof(double));
        double t = 1.0;
        while (it < max_iter) {
            divergence(&g, wa, x, v, z);
            for (int64_t r = 0; r < R; ++r) {
                const int64_t b = r * n;
                int nxt[3];
                row_next(&g, b, nxt);
                for (int64_t j = 0; j < n; ++j) {
                    const int64_t p = b + j;
                    double gv[3];
                    grad_at(&g, wa, z, b, j, nxt, gv);
                    for (int a = 0; a < k; ++a) {
                        const double c = wa[a] != 0.0 ? v[a * N + p] + tau * gv[a] : 0.0;
                        const double rr = RAD(a, p);
                        up[a * N + p] = c > rr ? rr : (c < -rr ? -rr : c);
                    }
                }
            }
            ++it;
            double tn = fista_t(t);
            double coef = (t - 1.0) / tn;
            const int check = it % 4 == 0 || it == max_iter;
            if (check) {
                double s = 0.0;
                for (int64_t r = 0; r < R; ++r) {
                    double sr = 0.0;
                    for (int64_t j = 0; j < n; ++j)
                        for (int a = 0; a < k; ++a) {
                            const int64_t i = a * N + r * n + j;
                            sr += (v[i] - up[i]) * (up[i] - u[i]);
                        }
                    s += sr;
                }
                if (s > 0.0) {
                    coef = 0.0;
                    tn = 1.0;
                }
            }
            for (int64_t i = 0; i < k * N; ++i) {
                v[i] = up[i] + coef * (up[i] - u[i]);
                u[i] = up[i];
            }
            t = tn;
            if (check) {
                divergence(&g, wa, x, u, z);
                gp = aniso_gap_impl(&g, wa, gamma, beta, z, u);
                if (gp <= target) break;
            }
        }
        divergence(&g, wa, x, u, z);
        gp = aniso_gap_impl(&g, wa, gamma, beta, z, u);
    }
#undef RAD
    aniso_dual_out(&g, w, u, dual);
    for (int64_t i = 0; i < N; ++i) zout[i] = z[i];
    *gap = gp;
    *iters = it;
    *converged = gp <= target;
    free(buf);
    return OR_OK;
}

/* The aniso certificate from (z, q) in C form, as the reference's tests
 * recompute it: gamma sum_a w_a sum beta |D_a z| - <z, x - z>. */
double or_aniso_gap(int k, const int64_t *shape, const double *w, const double *beta, double gamma,
                    const double *x, const double *z) {
    const geom g = make_geom(k, shape);
    const int64_t N = g.N;
    double tv = 0.0, zd = 0.0;
    for (int64_t p = 0; p < N; ++p) {
        for (int a = 0; a < k; ++a) {
            if (!g_has_next(&g, p, a)) continue;
            const double b = beta ? beta[a * N + p] : 1.0;
            tv += w[a] * b * fabs(z[p + g.S[a]] - z[p]);
        }
        zd += z[p] * (x[p] - z[p]);
    }
    return gamma * tv - zd;
}

/* ---------------------------------------------------------------------------
 * Certified relative error of an image's output (driver.hpp:1186-1236): r =
 * sqrt(2 gap)/|x| plus the output dtype's rounding u (1 + r) and eta sqrt(N)
 * relative to |x|; |x|^2 summed in eight partial sums by position.
 * ------------------------------------------------------------------------- */
double or_zerr(double gap, const double *x, int64_t N, double u, double eta, int scale_exp, double scaled_gap) {
    double gp = gap, et = eta;
    if (scale_exp) {
        gp = scaled_gap;
        et = ldexp(eta, -scale_exp);
    }
    if (isnan(gp)) return gp;
    const double e = sqrt(2.0 * stdmax(gp, 0.0));
    const double a = et * sqrt((double)N);
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int64_t i = 0; i < N; ++i) {
        const double v = scale_exp ? ldexp(x[i], -scale_exp) : x[i];
        acc[i & 7] += v * v;
    }
    const double s = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
    if (s >= 1e-280 && s <= 1e280) {
        const double nrm = sqrt(s);
        return e / nrm + (u * (1.0 + e / nrm) + a / nrm);
    }
    double m = 0.0;
    for (int64_t i = 0; i < N; ++i) m = fmax(m, fabs(scale_exp ? ldexp(x[i], -scale_exp) : x[i]));
    if (!(m > 0.0)) return isnan(m) ? m : 0.0;
    double t = 0.0;
    for (int64_t i = 0; i < N; ++i) {
        const double v = (scale_exp ? ldexp(x[i], -scale_exp) : x[i]) / m;
        t += v * v;
    }
    const double rt = sqrt(t);
    const double r = e / m / rt;
    return r + (u * (1.0 + r) + a / m / rt);
}
