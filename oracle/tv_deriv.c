/*
 * CPU oracle for the derivatives of image TV (implicit differentiation at a
 * solution). TEST INFRASTRUCTURE ONLY, like tv_oracle.c: only the tests and
 * the smoke check load it, as the checker.
 *
 * Isotropic TV (or_iso_deriv) restates src/ops/tv_iso.cpp:
 *   the structure (TVIso::save, :551-631): flat pixels (dual strictly inside
 *     its ball, |u_p| < r_p (1 - 1e-9), or |g_p| at the rounding level
 *     w_max 64 2.2e-16 max|d|), regions = components of the forward edges of
 *     flat pixels (roots at the smallest index, labels in root order),
 *     members by a counting sort, 1/|C|;
 *   the system (Sys, :672-895): gz = W D z, the capped curvature
 *     h_p = r_p / max(|g_p|, mu), mu = 0.1 w_max sqrt(2 gap / N), the Jacobi
 *     diagonal, apply = M^T (I + G^T H G) M, region sums in chunks of 4096
 *     member positions, region dots in blocks of 4096 regions, row sums in
 *     blocks of 16384 / n rows;
 *   cg (:899-943), jvp_img (:952-997), vjp_img (:999-1067).
 * Every sum runs in the reference's fixed order and every expression in its
 * operation order (built with -ffp-contract=off), so the outputs are the
 * reference's bits for any image size.
 *
 * Anisotropic TV (or_aniso_deriv) restates the device method's derivative,
 * bindings/python/fastprox/_tv_torch.py:346-362 (flat edges, signs) and
 * :528-585 (region means; the gradients in gamma, w and edge weights).
 *
 * SPDX-License-Identifier: MIT
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define SEG 4096      /* tv_iso.cpp:119 (kSeg) */
#define REGBLK 4096   /* tv_iso.cpp:120 (kRegBlk) */
#define CHUNK 16384   /* src/core/workspace.hpp:16 (kChunk) */
#define DUAL_ETA 1e-9 /* tv_iso.cpp:109 */
#define CURV_RMS 0.1  /* tv_iso.cpp:113 */
#define DERIV_FLOOR 1e-10

typedef struct {
    int k;
    int64_t L[3], S[3], N, n, R;
} dgeom;

static dgeom mk(int k, const int64_t *shape) {
    dgeom g;
    g.k = k;
    for (int a = 0; a < 3; ++a) { g.L[a] = 1; g.S[a] = 1; }
    for (int a = 0; a < k; ++a) g.L[a] = shape[a];
    g.S[k - 1] = 1;
    for (int a = k - 2; a >= 0; --a) g.S[a] = g.S[a + 1] * g.L[a + 1];
    g.N = g.S[0] * g.L[0];
    g.n = g.L[k - 1];
    g.R = g.N / g.n;
    return g;
}
static int64_t crd(const dgeom *g, int64_t o, int a) { return (o / g->S[a]) % g->L[a]; }
static int hasnx(const dgeom *g, int64_t o, int a) { return crd(g, o, a) + 1 < g->L[a]; }
static double dmax(double a, double b) { return a < b ? b : a; }

/* The structure and the work arrays of one image. */
typedef struct {
    dgeom g;
    double gamma;       /* at the solved scale */
    const double *al;   /* pixel weights or NULL */
    double wa[3];       /* Sys::wa */
    double *z, *u;      /* at the solved scale */
    unsigned char *flat;
    int32_t *label, *members, *rstart;
    double *inv;
    int64_t nc;
    double rel, mu;
    double *gz, *hbuf, *pix, *part, *piece, *rpart;
} sys_t;

static double rad(const sys_t *S, int64_t p) {
    if (!S->al) return S->gamma;
    if (S->al[p] == 0.0) return 0.0;
    return S->gamma * S->al[p];
}

/* tv_iso.cpp:144-146 */
static void rnext(const dgeom *g, const double *wa, int64_t b, int *nxt) {
    for (int a = 0; a < g->k; ++a) nxt[a] = wa[a] != 0.0 && (a == g->k - 1 || hasnx(g, b, a));
}

static int64_t uf_find(int32_t *parent, int64_t i) {
    while (parent[i] != i) {
        parent[i] = parent[parent[i]];
        i = parent[i];
    }
    return i;
}

/* The flat test's constants: the C kernels' by default (a float64 dual:
 * the rounding floor w_max 64 2.2e-16 max|d|, tv_iso.cpp:565, and
 * DUAL_ETA); for a float32 dual the torch device path's, w_max 64 eps
 * max|x| and max(1e-9, 64 eps) with eps = 2^-23 (_tv_torch.py:192-195,
 * 404-410). or_deriv_set_eps(2.2e-16, DUAL_ETA) restores the default. */
static double g_floor_eps = 2.2e-16, g_dual_eta = DUAL_ETA;

void or_deriv_set_eps(double floor_eps, double eta) {
    g_floor_eps = floor_eps;
    g_dual_eta = eta;
}

/* TVIso::save (tv_iso.cpp:551-631). wsave: the active axis weights of
 * prox_img; d, z, u at the solved scale; u NULL for a zero dual. */
static int build_structure(sys_t *S, const double *wsave, const double *d, const double *u, double gap,
                           int from_dual) {
    const dgeom *g = &S->g;
    const int k = g->k;
    const int64_t N = g->N, n = g->n;
    double scale = 0.0, wmax = 0.0, dd = 0.0;
    for (int64_t i = 0; i < N; ++i) {
        scale = dmax(scale, fabs(d[i]));
        dd += d[i] * d[i];
    }
    for (int a = 0; a < k; ++a) wmax = dmax(wmax, wsave[a]);
    const double thr = from_dual ? wmax * 64.0 * g_floor_eps * scale : 0.0;
    S->rel = dd > 0.0 ? sqrt(2.0 * dmax(gap, 0.0) / dd) : 0.0;
    S->mu = from_dual ? CURV_RMS * wmax * sqrt(2.0 * dmax(gap, 0.0) / (double)N) : 0.0;
    int32_t *parent = S->label;
    for (int64_t i = 0; i < N; ++i) parent[i] = (int32_t)i;
    int nxt[3] = {0, 0, 0};
    for (int64_t p = 0, j = 0; p < N; ++p, j = j + 1 == n ? 0 : j + 1) {
        if (j == 0) rnext(g, wsave, p, nxt);
        double nn = 0.0;
        int any = 0;
        for (int a = 0; a < k; ++a) {
            if (!(nxt[a] && (a + 1 < k || j + 1 < n))) continue;
            const double t = wsave[a] * (S->z[p + g->S[a]] - S->z[p]);
            nn += t * t;
            any = 1;
        }
        int fl = !any || sqrt(nn) <= thr;
        if (from_dual && !fl) {
            double uu = 0.0;
            for (int a = 0; a < k; ++a) uu += S->u[a * N + p] * S->u[a * N + p];
            fl = sqrt(uu) < rad(S, p) * (1.0 - g_dual_eta);
        }
        S->flat[p] = (unsigned char)fl;
        if (!fl || !any) continue;
        for (int a = 0; a < k; ++a) {
            if (!(nxt[a] && (a + 1 < k || j + 1 < n))) continue;
            const int64_t ri = uf_find(parent, p), rj = uf_find(parent, p + g->S[a]);
            if (ri == rj) continue;
            if (ri < rj) parent[rj] = (int32_t)ri;
            else parent[ri] = (int32_t)rj;
        }
    }
    (void)u;
    for (int64_t i = 0; i < N; ++i) parent[i] = (int32_t)uf_find(parent, i);
    int64_t nc = 0;
    for (int64_t i = 0; i < N; ++i)
        if (parent[i] == i) S->inv[i] = (double)nc++;
    for (int64_t i = 0; i < N; ++i) parent[i] = (int32_t)S->inv[parent[i]];
    for (int64_t C = 0; C < nc; ++C) S->inv[C] = 0.0;
    for (int64_t i = 0; i < N; ++i) S->inv[S->label[i]] += 1.0;
    for (int64_t C = 0; C < nc; ++C) S->inv[C] = 1.0 / S->inv[C];
    S->nc = nc;
    int32_t *rs = S->rstart;
    for (int64_t C = 0; C <= nc; ++C) rs[C] = 0;
    for (int64_t i = 0; i < N; ++i) ++rs[S->label[i] + 1];
    for (int64_t C = 0; C < nc; ++C) rs[C + 1] += rs[C];
    for (int64_t i = 0; i < N; ++i) S->members[rs[S->label[i]]++] = (int32_t)i;
    for (int64_t C = nc; C > 0; --C) rs[C] = rs[C - 1];
    rs[0] = 0;
    return 0;
}

/* grad_row (tv_iso.cpp:780-800) on every row. */
static void grad(const sys_t *S, const double *x, double *y) {
    const dgeom *g = &S->g;
    const int64_t N = g->N, n = g->n;
    for (int64_t r = 0; r < g->R; ++r) {
        const int64_t b = r * n;
        for (int a = 0; a < g->k; ++a) {
            double *ya = y + a * N;
            const double wv = S->wa[a];
            if (wv == 0.0) {
                for (int64_t j = 0; j < n; ++j) ya[b + j] = 0.0;
            } else if (a == g->k - 1) {
                for (int64_t j = 0; j + 1 < n; ++j) ya[b + j] = wv * (x[b + j + 1] - x[b + j]);
                ya[b + n - 1] = 0.0;
            } else if (hasnx(g, b, a)) {
                for (int64_t j = 0; j < n; ++j) ya[b + j] = wv * (x[b + j + g->S[a]] - x[b + j]);
            } else {
                for (int64_t j = 0; j < n; ++j) ya[b + j] = 0.0;
            }
        }
    }
}

/* add_diff_row (tv_iso.cpp:130-140): x += c (h[p] - h[p - e_a]) on the row at b. */
static void add_diff(const dgeom *g, int a, double c, const double *h, double *x, int64_t b) {
    const int64_t n = g->n, Sa = g->S[a];
    if (a == g->k - 1) {
        x[b] += c * h[b];
        for (int64_t j = 1; j < n; ++j) x[b + j] += c * (h[b + j] - h[b + j - 1]);
    } else if (crd(g, b, a) > 0) {
        for (int64_t j = 0; j < n; ++j) x[b + j] += c * (h[b + j] - h[b + j - Sa]);
    } else {
        for (int64_t j = 0; j < n; ++j) x[b + j] += c * h[b + j];
    }
}

static void sub_gradT_row(const sys_t *S, const double *h, double *x, int64_t b) {
    for (int a = 0; a < S->g.k; ++a)
        if (S->wa[a] != 0.0) add_diff(&S->g, a, S->wa[a], h + a * S->g.N, x, b);
}

/* hess_row (tv_iso.cpp:810-828) */
static void hess_row(const sys_t *S, double *y, int64_t b) {
    const int k = S->g.k;
    const int64_t N = S->g.N, n = S->g.n;
    for (int64_t p = b; p < b + n; ++p) {
        if (S->flat[p]) {
            for (int a = 0; a < k; ++a) y[a * N + p] = 0.0;
            continue;
        }
        double nn = 0.0, dot = 0.0;
        for (int a = 0; a < k; ++a) nn += S->gz[a * N + p] * S->gz[a * N + p];
        const double nr = sqrt(nn);
        for (int a = 0; a < k; ++a) dot += S->gz[a * N + p] * y[a * N + p];
        const double h = rad(S, p) / dmax(nr, S->mu);
        for (int a = 0; a < k; ++a) {
            const double na = S->gz[a * N + p] / nr;
            y[a * N + p] = h * (y[a * N + p] - na * dot / nr);
        }
    }
}

/* region_sum (tv_iso.cpp:736-776): out[C] = the sum (neg: minus the sum) of
 * v over C's members, by chunks of SEG positions and pieces. */
static void region_sum(const sys_t *S, const double *v, double *out, int neg) {
    const int64_t N = S->g.N, nch = (N + SEG - 1) / SEG;
    const int32_t *rs = S->rstart, *mem = S->members, *lab = S->label;
    double *piece = S->piece;
    for (int64_t j = 0; j < nch; ++j) {
        const int64_t lo = j * SEG, hi = N < lo + SEG ? N : lo + SEG;
        for (int64_t C = lab[mem[lo]], m = lo; m < hi; ++C) {
            const int64_t end = rs[C + 1] < hi ? rs[C + 1] : hi;
            double t = 0.0;
            for (; m < end; ++m) {
                if (neg) t -= v[mem[m]];
                else t += v[mem[m]];
            }
            if (rs[C] >= lo && rs[C + 1] <= hi) out[C] = t;
            else piece[2 * j + (rs[C] < lo ? 0 : 1)] = t;
        }
    }
    double run = 0.0;
    for (int64_t j = 0; j < nch; ++j) {
        const int64_t lo = j * SEG, hi = N < lo + SEG ? N : lo + SEG;
        const int64_t C = lab[mem[lo]];
        if (rs[C] < lo) {
            run += piece[2 * j];
            if (rs[C + 1] > hi) continue;
            out[C] = run;
        }
        const int64_t Ct = lab[mem[hi - 1]];
        if (rs[Ct] >= lo && rs[Ct + 1] > hi) run = piece[2 * j + 1];
    }
}

/* region_totals (tv_iso.cpp:723-734) of the terms f adds for C, with which =
 * the expression: 0 b^2 inv; 1 r pv (r = b, pv = r dinv, written);
 * 2 pv qv; 3 the x, r update with (r^2 dinv, r^2 inv). */
typedef struct {
    double *b, *x, *r, *pv, *qv, *dinv;
    double alpha;
} cgv;

static void region_totals(const sys_t *S, int which, cgv *c, double *s0, double *s1) {
    const int64_t nb = (S->nc + REGBLK - 1) / REGBLK;
    for (int64_t blk = 0; blk < nb; ++blk) {
        double t0 = 0.0, t1 = 0.0;
        const int64_t hi = S->nc < (blk + 1) * REGBLK ? S->nc : (blk + 1) * REGBLK;
        for (int64_t C = blk * REGBLK; C < hi; ++C) {
            switch (which) {
                case 0: t0 += c->b[C] * c->b[C] * S->inv[C]; break;
                case 1:
                    c->r[C] = c->b[C];
                    c->pv[C] = c->r[C] * c->dinv[C];
                    t0 += c->r[C] * c->pv[C];
                    break;
                case 2: t0 += c->pv[C] * c->qv[C]; break;
                default:
                    c->x[C] += c->alpha * c->pv[C];
                    c->r[C] -= c->alpha * c->qv[C];
                    t0 += c->r[C] * c->r[C] * c->dinv[C];
                    t1 += c->r[C] * c->r[C] * S->inv[C];
            }
        }
        S->rpart[2 * blk] = t0;
        S->rpart[2 * blk + 1] = t1;
    }
    *s0 = *s1 = 0.0;
    for (int64_t blk = 0; blk < nb; ++blk) {
        *s0 += S->rpart[2 * blk];
        *s1 += S->rpart[2 * blk + 1];
    }
}

/* jacobi (tv_iso.cpp:833-870) */
static void jacobi(const sys_t *S, double *dinv) {
    const dgeom *g = &S->g;
    const int k = g->k;
    const int64_t N = g->N, n = g->n;
    for (int64_t C = 0; C < S->nc; ++C) dinv[C] = 1.0 / S->inv[C];
    int nxt[3] = {0, 0, 0};
    for (int64_t p = 0, j = 0; p < N; ++p, j = j + 1 == n ? 0 : j + 1) {
        if (j == 0) rnext(g, S->wa, p, nxt);
        if (S->flat[p]) continue;
        double nn = 0.0;
        for (int a = 0; a < k; ++a) nn += S->gz[a * N + p] * S->gz[a * N + p];
        const double nr = sqrt(nn), hs = rad(S, p) / dmax(nr, S->mu);
        int64_t lab[4];
        int has[3];
        lab[0] = S->label[p];
        for (int a = 0; a < k; ++a) {
            has[a] = nxt[a] && (a + 1 < k || j + 1 < n);
            lab[a + 1] = has[a] ? S->label[p + g->S[a]] : -1;
        }
        for (int c = 0; c <= k; ++c) {
            const int64_t C = lab[c];
            if (C < 0) continue;
            int seen = 0;
            for (int e = 0; e < c; ++e) seen = seen || lab[e] == C;
            if (seen) continue;
            double hh = 0.0, nh = 0.0;
            for (int a = 0; a < k; ++a) {
                if (!has[a]) continue;
                const double h = S->wa[a] * ((lab[a + 1] == C ? 1.0 : 0.0) - (lab[0] == C ? 1.0 : 0.0));
                hh += h * h;
                nh += h * S->gz[a * N + p] / nr;
            }
            dinv[C] += hs * (hh - nh * nh);
        }
    }
    for (int64_t C = 0; C < S->nc; ++C) dinv[C] = 1.0 / dinv[C];
}

/* apply (tv_iso.cpp:873-888): out = M^T (I + G^T H G) M cv. */
static void apply(const sys_t *S, const double *cv, double *out) {
    const dgeom *g = &S->g;
    const int64_t n = g->n;
    for (int64_t p = 0; p < g->N; ++p) S->pix[p] = cv[S->label[p]];
    grad(S, S->pix, S->hbuf);
    for (int64_t r = 0; r < g->R; ++r) hess_row(S, S->hbuf, r * n);
    for (int64_t r = 0; r < g->R; ++r) {
        const int64_t b = r * n;
        for (int64_t p = b; p < b + n; ++p) S->pix[p] = -S->pix[p];
        sub_gradT_row(S, S->hbuf, S->pix, b);
    }
    region_sum(S, S->pix, out, 1);
}

/* cg (tv_iso.cpp:899-943) */
static void cg(const sys_t *S, double *b, double *x, double *r, double *pv, double *qv, double *dinv, double tol_in,
               int max_iter, double *residual, int *iters) {
    jacobi(S, dinv);
    cgv c = {b, x, r, pv, qv, dinv, 0.0};
    double bn, rz, unused;
    region_totals(S, 0, &c, &bn, &unused);
    for (int64_t C = 0; C < S->nc; ++C) x[C] = 0.0;
    if (bn == 0.0) {
        *residual = 0.0;
        *iters = 0;
        return;
    }
    region_totals(S, 1, &c, &rz, &unused);
    const double tol = tol_in > 0.0 ? tol_in : dmax(DERIV_FLOOR, S->rel);
    int it = 0;
    double res = 1.0;
    while (res > tol && it < max_iter) {
        apply(S, pv, qv);
        double pq;
        region_totals(S, 2, &c, &pq, &unused);
        if (!(pq > 0.0)) break;
        c.alpha = rz / pq;
        double rzn, rn;
        region_totals(S, 3, &c, &rzn, &rn);
        const double beta = rzn / rz;
        rz = rzn;
        for (int64_t C = 0; C < S->nc; ++C) pv[C] = r[C] * dinv[C] + beta * pv[C];
        ++it;
        res = sqrt(rn / bn);
    }
    *residual = res;
    *iters = it;
}

/* row_total (tv_iso.cpp:700-713): blocks of max(1, CHUNK / n) rows. */
typedef double (*row_fn)(const sys_t *, int64_t b, const void *ctx, double t);
static double row_total(const sys_t *S, row_fn f, const void *ctx) {
    const int64_t rb = CHUNK / S->g.n > 1 ? CHUNK / S->g.n : 1;
    const int64_t nb = (S->g.R + rb - 1) / rb;
    double tot = 0.0;
    for (int64_t blk = 0; blk < nb; ++blk) {
        double t = 0.0;
        for (int64_t r = blk * rb; r < S->g.R && r < (blk + 1) * rb; ++r) t = f(S, r * S->g.n, ctx, t);
        tot += t;
    }
    return tot;
}

typedef struct {
    const double *v, *ua, *gv;
    int a;
    double *acc;
    double sc;
    int64_t k;
} rctx;

static double row_du(const sys_t *S, int64_t b, const void *cp, double sum) {
    const rctx *c = (const rctx *)cp;
    const dgeom *g = &S->g;
    const int64_t n = g->n;
    if (c->a == g->k - 1) {
        for (int64_t j = 0; j + 1 < n; ++j) sum += (c->v[b + j + 1] - c->v[b + j]) * c->ua[b + j];
    } else if (hasnx(g, b, c->a)) {
        for (int64_t j = 0; j < n; ++j) sum += (c->v[b + j + g->S[c->a]] - c->v[b + j]) * c->ua[b + j];
    }
    return sum;
}

static double row_gam(const sys_t *S, int64_t b, const void *cp, double sum) {
    const rctx *c = (const rctx *)cp;
    const int k = S->g.k;
    const int64_t N = S->g.N;
    for (int64_t p = b; p < b + S->g.n; ++p) {
        if (S->flat[p]) continue;
        double nn = 0.0;
        for (int a = 0; a < k; ++a) nn += S->gz[a * N + p] * S->gz[a * N + p];
        const double nr = sqrt(nn);
        double t = 0.0;
        for (int a = 0; a < k; ++a) t += c->gv[a * N + p] * S->gz[a * N + p] / nr;
        if (S->al) {
            sum += S->al[p] * t;
            c->acc[k + p] -= c->sc * (S->gamma * t);
        } else {
            sum += t;
        }
    }
    return sum;
}

static double row_hw(const sys_t *S, int64_t b, const void *cp, double sum) {
    const rctx *c = (const rctx *)cp;
    const int64_t N = S->g.N;
    const int a = c->a;
    for (int64_t p = b; p < b + S->g.n; ++p)
        if (!S->flat[p]) sum += c->gv[a * N + p] * (S->gz[a * N + p] / S->wa[a]);
    return sum;
}

/*
 * One image's derivative at a solution: x (the input), z and u (the output
 * and dual of the solve, at the input's scale), gap (the solve's gap, at the
 * input's scale) and e (its scale exponent, tv_iso.cpp:367-377), as the
 * reference's prox_img leaves them for save (tv_iso.cpp:402); gamma, w and
 * alpha are the operator's and the step's.
 *   mode 0 (vjp): in = g, out = J^T g; acc (nparams + 1 doubles) gets the
 *     parameter gradients subtracted in (acc[k + p] for pixel weights,
 *     acc[nparams] for gamma), as vjp_img does;
 *   mode 1 (jvp): in = dx, out = J dx + the dgamma and dparams terms.
 * deriv_tol 0: the default tolerance. Returns 0, or 2 when out of memory.
 */
int or_iso_deriv(int k, const int64_t *shape, const double *w, const double *alpha, double gamma, const double *x,
                 const double *z_in, const double *u_in, double gap, int e, int mode, const double *in,
                 double dgamma, const double *dparams, double deriv_tol, int deriv_max_iter, double *out,
                 double *acc, double *residual, int *iters, int64_t *nregions) {
    sys_t S;
    memset(&S, 0, sizeof S);
    S.g = mk(k, shape);
    const int64_t N = S.g.N;
    const int64_t nch = (N + SEG - 1) / SEG, nrb = (N + REGBLK - 1) / REGBLK;
    double *mem = (double *)calloc((size_t)((4 + 3 * k + 8) * N + 2 * nch + 2 * nrb + 8), sizeof(double));
    int32_t *imem = (int32_t *)calloc((size_t)(3 * N + 2), sizeof(int32_t));
    unsigned char *fl = (unsigned char *)calloc((size_t)N + 1, 1);
    if (!mem || !imem || !fl) {
        free(mem);
        free(imem);
        free(fl);
        return 2;
    }
    double *d = mem, *z = d + N, *u = z + N;
    S.inv = u + k * N;
    S.gz = S.inv + N;
    S.hbuf = S.gz + k * N;
    S.pix = S.hbuf + k * N;
    double *cb = S.pix + N, *xr = cb + N, *rr = xr + N, *pv = rr + N, *qv = pv + N, *dg = qv + N, *rhs = dg + N;
    S.piece = rhs + N;
    S.rpart = S.piece + 2 * nch;
    S.label = imem;
    S.members = imem + N;
    S.rstart = imem + 2 * N;
    S.flat = fl;
    S.al = alpha;
    S.z = z;
    S.u = u;

    /* prox_img's view (tv_iso.cpp:342-402): active weights, NaN images, the
     * solved scale. */
    double wmax = 0.0, wsave[3] = {0.0, 0.0, 0.0};
    for (int a = 0; a < k; ++a)
        if (S.g.L[a] > 1) wmax = dmax(wmax, w[a]);
    int na = 0, nan = 0;
    for (int a = 0; a < k; ++a)
        if (gamma > 0.0 && w[a] > 0.0 && S.g.L[a] > 1 && w[a] >= 1e-100 * wmax) {
            wsave[a] = w[a];
            ++na;
        }
    for (int64_t i = 0; i < N; ++i) nan = nan || isnan(x[i]);
    for (int64_t i = 0; i < N; ++i) {
        d[i] = ldexp(x[i], -e);
        z[i] = nan ? d[i] : ldexp(z_in[i], -e);
    }
    const int use_dual = !nan && na >= 1;
    for (int64_t i = 0; i < k * N; ++i) u[i] = use_dual ? ldexp(u_in[i], -e) : 0.0;
    S.gamma = ldexp(gamma, -e);
    build_structure(&S, wsave, d, u, nan ? 0.0 : ldexp(gap, -2 * e), !nan && na >= 2);
    for (int a = 0; a < k; ++a) S.wa[a] = (S.gamma > 0.0 && w[a] > 0.0 && S.g.L[a] > 1) ? w[a] : 0.0;
    grad(&S, S.z, S.gz);
    const double sc = ldexp(1.0, e);
    const int64_t n = S.g.n;
    if (mode == 1) {
        double *cA = S.hbuf;
        for (int64_t r = 0; r < S.g.R; ++r) {
            const int64_t b = r * n;
            for (int64_t p = b; p < b + n; ++p) rhs[p] = in ? in[p] : 0.0;
            if (dparams)
                for (int a = 0; a < k; ++a)
                    if (dparams[a] != 0.0 && S.g.L[a] > 1) add_diff(&S.g, a, sc * dparams[a], u + a * N, rhs, b);
            for (int a = 0; a < k; ++a) {
                double *ca = cA + a * N;
                for (int64_t p = b; p < b + n; ++p) {
                    const double wd = S.wa[a] != 0.0 ? S.gz[a * N + p] / S.wa[a] : 0.0;
                    ca[p] = (dparams && S.wa[a] != 0.0) ? sc * dparams[a] * wd : 0.0;
                }
            }
            hess_row(&S, cA, b);
            for (int64_t p = b; p < b + n; ++p) {
                if (S.flat[p]) continue;
                double nn = 0.0;
                for (int a = 0; a < k; ++a) nn += S.gz[a * N + p] * S.gz[a * N + p];
                const double nr = sqrt(nn);
                double dr = dgamma;
                if (S.al) dr = dgamma * S.al[p] + (dparams ? S.gamma * (sc * dparams[k + p]) : 0.0);
                for (int a = 0; a < k; ++a) cA[a * N + p] += dr * S.gz[a * N + p] / nr;
            }
        }
        for (int64_t r = 0; r < S.g.R; ++r) sub_gradT_row(&S, cA, rhs, r * n);
        region_sum(&S, rhs, cb, 0);
        cg(&S, cb, xr, rr, pv, qv, dg, deriv_tol, deriv_max_iter, residual, iters);
        for (int64_t p = 0; p < N; ++p) out[p] = xr[S.label[p]];
    } else {
        for (int64_t p = 0; p < N; ++p) S.pix[p] = in[p];
        region_sum(&S, S.pix, cb, 0);
        cg(&S, cb, xr, rr, pv, qv, dg, deriv_tol, deriv_max_iter, residual, iters);
        double *vpix = S.pix;
        for (int64_t p = 0; p < N; ++p) {
            vpix[p] = xr[S.label[p]];
            out[p] = vpix[p];
        }
        const int64_t P = k + (alpha ? N : 0);
        double *gv = S.hbuf;
        grad(&S, vpix, gv);
        rctx c = {vpix, NULL, gv, 0, acc, sc, k};
        for (int a = 0; a < k; ++a) {
            if (S.g.L[a] == 1) continue;
            c.a = a;
            c.ua = u + a * N;
            acc[a] -= sc * row_total(&S, row_du, &c);
        }
        const double ggam = row_total(&S, row_gam, &c);
        acc[P] -= ggam;
        for (int64_t r = 0; r < S.g.R; ++r) hess_row(&S, gv, r * n);
        for (int a = 0; a < k; ++a) {
            if (S.wa[a] == 0.0) continue;
            c.a = a;
            acc[a] -= sc * row_total(&S, row_hw, &c);
        }
    }
    if (nregions) *nregions = S.nc;
    free(mem);
    free(imem);
    free(fl);
    return 0;
}

/* ---------------------------------------------------------------------------
 * Anisotropic TV, the device method's derivative (_tv_torch.py:346-362,
 * :528-585): an edge (p, p + e_a) is flat where |u_a[p]| < r (1 - eta), or
 * where |g| <= w_max 64 eps max|x| and r > 0 (r = gamma beta_a[p], or gamma);
 * flat edges of axes with w_a != 0 join their pixels; on every other edge n =
 * sign(u) (sign of the jump where beta is 0). dz is the region mean of
 * r = dx - sum_a dw_a D_a^T u_a - G^T (dr n); the VJP its transpose:
 *   gx = region means of g; ggamma = -sum_e r_e/gamma (G h)_e n_e;
 *   gw_a = -<D_a h, u_a>; gbeta_e = -gamma (G h)_e n_e.
 * u is the per-edge dual (k, *img) at the input's scale; eps the dtype's.
 * Regions by union-find (roots at the smallest index); sums in index order.
 * ------------------------------------------------------------------------- */
static int64_t aniso_regions(const dgeom *g, const double *w, const double *beta, double gamma, const double *x,
                             const double *z, const double *u, double eta, double eps, int32_t *label,
                             signed char *nsg, double *inv) {
    const int k = g->k;
    const int64_t N = g->N;
    double xm = 0.0, wmax = 0.0;
    for (int64_t i = 0; i < N; ++i) xm = dmax(xm, fabs(x[i]));
    for (int a = 0; a < k; ++a) wmax = dmax(wmax, w[a]);
    const double floor = wmax * 64.0 * eps * xm;
    for (int64_t i = 0; i < N; ++i) label[i] = (int32_t)i;
    for (int a = 0; a < k; ++a)
        for (int64_t p = 0; p < N; ++p) {
            const int has = hasnx(g, p, a);
            const double r = beta ? gamma * beta[a * N + p] : gamma;
            const double inside = r * (1.0 - eta);
            const double gj = has ? w[a] * (z[p + g->S[a]] - z[p]) : 0.0;
            const double ua = u[a * N + p];
            const int fl = fabs(ua) < inside || (fabs(gj) <= floor && inside > 0.0);
            double s = ua > 0.0 ? 1.0 : (ua < 0.0 ? -1.0 : 0.0);
            if (beta && beta[a * N + p] == 0.0) s = gj > 0.0 ? 1.0 : (gj < 0.0 ? -1.0 : 0.0);
            nsg[a * N + p] = (signed char)(fl ? 0 : (int)s);
            if (fl && has && w[a] != 0.0) {
                const int64_t ri = uf_find(label, p), rj = uf_find(label, p + g->S[a]);
                if (ri < rj) label[rj] = (int32_t)ri;
                else if (rj < ri) label[ri] = (int32_t)rj;
            }
        }
    for (int64_t i = 0; i < N; ++i) label[i] = (int32_t)uf_find(label, i);
    int64_t nc = 0;
    for (int64_t i = 0; i < N; ++i)
        if (label[i] == i) inv[i] = (double)nc++;
    for (int64_t i = 0; i < N; ++i) label[i] = (int32_t)inv[label[i]];
    for (int64_t C = 0; C < nc; ++C) inv[C] = 0.0;
    for (int64_t i = 0; i < N; ++i) inv[label[i]] += 1.0;
    for (int64_t C = 0; C < nc; ++C) inv[C] = 1.0 / inv[C];
    return nc;
}

/* mode 0 (vjp): in = g -> out = gx, acc[k + 1 (+ k N)]: gw_a, ggamma at
 * acc[k], gbeta at acc[k + 1 + a N + p] (written, not accumulated).
 * mode 1 (jvp): in = dx (or NULL), dgamma, dw (k or NULL), dbeta (k N or
 * NULL) -> out = dz. */
int or_aniso_deriv(int k, const int64_t *shape, const double *w, const double *beta, double gamma,
                   const double *x, const double *z, const double *u, double eta, double eps, int mode,
                   const double *in, double dgamma, const double *dw, const double *dbeta, double *out, double *acc,
                   int64_t *nregions) {
    const dgeom g = mk(k, shape);
    const int64_t N = g.N;
    int32_t *label = (int32_t *)malloc((size_t)N * sizeof(int32_t));
    signed char *nsg = (signed char *)malloc((size_t)(k * N));
    double *inv = (double *)malloc((size_t)N * sizeof(double));
    double *cs = (double *)calloc((size_t)N, sizeof(double));
    double *r = (double *)calloc((size_t)N, sizeof(double));
    if (!label || !nsg || !inv || !cs || !r) {
        free(label);
        free(nsg);
        free(inv);
        free(cs);
        free(r);
        return 2;
    }
    const int64_t nc = aniso_regions(&g, w, beta, gamma, x, z, u, eta, eps, label, nsg, inv);
    if (mode == 1) {
        for (int64_t p = 0; p < N; ++p) r[p] = in ? in[p] : 0.0;
        for (int a = 0; a < k; ++a)
            for (int64_t p = 0; p < N; ++p) {
                if (!hasnx(&g, p, a)) continue;
                const int64_t q = p + g.S[a];
                /* - dw_a D_a^T u_a: (D_a^T v)_p = v[p - e_a] - v[p] */
                if (dw && dw[a] != 0.0) {
                    r[p] += dw[a] * u[a * N + p];
                    r[q] -= dw[a] * u[a * N + p];
                }
                /* - G^T (dr n): (G^T c)_p = w_a (c[p - e_a] - c[p]) */
                double dr = 0.0;
                if (beta) dr = dgamma * beta[a * N + p] + (dbeta ? gamma * dbeta[a * N + p] : 0.0);
                else dr = dgamma;
                const double c = dr * (double)nsg[a * N + p];
                if (w[a] != 0.0 && c != 0.0) {
                    r[p] += w[a] * c;
                    r[q] -= w[a] * c;
                }
            }
        for (int64_t p = 0; p < N; ++p) cs[label[p]] += r[p];
        for (int64_t p = 0; p < N; ++p) out[p] = cs[label[p]] * inv[label[p]];
    } else {
        for (int64_t p = 0; p < N; ++p) cs[label[p]] += in[p];
        for (int64_t p = 0; p < N; ++p) out[p] = cs[label[p]] * inv[label[p]];
        double gg = 0.0;
        for (int a = 0; a < k; ++a) {
            double gw = 0.0;
            for (int64_t p = 0; p < N; ++p) {
                const int has = hasnx(&g, p, a);
                const double dh = has ? out[p + g.S[a]] - out[p] : 0.0;
                gw += dh * u[a * N + p];
                const double t = w[a] * dh * (double)nsg[a * N + p];
                const double sw = beta ? beta[a * N + p] : 1.0;
                gg += sw * t;
                if (beta) acc[k + 1 + a * N + p] = -(gamma * t);
            }
            acc[a] = -gw;
        }
        acc[k] = -gg;
    }
    if (nregions) *nregions = nc;
    free(label);
    free(nsg);
    free(inv);
    free(cs);
    free(r);
    return 0;
}
