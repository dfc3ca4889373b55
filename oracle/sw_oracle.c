/*
 * ORACLE — test infrastructure only, never part of the product.
 *
 * Plain-C restatement of the reference CPU algorithm for soft Smith-Waterman
 * (linear gap and affine gap): forward partition, backward marginals and
 * scalar-parameter gradients, Hessian-vector product, and parameter
 * sensitivities.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker.
 *
 * The file is compiled twice by oracle/Makefile: REAL=float gives the FP32
 * restatement (bit-for-bit the same arithmetic order as the reference, so it
 * can be pinned against oracle/_ref and the committed golden vectors) and
 * REAL=double gives the FP64 oracle used for error reports.
 *
 * Every function follows one reference function; the citation is on it.
 * Layouts follow the reference: alpha/beta tables are [B, S*(L1+1)*(L2+1)]
 * (S = 1 linear, S = 3 affine with M|I|D planes), scores are [B, L1, L2]
 * row-major, lengths are int32 [B, 2].
 *
 * Numerics contract restated from src/common/numerics.h:29-127:
 *   NINF = -1e30, safe_exp(x) = 0 for x < -88 else exp(min(x, 88)),
 *   Kahan-compensated sums, logsumexp with max shift returning NINF when the
 *   max is <= NINF.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef REAL
#define REAL float
#endif

#if REAL_IS_DOUBLE
#define OEXP exp
#define OLOG log
#define OFN(name) oracle_f64_##name
#else
#define OEXP expf
#define OLOG logf
#define OFN(name) oracle_f32_##name
#endif

typedef REAL real;

static const real NINF_ = (real)-1e30;

/* numerics.h:34-42 */
static real sexp(real x) {
    if (x < (real)-88.0) return (real)0;
    if (x > (real)88.0) x = (real)88.0;
    return OEXP(x);
}

/* numerics.h:49-68: Kahan accumulator */
typedef struct { real s, c; } kahan;
static void kadd(kahan* k, real v) {
    real y = v - k->c;
    real t = k->s + y;
    k->c = (t - k->s) - y;
    k->s = t;
}

/* std::max semantics: (a < b) ? b : a, folded left to right like std::max({...}) */
static real rmax(real a, real b) { return (a < b) ? b : a; }

/* numerics.h:75-127: temperature-scaled logsumexp of n terms, Kahan sum of
 * the shifted exponentials in argument order. */
static real lse_n(const real* x, int n, real T) {
    real m = x[0];
    for (int k = 1; k < n; ++k) m = rmax(m, x[k]);
    if (m <= NINF_) return NINF_;
    kahan acc = {0, 0};
    for (int k = 0; k < n; ++k) kadd(&acc, sexp((x[k] - m) / T));
    return m + T * OLOG(acc.s);
}

/* src/sw_affine/kernels_cpu.cpp:56-88 and src/sw/kernels_cpu.cpp:43-64:
 * normalised soft-max weights; entries at NINF get weight 0, the others are
 * divided by their Kahan sum when it is positive. */
static void wsoft(const real* x, int n, real T, real* w) {
    real m = x[0];
    for (int k = 1; k < n; ++k) m = rmax(m, x[k]);
    if (m <= NINF_) {
        for (int k = 0; k < n; ++k) w[k] = 0;
        return;
    }
    kahan acc = {0, 0};
    for (int k = 0; k < n; ++k) {
        if (x[k] > NINF_) { w[k] = sexp((x[k] - m) / T); kadd(&acc, w[k]); }
        else w[k] = 0;
    }
    if (acc.s > 0)
        for (int k = 0; k < n; ++k) w[k] /= acc.s;
}

/* clamp_log_posterior + safe_exp, kernels_cpu.cpp:43-52 (affine) */
static real post_w(real a, real S, real T) {
    real lp = (a - S) / T;
    if (lp > 0) lp = 0;
    return sexp(lp);
}

static int is_bad(real v) { return isinf(v) || isnan(v) || v == (real)0; }

/* ------------------------------------------------------------------------ */
/* Affine gap, three states M, I, D.                                         */
/* ------------------------------------------------------------------------ */

typedef struct {
    size_t cs;     /* cells per plane, (L1max+1)*(L2max+1) */
    size_t cols;   /* L2max + 1 */
    size_t ss;     /* score stride per pair, L1max*L2max */
    int ld;        /* score row stride, L2max */
} geom;

static geom mk_geom(int L1, int L2) {
    geom g;
    g.cols = (size_t)L2 + 1;
    g.cs = ((size_t)L1 + 1) * g.cols;
    g.ss = (size_t)L1 * (size_t)L2;
    g.ld = L2;
    return g;
}

#define CELL(g, i, j) ((size_t)(i) * (g).cols + (size_t)(j))
#define SIDX(g, i, j) ((size_t)((i) - 1) * (size_t)(g).ld + (size_t)((j) - 1))

/* Candidates of the three recurrences at cell (i, j), kernels_cpu.cpp:140-161 */
static void aff_cands(const real* a, geom g, int i, int j, real sc, real go, real ge,
                      real* cm, real* ci, real* cd) {
    size_t dg = CELL(g, i - 1, j - 1), up = CELL(g, i - 1, j), lf = CELL(g, i, j - 1);
    int in = (i > 1 && j > 1);
    cm[0] = in ? a[0 * g.cs + dg] + sc : NINF_;
    cm[1] = in ? a[1 * g.cs + dg] + sc : NINF_;
    cm[2] = in ? a[2 * g.cs + dg] + sc : NINF_;
    cm[3] = sc;
    ci[0] = (i > 1) ? a[0 * g.cs + up] + go : NINF_;
    ci[1] = (i > 1) ? a[1 * g.cs + up] + ge : NINF_;
    cd[0] = (j > 1) ? a[0 * g.cs + lf] + go : NINF_;
    cd[1] = (j > 1) ? a[2 * g.cs + lf] + ge : NINF_;
}

/* sw_affine_forward_cpu, src/sw_affine/kernels_cpu.cpp:94-195 */
void OFN(aff_forward)(const real* scores, real* alpha, real* part, const int* len,
                      int B, int L1m, int L2m, real go, real ge, real T) {
    geom g = mk_geom(L1m, L2m);
    for (int b = 0; b < B; ++b) {
        const real* s = scores + (size_t)b * g.ss;
        real* a = alpha + (size_t)b * 3 * g.cs;
        int L1 = len[2 * b], L2 = len[2 * b + 1];
        for (size_t t = 0; t < 3 * g.cs; ++t) a[t] = NINF_;
        a[0] = 0;
        for (int k = 2; k <= L1 + L2; ++k) {
            int lo = k - L2 > 1 ? k - L2 : 1, hi = k - 1 < L1 ? k - 1 : L1;
            for (int i = lo; i <= hi; ++i) {
                int j = k - i;
                real cm[4], ci[2], cd[2];
                aff_cands(a, g, i, j, s[SIDX(g, i, j)], go, ge, cm, ci, cd);
                size_t c = CELL(g, i, j);
                a[0 * g.cs + c] = lse_n(cm, 4, T);
                a[1 * g.cs + c] = lse_n(ci, 2, T);
                a[2 * g.cs + c] = lse_n(cd, 2, T);
            }
        }
        real m = NINF_;
        for (int i = 0; i <= L1; ++i)
            for (int j = 0; j <= L2; ++j)
                for (int st = 0; st < 3; ++st) m = rmax(m, a[st * g.cs + CELL(g, i, j)]);
        kahan z = {0, 0};
        for (int i = 0; i <= L1; ++i)
            for (int j = 0; j <= L2; ++j)
                for (int st = 0; st < 3; ++st) {
                    real v = a[st * g.cs + CELL(g, i, j)];
                    if (v > NINF_) kadd(&z, sexp((v - m) / T));
                }
        part[b] = m + T * OLOG(z.s);
    }
}

/* sw_affine_backward_cpu, src/sw_affine/kernels_cpu.cpp:201-356 */
void OFN(aff_backward)(const real* alpha, const real* scores, const real* part, real* beta,
                       real* post, real* gopen, real* gext, real* gT, const int* len,
                       int B, int L1m, int L2m, real go, real ge, real T) {
    geom g = mk_geom(L1m, L2m);
    for (int b = 0; b < B; ++b) {
        const real* a = alpha + (size_t)b * 3 * g.cs;
        const real* s = scores + (size_t)b * g.ss;
        real* be = beta + (size_t)b * 3 * g.cs;
        real* P = post + (size_t)b * g.ss;
        real S = part[b];
        int L1 = len[2 * b], L2 = len[2 * b + 1];
        memset(P, 0, g.ss * sizeof(real));
        memset(be, 0, 3 * g.cs * sizeof(real));
        for (int i = 0; i <= L1; ++i)
            for (int j = 0; j <= L2; ++j)
                for (int st = 0; st < 3; ++st) {
                    size_t x = st * g.cs + CELL(g, i, j);
                    be[x] = post_w(a[x], S, T);
                }
        real so = 0, se = 0;
        for (int k = L1 + L2; k >= 2; --k) {
            int lo = k - L2 > 1 ? k - L2 : 1, hi = k - 1 < L1 ? k - 1 : L1;
            for (int i = lo; i <= hi; ++i) {
                int j = k - i;
                size_t c = CELL(g, i, j), dg = CELL(g, i - 1, j - 1);
                size_t up = CELL(g, i - 1, j), lf = CELL(g, i, j - 1);
                real cm[4], ci[2], cd[2], w[4];
                aff_cands(a, g, i, j, s[SIDX(g, i, j)], go, ge, cm, ci, cd);
                real bm = be[0 * g.cs + c];
                if (!is_bad(bm)) {
                    wsoft(cm, 4, T, w);
                    P[SIDX(g, i, j)] += bm * (w[0] + w[1] + w[2] + w[3]);
                    if (i > 1 && j > 1) {
                        if (w[0] > 0) be[0 * g.cs + dg] += bm * w[0];
                        if (w[1] > 0) be[1 * g.cs + dg] += bm * w[1];
                        if (w[2] > 0) be[2 * g.cs + dg] += bm * w[2];
                    }
                }
                real bi = be[1 * g.cs + c];
                if (!is_bad(bi)) {
                    wsoft(ci, 2, T, w);
                    if (i > 1) {
                        if (w[0] > 0) { be[0 * g.cs + up] += bi * w[0]; so += bi * w[0]; }
                        if (w[1] > 0) { be[1 * g.cs + up] += bi * w[1]; se += bi * w[1]; }
                    }
                }
                real bd = be[2 * g.cs + c];
                if (!is_bad(bd)) {
                    wsoft(cd, 2, T, w);
                    if (j > 1) {
                        if (w[0] > 0) { be[0 * g.cs + lf] += bd * w[0]; so += bd * w[0]; }
                        if (w[1] > 0) { be[2 * g.cs + lf] += bd * w[1]; se += bd * w[1]; }
                    }
                }
            }
        }
        kahan ms = {0, 0};
        for (size_t x = 0; x < g.ss; ++x) kadd(&ms, P[x] * s[x]);
        real gap_cost = so * go + se * ge;
        real tot = ms.s + gap_cost;
        gopen[b] = so;
        gext[b] = se;
        gT[b] = (S - tot) / T;
    }
}

/* sw_affine_hvp_cpu, src/sw_affine/kernels_cpu.cpp:362-631 */
void OFN(aff_hvp)(const real* alpha, const real* scores, const real* part, const real* V,
                  real* dalpha, real* dpart, real* dbeta, real* H, const int* len,
                  int B, int L1m, int L2m, real go, real ge, real T) {
    geom g = mk_geom(L1m, L2m);
    real* be = (real*)malloc(3 * g.cs * sizeof(real));
    for (int b = 0; b < B; ++b) {
        const real* a = alpha + (size_t)b * 3 * g.cs;
        const real* s = scores + (size_t)b * g.ss;
        const real* v = V + (size_t)b * g.ss;
        real* da = dalpha + (size_t)b * 3 * g.cs;
        real* db = dbeta + (size_t)b * 3 * g.cs;
        real* h = H + (size_t)b * g.ss;
        real S = part[b];
        int L1 = len[2 * b], L2 = len[2 * b + 1];
        memset(da, 0, 3 * g.cs * sizeof(real));
        memset(db, 0, 3 * g.cs * sizeof(real));
        memset(h, 0, g.ss * sizeof(real));
        /* tangent forward */
        for (int k = 2; k <= L1 + L2; ++k) {
            int lo = k - L2 > 1 ? k - L2 : 1, hi = k - 1 < L1 ? k - 1 : L1;
            for (int i = lo; i <= hi; ++i) {
                int j = k - i;
                size_t c = CELL(g, i, j), dg = CELL(g, i - 1, j - 1);
                size_t up = CELL(g, i - 1, j), lf = CELL(g, i, j - 1);
                real vij = v[SIDX(g, i, j)];
                real cm[4], ci[2], cd[2], wm[4], wi[2], wd[2];
                aff_cands(a, g, i, j, s[SIDX(g, i, j)], go, ge, cm, ci, cd);
                int in = (i > 1 && j > 1);
                wsoft(cm, 4, T, wm);
                real t0 = in ? da[0 * g.cs + dg] + vij : 0;
                real t1 = in ? da[1 * g.cs + dg] + vij : 0;
                real t2 = in ? da[2 * g.cs + dg] + vij : 0;
                da[0 * g.cs + c] = wm[0] * t0 + wm[1] * t1 + wm[2] * t2 + wm[3] * vij;
                wsoft(ci, 2, T, wi);
                real u0 = (i > 1) ? da[0 * g.cs + up] : 0;
                real u1 = (i > 1) ? da[1 * g.cs + up] : 0;
                da[1 * g.cs + c] = wi[0] * u0 + wi[1] * u1;
                wsoft(cd, 2, T, wd);
                real l0 = (j > 1) ? da[0 * g.cs + lf] : 0;
                real l2 = (j > 1) ? da[2 * g.cs + lf] : 0;
                da[2 * g.cs + c] = wd[0] * l0 + wd[1] * l2;
            }
        }
        kahan ds = {0, 0};
        for (int i = 0; i <= L1; ++i)
            for (int j = 0; j <= L2; ++j)
                for (int st = 0; st < 3; ++st) {
                    size_t x = st * g.cs + CELL(g, i, j);
                    kadd(&ds, post_w(a[x], S, T) * da[x]);
                }
        real dS = ds.s;
        dpart[b] = dS;
        memset(be, 0, 3 * g.cs * sizeof(real));
        for (int i = 0; i <= L1; ++i)
            for (int j = 0; j <= L2; ++j)
                for (int st = 0; st < 3; ++st) {
                    size_t x = st * g.cs + CELL(g, i, j);
                    be[x] = post_w(a[x], S, T);
                }
        for (int i = 0; i <= L1; ++i)
            for (int j = 0; j <= L2; ++j)
                for (int st = 0; st < 3; ++st) {
                    size_t x = st * g.cs + CELL(g, i, j);
                    db[x] = be[x] * (da[x] - dS) / T;
                }
        /* tangent backward */
        for (int k = L1 + L2; k >= 2; --k) {
            int lo = k - L2 > 1 ? k - L2 : 1, hi = k - 1 < L1 ? k - 1 : L1;
            for (int i = lo; i <= hi; ++i) {
                int j = k - i;
                size_t c = CELL(g, i, j), dg = CELL(g, i - 1, j - 1);
                size_t up = CELL(g, i - 1, j), lf = CELL(g, i, j - 1);
                real vij = v[SIDX(g, i, j)];
                real cm[4], ci[2], cd[2], w[4];
                aff_cands(a, g, i, j, s[SIDX(g, i, j)], go, ge, cm, ci, cd);
                int in = (i > 1 && j > 1);
                /* M */
                real bm = be[0 * g.cs + c], dbm = db[0 * g.cs + c];
                wsoft(cm, 4, T, w);
                real t0 = in ? da[0 * g.cs + dg] + vij : 0;
                real t1 = in ? da[1 * g.cs + dg] + vij : 0;
                real t2 = in ? da[2 * g.cs + dg] + vij : 0;
                real t3 = vij;
                real mean = w[0] * t0 + w[1] * t1 + w[2] * t2 + w[3] * t3;
                real dw0 = w[0] * (t0 - mean) / T, dw1 = w[1] * (t1 - mean) / T;
                real dw2 = w[2] * (t2 - mean) / T, dw3 = w[3] * (t3 - mean) / T;
                h[SIDX(g, i, j)] += dbm * (w[0] + w[1] + w[2] + w[3]) +
                                    bm * (dw0 + dw1 + dw2 + dw3);
                if (in) {
                    be[0 * g.cs + dg] += bm * w[0];
                    be[1 * g.cs + dg] += bm * w[1];
                    be[2 * g.cs + dg] += bm * w[2];
                    db[0 * g.cs + dg] += dbm * w[0] + bm * dw0;
                    db[1 * g.cs + dg] += dbm * w[1] + bm * dw1;
                    db[2 * g.cs + dg] += dbm * w[2] + bm * dw2;
                }
                /* I */
                real bi = be[1 * g.cs + c], dbi = db[1 * g.cs + c];
                wsoft(ci, 2, T, w);
                real u0 = (i > 1) ? da[0 * g.cs + up] : 0;
                real u1 = (i > 1) ? da[1 * g.cs + up] : 0;
                mean = w[0] * u0 + w[1] * u1;
                dw0 = w[0] * (u0 - mean) / T;
                dw1 = w[1] * (u1 - mean) / T;
                if (i > 1) {
                    be[0 * g.cs + up] += bi * w[0];
                    be[1 * g.cs + up] += bi * w[1];
                    db[0 * g.cs + up] += dbi * w[0] + bi * dw0;
                    db[1 * g.cs + up] += dbi * w[1] + bi * dw1;
                }
                /* D */
                real bd = be[2 * g.cs + c], dbd = db[2 * g.cs + c];
                wsoft(cd, 2, T, w);
                real l0 = (j > 1) ? da[0 * g.cs + lf] : 0;
                real l2 = (j > 1) ? da[2 * g.cs + lf] : 0;
                mean = w[0] * l0 + w[1] * l2;
                dw0 = w[0] * (l0 - mean) / T;
                dw1 = w[1] * (l2 - mean) / T;
                if (j > 1) {
                    be[0 * g.cs + lf] += bd * w[0];
                    be[2 * g.cs + lf] += bd * w[1];
                    db[0 * g.cs + lf] += dbd * w[0] + bd * dw0;
                    db[2 * g.cs + lf] += dbd * w[1] + bd * dw1;
                }
            }
        }
    }
    free(be);
}

/* sw_affine_param_grad_cpu, src/sw_affine/kernels_cpu.cpp:643-921.
 * ptype 0 = gap_open, 1 = gap_ext, 2 = temperature. */
void OFN(aff_param_grad)(const real* alpha, const real* scores, const real* part,
                         const real* dSdt, real* U, real* beta, real* W, real* dP,
                         const int* len, int B, int L1m, int L2m, real go, real ge, real T,
                         int ptype) {
    geom g = mk_geom(L1m, L2m);
    const real iT2 = (real)1 / (T * T);
    for (int b = 0; b < B; ++b) {
        const real* a = alpha + (size_t)b * 3 * g.cs;
        const real* s = scores + (size_t)b * g.ss;
        real* u = U + (size_t)b * 3 * g.cs;
        real* be = beta + (size_t)b * 3 * g.cs;
        real* w = W + (size_t)b * 3 * g.cs;
        real* out = dP + (size_t)b * g.ss;
        real S = part[b], dS = dSdt[b];
        int L1 = len[2 * b], L2 = len[2 * b + 1];
        memset(u, 0, 3 * g.cs * sizeof(real));
        memset(be, 0, 3 * g.cs * sizeof(real));
        memset(w, 0, 3 * g.cs * sizeof(real));
        memset(out, 0, g.ss * sizeof(real));
        /* U forward */
        for (int k = 2; k <= L1 + L2; ++k) {
            int lo = k - L2 > 1 ? k - L2 : 1, hi = k - 1 < L1 ? k - 1 : L1;
            for (int i = lo; i <= hi; ++i) {
                int j = k - i;
                size_t c = CELL(g, i, j), dg = CELL(g, i - 1, j - 1);
                size_t up = CELL(g, i - 1, j), lf = CELL(g, i, j - 1);
                real cm[4], ci[2], cd[2], wm[4], wi[2], wd[2];
                aff_cands(a, g, i, j, s[SIDX(g, i, j)], go, ge, cm, ci, cd);
                int in = (i > 1 && j > 1);
                wsoft(cm, 4, T, wm);
                real d0 = in ? u[0 * g.cs + dg] : 0, d1 = in ? u[1 * g.cs + dg] : 0;
                real d2 = in ? u[2 * g.cs + dg] : 0;
                real um = wm[0] * d0 + wm[1] * d1 + wm[2] * d2;
                if (ptype == 2) {
                    real e = wm[0] * cm[0] + wm[1] * cm[1] + wm[2] * cm[2] + wm[3] * cm[3];
                    um += (a[0 * g.cs + c] - e) / T;
                }
                u[0 * g.cs + c] = um;
                wsoft(ci, 2, T, wi);
                real i0 = (i > 1) ? u[0 * g.cs + up] : 0, i1 = (i > 1) ? u[1 * g.cs + up] : 0;
                if (ptype == 0) i0 += 1;
                if (ptype == 1) i1 += 1;
                real ui = wi[0] * i0 + wi[1] * i1;
                if (ptype == 2) {
                    real e = wi[0] * ci[0] + wi[1] * ci[1];
                    ui += (a[1 * g.cs + c] - e) / T;
                }
                u[1 * g.cs + c] = ui;
                wsoft(cd, 2, T, wd);
                real l0 = (j > 1) ? u[0 * g.cs + lf] : 0, l2 = (j > 1) ? u[2 * g.cs + lf] : 0;
                if (ptype == 0) l0 += 1;
                if (ptype == 1) l2 += 1;
                real ud = wd[0] * l0 + wd[1] * l2;
                if (ptype == 2) {
                    real e = wd[0] * cd[0] + wd[1] * cd[1];
                    ud += (a[2 * g.cs + c] - e) / T;
                }
                u[2 * g.cs + c] = ud;
            }
        }
        for (int i = 0; i <= L1; ++i)
            for (int j = 0; j <= L2; ++j)
                for (int st = 0; st < 3; ++st) {
                    size_t x = st * g.cs + CELL(g, i, j);
                    be[x] = (a[x] <= NINF_) ? (real)0 : post_w(a[x], S, T);
                }
        /* W backward */
        for (int k = L1 + L2; k >= 2; --k) {
            int lo = k - L2 > 1 ? k - L2 : 1, hi = k - 1 < L1 ? k - 1 : L1;
            for (int i = lo; i <= hi; ++i) {
                int j = k - i;
                size_t c = CELL(g, i, j), dg = CELL(g, i - 1, j - 1);
                size_t up = CELL(g, i - 1, j), lf = CELL(g, i, j - 1);
                size_t xm = 0 * g.cs + c, xi = 1 * g.cs + c, xd = 2 * g.cs + c;
                real sc = s[SIDX(g, i, j)];
                real cw[4];
                /* M */
                real bm = be[xm];
                real b0 = (a[xm] <= NINF_) ? (real)0 : post_w(a[xm], S, T);
                real Wm = w[xm] + b0 * (u[xm] - dS) / T;
                if (ptype == 2 && a[xm] > NINF_) Wm -= b0 * (a[xm] - S) / (T * T);
                real cm[4], ci[2], cd[2];
                aff_cands(a, g, i, j, sc, go, ge, cm, ci, cd);
                wsoft(cm, 4, T, cw);
                int in = (i > 1 && j > 1);
                real d0 = in ? u[0 * g.cs + dg] : 0, d1 = in ? u[1 * g.cs + dg] : 0;
                real d2 = in ? u[2 * g.cs + dg] : 0, d3 = 0;
                real mean = cw[0] * d0 + cw[1] * d1 + cw[2] * d2 + cw[3] * d3;
                real dw0 = cw[0] * (d0 - mean) / T, dw1 = cw[1] * (d1 - mean) / T;
                real dw2 = cw[2] * (d2 - mean) / T, dw3 = cw[3] * (d3 - mean) / T;
                if (ptype == 2) {
                    real e = cw[0] * cm[0] + cw[1] * cm[1] + cw[2] * cm[2] + cw[3] * cm[3];
                    dw0 += cw[0] * (e - cm[0]) * iT2;
                    dw1 += cw[1] * (e - cm[1]) * iT2;
                    dw2 += cw[2] * (e - cm[2]) * iT2;
                    dw3 += cw[3] * (e - cm[3]) * iT2;
                }
                (void)dw3;
                out[SIDX(g, i, j)] += Wm;
                if (in) {
                    be[0 * g.cs + dg] += bm * cw[0];
                    be[1 * g.cs + dg] += bm * cw[1];
                    be[2 * g.cs + dg] += bm * cw[2];
                    w[0 * g.cs + dg] += Wm * cw[0] + bm * dw0;
                    w[1 * g.cs + dg] += Wm * cw[1] + bm * dw1;
                    w[2 * g.cs + dg] += Wm * cw[2] + bm * dw2;
                }
                /* I */
                real bi = be[xi];
                real bi0 = (a[xi] <= NINF_) ? (real)0 : post_w(a[xi], S, T);
                real Wi = w[xi] + bi0 * (u[xi] - dS) / T;
                if (ptype == 2 && a[xi] > NINF_) Wi -= bi0 * (a[xi] - S) / (T * T);
                if (i > 1) {
                    real c0 = a[0 * g.cs + up] + go, c1 = a[1 * g.cs + up] + ge;
                    real cc[2] = {c0, c1}, ww[2];
                    wsoft(cc, 2, T, ww);
                    real e0 = u[0 * g.cs + up], e1 = u[1 * g.cs + up];
                    if (ptype == 0) e0 += 1;
                    if (ptype == 1) e1 += 1;
                    real mu = ww[0] * e0 + ww[1] * e1;
                    real q0 = ww[0] * (e0 - mu) / T, q1 = ww[1] * (e1 - mu) / T;
                    if (ptype == 2) {
                        real e = ww[0] * c0 + ww[1] * c1;
                        q0 += ww[0] * (e - c0) * iT2;
                        q1 += ww[1] * (e - c1) * iT2;
                    }
                    be[0 * g.cs + up] += bi * ww[0];
                    be[1 * g.cs + up] += bi * ww[1];
                    w[0 * g.cs + up] += Wi * ww[0] + bi * q0;
                    w[1 * g.cs + up] += Wi * ww[1] + bi * q1;
                }
                /* D */
                real bd = be[xd];
                real bd0 = (a[xd] <= NINF_) ? (real)0 : post_w(a[xd], S, T);
                real Wd = w[xd] + bd0 * (u[xd] - dS) / T;
                if (ptype == 2 && a[xd] > NINF_) Wd -= bd0 * (a[xd] - S) / (T * T);
                if (j > 1) {
                    real c0 = a[0 * g.cs + lf] + go, c1 = a[2 * g.cs + lf] + ge;
                    real cc[2] = {c0, c1}, ww[2];
                    wsoft(cc, 2, T, ww);
                    real e0 = u[0 * g.cs + lf], e1 = u[2 * g.cs + lf];
                    if (ptype == 0) e0 += 1;
                    if (ptype == 1) e1 += 1;
                    real mu = ww[0] * e0 + ww[1] * e1;
                    real q0 = ww[0] * (e0 - mu) / T, q1 = ww[1] * (e1 - mu) / T;
                    if (ptype == 2) {
                        real e = ww[0] * c0 + ww[1] * c1;
                        q0 += ww[0] * (e - c0) * iT2;
                        q1 += ww[1] * (e - c1) * iT2;
                    }
                    be[0 * g.cs + lf] += bd * ww[0];
                    be[2 * g.cs + lf] += bd * ww[1];
                    w[0 * g.cs + lf] += Wd * ww[0] + bd * q0;
                    w[2 * g.cs + lf] += Wd * ww[1] + bd * q1;
                }
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Linear gap, one state.                                                    */
/* ------------------------------------------------------------------------ */

static void lin_cands(const real* a, geom g, int i, int j, real sc, real gap, real* c) {
    c[0] = (i > 1 && j > 1) ? a[CELL(g, i - 1, j - 1)] + sc : NINF_;
    c[1] = (i > 1) ? a[CELL(g, i - 1, j)] + gap : NINF_;
    c[2] = (j > 1) ? a[CELL(g, i, j - 1)] + gap : NINF_;
    c[3] = sc;
}

/* sw_regular_forward_cpu, src/sw/kernels_cpu.cpp:70-148 */
void OFN(lin_forward)(const real* scores, real* alpha, real* part, const int* len,
                      int B, int L1m, int L2m, real gap, real T) {
    geom g = mk_geom(L1m, L2m);
    for (int b = 0; b < B; ++b) {
        const real* s = scores + (size_t)b * g.ss;
        real* a = alpha + (size_t)b * g.cs;
        int L1 = len[2 * b], L2 = len[2 * b + 1];
        for (size_t t = 0; t < g.cs; ++t) a[t] = NINF_;
        a[0] = 0;
        for (int k = 2; k <= L1 + L2; ++k) {
            int lo = k - L2 > 1 ? k - L2 : 1, hi = k - 1 < L1 ? k - 1 : L1;
            for (int i = lo; i <= hi; ++i) {
                int j = k - i;
                real c[4];
                lin_cands(a, g, i, j, s[SIDX(g, i, j)], gap, c);
                a[CELL(g, i, j)] = lse_n(c, 4, T);
            }
        }
        real m = NINF_;
        for (int i = 0; i <= L1; ++i)
            for (int j = 0; j <= L2; ++j) m = rmax(m, a[CELL(g, i, j)]);
        kahan z = {0, 0};
        for (int i = 0; i <= L1; ++i)
            for (int j = 0; j <= L2; ++j) {
                real v = a[CELL(g, i, j)];
                if (v > NINF_) kadd(&z, sexp((v - m) / T));
            }
        part[b] = m + T * OLOG(z.s);
    }
}

/* sw_regular_backward_cpu, src/sw/kernels_cpu.cpp:154-271 */
void OFN(lin_backward)(const real* alpha, const real* scores, const real* part, real* beta,
                       real* post, real* ggap, real* gT, const int* len,
                       int B, int L1m, int L2m, real gap, real T) {
    geom g = mk_geom(L1m, L2m);
    for (int b = 0; b < B; ++b) {
        const real* a = alpha + (size_t)b * g.cs;
        const real* s = scores + (size_t)b * g.ss;
        real* be = beta + (size_t)b * g.cs;
        real* P = post + (size_t)b * g.ss;
        real S = part[b];
        int L1 = len[2 * b], L2 = len[2 * b + 1];
        memset(P, 0, g.ss * sizeof(real));
        for (size_t t = 0; t < g.cs; ++t) be[t] = NINF_;
        for (int i = 0; i <= L1; ++i)
            for (int j = 0; j <= L2; ++j) be[CELL(g, i, j)] = post_w(a[CELL(g, i, j)], S, T);
        real sg = 0;
        for (int k = L1 + L2; k >= 2; --k) {
            int lo = k - L2 > 1 ? k - L2 : 1, hi = k - 1 < L1 ? k - 1 : L1;
            for (int i = lo; i <= hi; ++i) {
                int j = k - i;
                size_t c = CELL(g, i, j);
                real bij = be[c];
                if (is_bad(bij)) continue;
                real cc[4], w[4];
                lin_cands(a, g, i, j, s[SIDX(g, i, j)], gap, cc);
                wsoft(cc, 4, T, w);
                P[SIDX(g, i, j)] += bij * (w[0] + w[3]);
                if (i > 1 && j > 1 && w[0] > 0) be[CELL(g, i - 1, j - 1)] += bij * w[0];
                if (i > 1 && w[1] > 0) { be[CELL(g, i - 1, j)] += bij * w[1]; sg += bij * w[1]; }
                if (j > 1 && w[2] > 0) { be[CELL(g, i, j - 1)] += bij * w[2]; sg += bij * w[2]; }
            }
        }
        kahan ms = {0, 0};
        for (size_t x = 0; x < g.ss; ++x) kadd(&ms, P[x] * s[x]);
        real tot = ms.s + sg * gap;
        ggap[b] = sg;
        gT[b] = (S - tot) / T;
    }
}

/* sw_regular_hvp_cpu, src/sw/kernels_cpu.cpp:279-477 */
void OFN(lin_hvp)(const real* alpha, const real* scores, const real* part, const real* V,
                  real* dalpha, real* dpart, real* beta, real* dbeta, real* H, const int* len,
                  int B, int L1m, int L2m, real gap, real T) {
    geom g = mk_geom(L1m, L2m);
    for (int b = 0; b < B; ++b) {
        const real* a = alpha + (size_t)b * g.cs;
        const real* s = scores + (size_t)b * g.ss;
        const real* v = V + (size_t)b * g.ss;
        real* da = dalpha + (size_t)b * g.cs;
        real* be = beta + (size_t)b * g.cs;
        real* db = dbeta + (size_t)b * g.cs;
        real* h = H + (size_t)b * g.ss;
        real S = part[b];
        int L1 = len[2 * b], L2 = len[2 * b + 1];
        memset(da, 0, g.cs * sizeof(real));
        memset(be, 0, g.cs * sizeof(real));
        memset(db, 0, g.cs * sizeof(real));
        memset(h, 0, g.ss * sizeof(real));
        for (int k = 2; k <= L1 + L2; ++k) {
            int lo = k - L2 > 1 ? k - L2 : 1, hi = k - 1 < L1 ? k - 1 : L1;
            for (int i = lo; i <= hi; ++i) {
                int j = k - i;
                real vij = v[SIDX(g, i, j)];
                real cc[4], w[4];
                lin_cands(a, g, i, j, s[SIDX(g, i, j)], gap, cc);
                wsoft(cc, 4, T, w);
                real t0 = (i > 1 && j > 1) ? da[CELL(g, i - 1, j - 1)] + vij : 0;
                real t1 = (i > 1) ? da[CELL(g, i - 1, j)] : 0;
                real t2 = (j > 1) ? da[CELL(g, i, j - 1)] : 0;
                da[CELL(g, i, j)] = w[0] * t0 + w[1] * t1 + w[2] * t2 + w[3] * vij;
            }
        }
        kahan ds = {0, 0};
        for (int i = 0; i <= L1; ++i)
            for (int j = 0; j <= L2; ++j) {
                size_t x = CELL(g, i, j);
                kadd(&ds, post_w(a[x], S, T) * da[x]);
            }
        real dS = ds.s;
        dpart[b] = dS;
        for (int i = 0; i <= L1; ++i)
            for (int j = 0; j <= L2; ++j) be[CELL(g, i, j)] = post_w(a[CELL(g, i, j)], S, T);
        for (int i = 0; i <= L1; ++i)
            for (int j = 0; j <= L2; ++j) {
                size_t x = CELL(g, i, j);
                db[x] = be[x] * (da[x] - dS) / T;
            }
        for (int k = L1 + L2; k >= 2; --k) {
            int lo = k - L2 > 1 ? k - L2 : 1, hi = k - 1 < L1 ? k - 1 : L1;
            for (int i = lo; i <= hi; ++i) {
                int j = k - i;
                size_t c = CELL(g, i, j);
                real vij = v[SIDX(g, i, j)];
                real bij = be[c], dbij = db[c];
                if (is_bad(bij)) continue;
                real cc[4], w[4];
                lin_cands(a, g, i, j, s[SIDX(g, i, j)], gap, cc);
                wsoft(cc, 4, T, w);
                real t0 = (i > 1 && j > 1) ? da[CELL(g, i - 1, j - 1)] + vij : 0;
                real t1 = (i > 1) ? da[CELL(g, i - 1, j)] : 0;
                real t2 = (j > 1) ? da[CELL(g, i, j - 1)] : 0;
                real t3 = vij;
                real mean = w[0] * t0 + w[1] * t1 + w[2] * t2 + w[3] * t3;
                real dw0 = w[0] * (t0 - mean) / T, dw1 = w[1] * (t1 - mean) / T;
                real dw2 = w[2] * (t2 - mean) / T, dw3 = w[3] * (t3 - mean) / T;
                h[SIDX(g, i, j)] += dbij * (w[0] + w[3]) + bij * (dw0 + dw3);
                if (i > 1 && j > 1 && w[0] > 0) be[CELL(g, i - 1, j - 1)] += bij * w[0];
                if (i > 1 && w[1] > 0) be[CELL(g, i - 1, j)] += bij * w[1];
                if (j > 1 && w[2] > 0) be[CELL(g, i, j - 1)] += bij * w[2];
                if (i > 1 && j > 1 && w[0] > 0) db[CELL(g, i - 1, j - 1)] += dbij * w[0] + bij * dw0;
                if (i > 1 && w[1] > 0) db[CELL(g, i - 1, j)] += dbij * w[1] + bij * dw1;
                if (j > 1 && w[2] > 0) db[CELL(g, i, j - 1)] += dbij * w[2] + bij * dw2;
            }
        }
    }
}

/* Plain (non-Kahan) normalised weights with the 1e-20 floor used by the
 * linear sensitivity pass, src/sw/kernels_cpu.cpp:551-575 and :639-658.
 * Returns 0 when every candidate is NINF. */
static int lin_plain_w(const real* c, real T, real* w) {
    real m = rmax(rmax(rmax(c[0], c[1]), c[2]), c[3]);
    if (m <= NINF_) return 0;
    real sum = 0;
    for (int k = 0; k < 4; ++k) { w[k] = sexp((c[k] - m) / T); }
    sum = w[0] + w[1] + w[2] + w[3];
    real inv = (sum > (real)1e-20) ? (real)1 / sum : (real)0;
    for (int k = 0; k < 4; ++k) w[k] *= inv;
    return 1;
}

/* sw_regular_param_grad_cpu, src/sw/kernels_cpu.cpp:491-710.
 * ptype 0 = gap, 1 = temperature. */
void OFN(lin_param_grad)(const real* alpha, const real* scores, const real* part,
                         const real* dSdt, real* U, real* beta, real* W, real* dP,
                         const int* len, int B, int L1m, int L2m, real gap, real T, int ptype) {
    geom g = mk_geom(L1m, L2m);
    for (int b = 0; b < B; ++b) {
        const real* a = alpha + (size_t)b * g.cs;
        const real* s = scores + (size_t)b * g.ss;
        real* u = U + (size_t)b * g.cs;
        real* be = beta + (size_t)b * g.cs;
        real* w = W + (size_t)b * g.cs;
        real* out = dP + (size_t)b * g.ss;
        real S = part[b], dS = dSdt[b];
        int L1 = len[2 * b], L2 = len[2 * b + 1];
        memset(u, 0, g.cs * sizeof(real));
        memset(be, 0, g.cs * sizeof(real));
        memset(w, 0, g.cs * sizeof(real));
        memset(out, 0, g.ss * sizeof(real));
        for (int k = 2; k <= L1 + L2; ++k) {
            int lo = k - L2 > 1 ? k - L2 : 1, hi = k - 1 < L1 ? k - 1 : L1;
            for (int i = lo; i <= hi; ++i) {
                int j = k - i;
                size_t c = CELL(g, i, j);
                real cc[4], ww[4];
                lin_cands(a, g, i, j, s[SIDX(g, i, j)], gap, cc);
                if (!lin_plain_w(cc, T, ww)) { u[c] = 0; continue; }
                real d0 = (i > 1 && j > 1) ? u[CELL(g, i - 1, j - 1)] : 0;
                real d1 = (i > 1) ? u[CELL(g, i - 1, j)] : 0;
                real d2 = (j > 1) ? u[CELL(g, i, j - 1)] : 0;
                real d3 = 0;
                if (ptype == 0) { d1 += 1; d2 += 1; }
                real uv = ww[0] * d0 + ww[1] * d1 + ww[2] * d2 + ww[3] * d3;
                if (ptype == 1) {
                    real mv = ww[0] * cc[0] + ww[1] * cc[1] + ww[2] * cc[2] + ww[3] * cc[3];
                    uv += (a[c] - mv) / T;
                }
                u[c] = uv;
            }
        }
        for (int i = 0; i <= L1; ++i)
            for (int j = 0; j <= L2; ++j) be[CELL(g, i, j)] = post_w(a[CELL(g, i, j)], S, T);
        for (int k = L1 + L2; k >= 2; --k) {
            int lo = k - L2 > 1 ? k - L2 : 1, hi = k - 1 < L1 ? k - 1 : L1;
            for (int i = lo; i <= hi; ++i) {
                int j = k - i;
                size_t c = CELL(g, i, j);
                real bc = be[c];
                real b0 = 0;
                if (a[c] > NINF_) b0 = post_w(a[c], S, T);
                real dw0 = b0 * (u[c] - dS) / T;
                if (ptype == 1 && a[c] > NINF_) dw0 -= b0 * (a[c] - S) / (T * T);
                real wc = w[c] + dw0;
                if (bc <= (real)1e-20 && (wc < 0 ? -wc : wc) <= (real)1e-20) continue;
                real cc[4], ww[4];
                lin_cands(a, g, i, j, s[SIDX(g, i, j)], gap, cc);
                if (!lin_plain_w(cc, T, ww)) continue;
                real d0 = (i > 1 && j > 1) ? u[CELL(g, i - 1, j - 1)] : 0;
                real d1 = (i > 1) ? u[CELL(g, i - 1, j)] : 0;
                real d2 = (j > 1) ? u[CELL(g, i, j - 1)] : 0;
                real d3 = 0;
                if (ptype == 0) { d1 += 1; d2 += 1; }
                real mu = ww[0] * d0 + ww[1] * d1 + ww[2] * d2 + ww[3] * d3;
                real q[4];
                q[0] = ww[0] * (d0 - mu) / T;
                q[1] = ww[1] * (d1 - mu) / T;
                q[2] = ww[2] * (d2 - mu) / T;
                q[3] = ww[3] * (d3 - mu) / T;
                if (ptype == 1) {
                    real mv = ww[0] * cc[0] + ww[1] * cc[1] + ww[2] * cc[2] + ww[3] * cc[3];
                    real iT2 = (real)1 / (T * T);
                    for (int k2 = 0; k2 < 4; ++k2) q[k2] += ww[k2] * (mv - cc[k2]) * iT2;
                }
                out[SIDX(g, i, j)] += wc * (ww[0] + ww[3]) + bc * (q[0] + q[3]);
                if (cc[0] > NINF_) {
                    be[CELL(g, i - 1, j - 1)] += bc * ww[0];
                    w[CELL(g, i - 1, j - 1)] += wc * ww[0] + bc * q[0];
                }
                if (cc[1] > NINF_) {
                    be[CELL(g, i - 1, j)] += bc * ww[1];
                    w[CELL(g, i - 1, j)] += wc * ww[1] + bc * q[1];
                }
                if (cc[2] > NINF_) {
                    be[CELL(g, i, j - 1)] += bc * ww[2];
                    w[CELL(g, i, j - 1)] += wc * ww[2] + bc * q[2];
                }
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Needleman-Wunsch, affine gap (global alignment), three states M, I, D.    */
/* ------------------------------------------------------------------------ */

/* softmax3 of src/nw_affine/kernels_cpu.cpp:58-70: Kahan sum over the
 * entries above NINF only, NINF when the sum is not positive. */
static real nw_lse3(real a, real b, real c, real T) {
    real m = rmax(rmax(a, b), c);
    if (m <= NINF_) return NINF_;
    kahan acc = {0, 0};
    if (a > NINF_) kadd(&acc, sexp((a - m) / T));
    if (b > NINF_) kadd(&acc, sexp((b - m) / T));
    if (c > NINF_) kadd(&acc, sexp((c - m) / T));
    if (acc.s <= 0) return NINF_;
    return m + T * OLOG(acc.s);
}

/* softmax3_weights, src/nw_affine/kernels_cpu.cpp:73-94: plain sum. */
static void nw_w3(real a, real b, real c, real T, real* w) {
    real m = rmax(rmax(a, b), c);
    if (m <= NINF_) { w[0] = w[1] = w[2] = 0; return; }
    real ea = (a > NINF_) ? sexp((a - m) / T) : 0;
    real eb = (b > NINF_) ? sexp((b - m) / T) : 0;
    real ec = (c > NINF_) ? sexp((c - m) / T) : 0;
    real tot = ea + eb + ec;
    if (tot > 0) { w[0] = ea / tot; w[1] = eb / tot; w[2] = ec / tot; }
    else w[0] = w[1] = w[2] = 0;
}

/* nw_affine_forward_cpu, src/nw_affine/kernels_cpu.cpp:100-193: boundary
 * states M(0,0) = 0, I(i,0) = go + (i-1) ge, D(0,j) = go + (j-1) ge, no
 * restart, value = LSE(M, I, D) at (L1, L2). */
void OFN(nwa_forward)(const real* scores, real* alpha, real* part, const int* len,
                      int B, int L1m, int L2m, real go, real ge, real T) {
    geom g = mk_geom(L1m, L2m);
    for (int b = 0; b < B; ++b) {
        const real* s = scores + (size_t)b * g.ss;
        real* M = alpha + (size_t)b * 3 * g.cs;
        real* I = M + g.cs;
        real* D = I + g.cs;
        int L1 = len[2 * b], L2 = len[2 * b + 1];
        for (size_t t = 0; t < 3 * g.cs; ++t) M[t] = NINF_;
        M[0] = 0;
        for (int i = 1; i <= L1; ++i) I[CELL(g, i, 0)] = go + (real)(i - 1) * ge;
        for (int j = 1; j <= L2; ++j) D[CELL(g, 0, j)] = go + (real)(j - 1) * ge;
        for (int k = 2; k <= L1 + L2; ++k) {
            int lo = k - L2 > 1 ? k - L2 : 1, hi = k - 1 < L1 ? k - 1 : L1;
            for (int i = lo; i <= hi; ++i) {
                int j = k - i;
                size_t c = CELL(g, i, j), dg = CELL(g, i - 1, j - 1);
                size_t up = CELL(g, i - 1, j), lf = CELL(g, i, j - 1);
                M[c] = s[SIDX(g, i, j)] + nw_lse3(M[dg], I[dg], D[dg], T);
                I[c] = nw_lse3(M[up] + go, I[up] + ge, D[up] + go, T);
                D[c] = nw_lse3(M[lf] + go, I[lf] + go, D[lf] + ge, T);
            }
        }
        size_t f = CELL(g, L1, L2);
        part[b] = nw_lse3(M[f], I[f], D[f], T);
    }
}

/* nw_affine_backward_cpu, src/nw_affine/kernels_cpu.cpp:199-335: push-form
 * adjoint seeded with the terminal LSE weights; adjoints at or below 1e-20
 * are not propagated; the leading-gap boundary states add one open and
 * (length - 1) extends each; grad_T = (S - E[path score]) / T. */
void OFN(nwa_backward)(const real* alpha, const real* scores, const real* part, real* beta,
                       real* post, real* gopen, real* gext, real* gT, const int* len,
                       int B, int L1m, int L2m, real go, real ge, real T) {
    geom g = mk_geom(L1m, L2m);
    const real tiny = (real)1e-20;
    for (int b = 0; b < B; ++b) {
        const real* M = alpha + (size_t)b * 3 * g.cs;
        const real* I = M + g.cs;
        const real* D = I + g.cs;
        const real* s = scores + (size_t)b * g.ss;
        real* bM = beta + (size_t)b * 3 * g.cs;
        real* bI = bM + g.cs;
        real* bD = bI + g.cs;
        real* P = post + (size_t)b * g.ss;
        int L1 = len[2 * b], L2 = len[2 * b + 1];
        memset(bM, 0, 3 * g.cs * sizeof(real));
        memset(P, 0, g.ss * sizeof(real));
        real w[3];
        size_t f = CELL(g, L1, L2);
        nw_w3(M[f], I[f], D[f], T, w);
        bM[f] = w[0]; bI[f] = w[1]; bD[f] = w[2];
        real so = 0, se = 0;
        kahan ms = {0, 0};
        for (int k = L1 + L2; k >= 2; --k) {
            int lo = k - L2 > 1 ? k - L2 : 1, hi = k - 1 < L1 ? k - 1 : L1;
            for (int i = lo; i <= hi; ++i) {
                int j = k - i;
                size_t c = CELL(g, i, j), dg = CELL(g, i - 1, j - 1);
                size_t up = CELL(g, i - 1, j), lf = CELL(g, i, j - 1);
                real sc = s[SIDX(g, i, j)];
                real bm = bM[c], bi = bI[c], bd = bD[c];
                if (bm > tiny && nw_lse3(M[dg], I[dg], D[dg], T) > NINF_) {
                    nw_w3(M[dg], I[dg], D[dg], T, w);
                    if (M[dg] > NINF_) bM[dg] += bm * w[0];
                    if (I[dg] > NINF_) bI[dg] += bm * w[1];
                    if (D[dg] > NINF_) bD[dg] += bm * w[2];
                    P[SIDX(g, i, j)] += bm;
                    kadd(&ms, bm * sc);
                }
                if (bi > tiny) {
                    real x0 = M[up] + go, x1 = I[up] + ge, x2 = D[up] + go;
                    if (nw_lse3(x0, x1, x2, T) > NINF_) {
                        nw_w3(x0, x1, x2, T, w);
                        bM[up] += bi * w[0]; bI[up] += bi * w[1]; bD[up] += bi * w[2];
                        so += bi * (w[0] + w[2]);
                        se += bi * w[1];
                    }
                }
                if (bd > tiny) {
                    real x0 = M[lf] + go, x1 = I[lf] + go, x2 = D[lf] + ge;
                    if (nw_lse3(x0, x1, x2, T) > NINF_) {
                        nw_w3(x0, x1, x2, T, w);
                        bM[lf] += bd * w[0]; bI[lf] += bd * w[1]; bD[lf] += bd * w[2];
                        so += bd * (w[0] + w[1]);
                        se += bd * w[2];
                    }
                }
            }
        }
        for (int i = 1; i <= L1; ++i) {
            real v = bI[CELL(g, i, 0)];
            if (v > tiny) { so += v; se += v * (real)(i - 1); }
        }
        for (int j = 1; j <= L2; ++j) {
            real v = bD[CELL(g, 0, j)];
            if (v > tiny) { so += v; se += v * (real)(j - 1); }
        }
        gopen[b] = so;
        gext[b] = se;
        gT[b] = (part[b] - (ms.s + so * go + se * ge)) / T;
    }
}

/* ------------------------------------------------------------------------ */
/* Saigo-Vert, affine gap: SW affine's states plus the I -> D open edge, no  */
/* D -> I edge, paths end in M only.                                         */
/* ------------------------------------------------------------------------ */

/* Candidates at cell (i, j), src/sv_affine/kernels_cpu.cpp:138-167 */
static void sva_cands(const real* a, geom g, int i, int j, real sc, real go, real ge,
                      real* cm, real* ci, real* cd) {
    size_t dg = CELL(g, i - 1, j - 1), up = CELL(g, i - 1, j), lf = CELL(g, i, j - 1);
    int in = (i > 1 && j > 1);
    cm[0] = in ? a[0 * g.cs + dg] + sc : NINF_;
    cm[1] = in ? a[1 * g.cs + dg] + sc : NINF_;
    cm[2] = in ? a[2 * g.cs + dg] + sc : NINF_;
    cm[3] = sc;
    ci[0] = (i > 1) ? a[0 * g.cs + up] + go : NINF_;
    ci[1] = (i > 1) ? a[1 * g.cs + up] + ge : NINF_;
    cd[0] = (j > 1) ? a[0 * g.cs + lf] + go : NINF_;
    cd[1] = (j > 1) ? a[1 * g.cs + lf] + go : NINF_;
    cd[2] = (j > 1) ? a[2 * g.cs + lf] + ge : NINF_;
}

/* sv_affine_forward_cpu, src/sv_affine/kernels_cpu.cpp:96-196: partition
 * LSE(0, {M}), Kahan sum of the empty alignment first, then the M states in
 * row-major order. */
void OFN(sva_forward)(const real* scores, real* alpha, real* part, const int* len,
                      int B, int L1m, int L2m, real go, real ge, real T) {
    geom g = mk_geom(L1m, L2m);
    for (int b = 0; b < B; ++b) {
        const real* s = scores + (size_t)b * g.ss;
        real* a = alpha + (size_t)b * 3 * g.cs;
        int L1 = len[2 * b], L2 = len[2 * b + 1];
        for (size_t t = 0; t < 3 * g.cs; ++t) a[t] = NINF_;
        a[0] = 0;
        for (int k = 2; k <= L1 + L2; ++k) {
            int lo = k - L2 > 1 ? k - L2 : 1, hi = k - 1 < L1 ? k - 1 : L1;
            for (int i = lo; i <= hi; ++i) {
                int j = k - i;
                real cm[4], ci[2], cd[3];
                sva_cands(a, g, i, j, s[SIDX(g, i, j)], go, ge, cm, ci, cd);
                size_t c = CELL(g, i, j);
                a[0 * g.cs + c] = lse_n(cm, 4, T);
                a[1 * g.cs + c] = lse_n(ci, 2, T);
                a[2 * g.cs + c] = lse_n(cd, 3, T);
            }
        }
        real m = 0;
        for (int i = 1; i <= L1; ++i)
            for (int j = 1; j <= L2; ++j) m = rmax(m, a[CELL(g, i, j)]);
        kahan z = {0, 0};
        kadd(&z, sexp((0 - m) / T));
        for (int i = 1; i <= L1; ++i)
            for (int j = 1; j <= L2; ++j) {
                real v = a[CELL(g, i, j)];
                if (v > NINF_) kadd(&z, sexp((v - m) / T));
            }
        part[b] = m + T * OLOG(z.s);
    }
}

/* sv_affine_backward_cpu, src/sv_affine/kernels_cpu.cpp:203-356: terminal
 * weights on the M states only, push-form adjoint, I -> D counted as an
 * open.  With `linear` (sv_linear_backward_cpu,
 * src/sv_linear/kernels_cpu.cpp:203-365: go = ge = gap) every gap edge goes
 * into one Kahan-summed count, written to gopen, and the expected total is a
 * Kahan sum of the match and gap terms. */
static void sva_backward_core(const real* alpha, const real* scores, const real* part,
                              real* beta, real* post, real* gopen, real* gext, real* gT,
                              const int* len, int B, int L1m, int L2m, real go, real ge, real T,
                              int linear) {
    geom g = mk_geom(L1m, L2m);
    for (int b = 0; b < B; ++b) {
        const real* a = alpha + (size_t)b * 3 * g.cs;
        const real* s = scores + (size_t)b * g.ss;
        real* be = beta + (size_t)b * 3 * g.cs;
        real* P = post + (size_t)b * g.ss;
        real S = part[b];
        int L1 = len[2 * b], L2 = len[2 * b + 1];
        memset(P, 0, g.ss * sizeof(real));
        memset(be, 0, 3 * g.cs * sizeof(real));
        for (int i = 1; i <= L1; ++i)
            for (int j = 1; j <= L2; ++j) be[CELL(g, i, j)] = post_w(a[CELL(g, i, j)], S, T);
        real so = 0, se = 0;
        kahan gl = {0, 0};  /* the linear gap count */
#define SVA_OPEN(v) do { if (linear) kadd(&gl, (v)); else so += (v); } while (0)
#define SVA_EXT(v) do { if (linear) kadd(&gl, (v)); else se += (v); } while (0)
        for (int k = L1 + L2; k >= 2; --k) {
            int lo = k - L2 > 1 ? k - L2 : 1, hi = k - 1 < L1 ? k - 1 : L1;
            for (int i = lo; i <= hi; ++i) {
                int j = k - i;
                size_t c = CELL(g, i, j), dg = CELL(g, i - 1, j - 1);
                size_t up = CELL(g, i - 1, j), lf = CELL(g, i, j - 1);
                real cm[4], ci[2], cd[3], w[4];
                sva_cands(a, g, i, j, s[SIDX(g, i, j)], go, ge, cm, ci, cd);
                real bm = be[0 * g.cs + c];
                if (!is_bad(bm)) {
                    wsoft(cm, 4, T, w);
                    P[SIDX(g, i, j)] += bm * (w[0] + w[1] + w[2] + w[3]);
                    if (i > 1 && j > 1) {
                        if (w[0] > 0) be[0 * g.cs + dg] += bm * w[0];
                        if (w[1] > 0) be[1 * g.cs + dg] += bm * w[1];
                        if (w[2] > 0) be[2 * g.cs + dg] += bm * w[2];
                    }
                }
                real bi = be[1 * g.cs + c];
                if (!is_bad(bi)) {
                    wsoft(ci, 2, T, w);
                    if (i > 1) {
                        if (w[0] > 0) { be[0 * g.cs + up] += bi * w[0]; SVA_OPEN(bi * w[0]); }
                        if (w[1] > 0) { be[1 * g.cs + up] += bi * w[1]; SVA_EXT(bi * w[1]); }
                    }
                }
                real bd = be[2 * g.cs + c];
                if (!is_bad(bd)) {
                    wsoft(cd, 3, T, w);
                    if (j > 1) {
                        if (w[0] > 0) { be[0 * g.cs + lf] += bd * w[0]; SVA_OPEN(bd * w[0]); }
                        if (w[1] > 0) { be[1 * g.cs + lf] += bd * w[1]; SVA_OPEN(bd * w[1]); }
                        if (w[2] > 0) { be[2 * g.cs + lf] += bd * w[2]; SVA_EXT(bd * w[2]); }
                    }
                }
            }
        }
#undef SVA_OPEN
#undef SVA_EXT
        kahan ms = {0, 0};
        for (size_t x = 0; x < g.ss; ++x) kadd(&ms, P[x] * s[x]);
        if (linear) {  /* sv_linear kernels_cpu.cpp:353-360 */
            real cost = gl.s * go;
            kahan tot = {0, 0};
            kadd(&tot, ms.s);
            kadd(&tot, cost);
            gopen[b] = gl.s;
            gT[b] = (S - tot.s) / T;
        } else {
            real gap_cost = so * go + se * ge;  /* sv_affine kernels_cpu.cpp:348-349 */
            gopen[b] = so;
            gext[b] = se;
            gT[b] = (S - (ms.s + gap_cost)) / T;
        }
    }
}

void OFN(sva_backward)(const real* alpha, const real* scores, const real* part, real* beta,
                       real* post, real* gopen, real* gext, real* gT, const int* len,
                       int B, int L1m, int L2m, real go, real ge, real T) {
    sva_backward_core(alpha, scores, part, beta, post, gopen, gext, gT, len, B, L1m, L2m, go, ge,
                      T, 0);
}

/* sv_linear_forward_cpu, src/sv_linear/kernels_cpu.cpp:94-196: the affine
 * recurrence with one gap score in both roles (the same operations). */
void OFN(svl_forward)(const real* scores, real* alpha, real* part, const int* len,
                      int B, int L1m, int L2m, real gap, real T) {
    OFN(sva_forward)(scores, alpha, part, len, B, L1m, L2m, gap, gap, T);
}

void OFN(svl_backward)(const real* alpha, const real* scores, const real* part, real* beta,
                       real* post, real* ggap, real* gT, const int* len,
                       int B, int L1m, int L2m, real gap, real T) {
    sva_backward_core(alpha, scores, part, beta, post, ggap, NULL, gT, len, B, L1m, L2m, gap,
                      gap, T, 1);
}

/* ------------------------------------------------------------------------ */
/* Needleman-Wunsch, linear gap: one state, A(i,0) = i gap, A(0,j) = j gap, */
/* value A(L1, L2).                                                          */
/* ------------------------------------------------------------------------ */

/* nw_forward_cpu, src/nw/kernels_cpu.cpp:78-156 */
void OFN(nwl_forward)(const real* scores, real* alpha, real* part, const int* len,
                      int B, int L1m, int L2m, real gap, real T) {
    geom g = mk_geom(L1m, L2m);
    for (int b = 0; b < B; ++b) {
        const real* s = scores + (size_t)b * g.ss;
        real* a = alpha + (size_t)b * g.cs;
        int L1 = len[2 * b], L2 = len[2 * b + 1];
        for (size_t t = 0; t < g.cs; ++t) a[t] = NINF_;
        a[0] = 0;
        for (int i = 1; i <= L1; ++i) a[CELL(g, i, 0)] = (real)i * gap;
        for (int j = 1; j <= L2; ++j) a[CELL(g, 0, j)] = (real)j * gap;
        for (int k = 2; k <= L1 + L2; ++k) {
            int lo = k - L2 > 1 ? k - L2 : 1, hi = k - 1 < L1 ? k - 1 : L1;
            for (int i = lo; i <= hi; ++i) {
                int j = k - i;
                real v[3] = {a[CELL(g, i - 1, j - 1)] + s[SIDX(g, i, j)],
                             a[CELL(g, i - 1, j)] + gap, a[CELL(g, i, j - 1)] + gap};
                a[CELL(g, i, j)] = lse_n(v, 3, T);
            }
        }
        part[b] = a[CELL(g, L1, L2)];
    }
}

/* nw_backward_cpu, src/nw/kernels_cpu.cpp:162-272: unit terminal adjoint,
 * map = adjoint x diagonal weight, Kahan-summed gap count including the
 * boundary cells (i gap, j gap). */
void OFN(nwl_backward)(const real* alpha, const real* scores, const real* part, real* beta,
                       real* post, real* ggap, real* gT, const int* len,
                       int B, int L1m, int L2m, real gap, real T) {
    geom g = mk_geom(L1m, L2m);
    for (int b = 0; b < B; ++b) {
        const real* a = alpha + (size_t)b * g.cs;
        const real* s = scores + (size_t)b * g.ss;
        real* be = beta + (size_t)b * g.cs;
        real* P = post + (size_t)b * g.ss;
        int L1 = len[2 * b], L2 = len[2 * b + 1];
        memset(P, 0, g.ss * sizeof(real));
        memset(be, 0, g.cs * sizeof(real));
        be[CELL(g, L1, L2)] = 1;
        kahan gg = {0, 0}, ms = {0, 0};
        for (int k = L1 + L2; k >= 2; --k) {
            int lo = k - L2 > 1 ? k - L2 : 1, hi = k - 1 < L1 ? k - 1 : L1;
            for (int i = lo; i <= hi; ++i) {
                int j = k - i;
                size_t c = CELL(g, i, j), dg = CELL(g, i - 1, j - 1);
                size_t up = CELL(g, i - 1, j), lf = CELL(g, i, j - 1);
                real bij = be[c];
                if (bij == 0) continue;
                real sc = s[SIDX(g, i, j)];
                real v[3] = {a[dg] + sc, a[up] + gap, a[lf] + gap}, w[3];
                wsoft(v, 3, T, w);
                real pij = bij * w[0];
                P[SIDX(g, i, j)] = pij;
                kadd(&ms, pij * sc);
                kadd(&gg, bij * (w[1] + w[2]));
                if (w[0] > 0) be[dg] += bij * w[0];
                if (w[1] > 0) be[up] += bij * w[1];
                if (w[2] > 0) be[lf] += bij * w[2];
            }
        }
        for (int i = 1; i <= L1; ++i) kadd(&gg, be[CELL(g, i, 0)] * (real)i);
        for (int j = 1; j <= L2; ++j) kadd(&gg, be[CELL(g, 0, j)] * (real)j);
        ggap[b] = gg.s;
        gT[b] = (part[b] - (ms.s + gg.s * gap)) / T;
    }
}
