/*
 * icm_oracle.c -- CPU restatement of the reference ICM equity algorithm.
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the CPU baseline).  The product
 * library never links or calls this file.
 *
 * What is restated, with the reference lines followed:
 *   - erfc-trap nodes ............ src/cpu/icm.c:117-132
 *   - V1 / V2 closed forms ....... src/cpu/icm.c:145-170
 *   - naive engine ............... src/cpu/icm.c:1880-1914
 *   - linear engine (flat) ....... src/cpu/icm.c:2032-2075
 *   - hybrid engine .............. src/cpu/icm.c:2126-2318
 *       tree geometry ............ src/cpu/icm.c:1205-1263 (k_pad = k)
 *       tree build ............... src/cpu/icm.c:1605-1654
 *       propagate (+subset prune)  src/cpu/icm.c:1667-1819
 *       hot mask ................. src/cpu/icm.c:2430-2473
 *   - integration loop ........... src/cpu/icm.c:2523-2611
 *
 * Differences that change rounding only (all engines compute the same
 * quadrature sum):
 *   - k is not padded to a "smooth" FFT size (best_k_pad, icm.c:661): the
 *     payout is zero beyond k, so the padded and unpadded trees agree.
 *   - FFT products use this file's own radix-2 FFT with twiddles rounded
 *     from long-double cos/sin, sized to the full linear convolution (no
 *     cyclic-wrap correction), instead of FFTW with wrap correction.
 *   - thread-local equity partials use a static q schedule and are merged
 *     in thread order, so the result is deterministic for a thread count.
 */
#include "icm_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ */
/* Quadrature: src/cpu/icm.c:117-132                                  */
/* ------------------------------------------------------------------ */

static double log_Phi(double y) {
    if (y >= 0) return log1p(-erfc(y / sqrt(2.0)) / 2.0);
    return log(erfc(-y / sqrt(2.0)) / 2.0);
}

void orc_make_nodes(int Q, double Smax, double *logv, double *w) {
    double y_lo = -7.7;
    double y_hi = sqrt(2.0) * sqrt(log(Smax) + 25.0);
    if (y_hi < 6.5) y_hi = 6.5;
    double h = (Q > 1) ? (y_hi - y_lo) / (Q - 1) : 0.0;
    for (int q = 0; q < Q; q++) {
        double y = y_lo + q * h;
        logv[q] = log_Phi(y);
        w[q] = h * exp(-0.5 * y * y) / sqrt(2.0 * 3.14159265358979323846);
    }
}

/* ------------------------------------------------------------------ */
/* Closed forms: src/cpu/icm.c:145-170                                */
/* ------------------------------------------------------------------ */

void orc_v1_exact(int n, const double *S, double *V1) {
    for (int i = 0; i < n; i++) {
        double v = 1.0;
        for (int j = 0; j < n; j++)
            if (j != i) v += S[i] / (S[i] + S[j]);
        V1[i] = v;
    }
}

void orc_v2_exact(int n, const double *S, double *V2) {
    for (int i = 0; i < n; i++) {
        double v = 0;
        for (int j = 0; j < n; j++) {
            if (j == i) continue;
            for (int l = j + 1; l < n; l++) {
                if (l == i) continue;
                v += S[i] / (S[i] + S[j] + S[l]);
            }
        }
        V2[i] = v;
    }
}

/* ------------------------------------------------------------------ */
/* Radix-2 complex FFT (oracle's own; replaces FFTW)                  */
/* ------------------------------------------------------------------ */

typedef struct { double re, im; } cplx;

#define ORC_MAX_LOG2 27
/* One table for the largest size prepared so far:
 * g_tw[j] = exp(-2 pi i j / 2^g_tw_p), j < 2^(g_tw_p - 1). */
static cplx *g_tw = NULL;
static int g_tw_p = 0;

/* Not thread-safe: called before any parallel region. */
static void tw_prepare(int pmax) {
    if (pmax > ORC_MAX_LOG2) pmax = ORC_MAX_LOG2;
    if (pmax <= g_tw_p) return;
    size_t N = (size_t)1 << pmax;
    size_t H = N / 2;
    cplx *t = (cplx *)malloc(H * sizeof(cplx));
    const long double PI = 3.141592653589793238462643383279502884L;
    for (size_t j = 0; j < H; j++) {
        long double ang = -2.0L * PI * (long double)j / (long double)N;
        t[j].re = (double)cosl(ang);
        t[j].im = (double)sinl(ang);
    }
    free(g_tw);
    g_tw = t;
    g_tw_p = pmax;
}

static int ceil_log2(size_t v) {
    int p = 0;
    while (((size_t)1 << p) < v) p++;
    return p;
}

static void fft_c(cplx *x, int p, int inverse) {
    size_t N = (size_t)1 << p;
    for (size_t i = 1, j = 0; i < N; i++) {
        size_t bit = N >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) { cplx t = x[i]; x[i] = x[j]; x[j] = t; }
    }
    const cplx *tw = g_tw;
    for (int s = 1; s <= p; s++) {
        size_t m = (size_t)1 << s, h = m >> 1;
        size_t stride = ((size_t)1 << (g_tw_p - s));
        for (size_t j = 0; j < h; j++) {
            cplx w = tw[j * stride];
            if (inverse) w.im = -w.im;
            for (size_t b = 0; b < N; b += m) {
                cplx u = x[b + j], v = x[b + j + h];
                double tr = v.re * w.re - v.im * w.im;
                double ti = v.re * w.im + v.im * w.re;
                x[b + j].re = u.re + tr;      x[b + j].im = u.im + ti;
                x[b + j + h].re = u.re - tr;  x[b + j + h].im = u.im - ti;
            }
        }
    }
}

/* Split the spectrum Z of (x + i y), x and y real, into X and Y at bin j. */
static inline void split2(cplx zj, cplx zm, cplx *X, cplx *Y) {
    /* X = (Z[j] + conj(Z[-j]))/2 ; Y = (Z[j] - conj(Z[-j]))/(2i) */
    X->re = 0.5 * (zj.re + zm.re);
    X->im = 0.5 * (zj.im - zm.im);
    Y->re = 0.5 * (zj.im + zm.im);
    Y->im = -0.5 * (zj.re - zm.re);
}

static inline cplx cmul(cplx a, cplx b) {
    cplx r = { a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re };
    return r;
}
static inline cplx cmulc(cplx a, cplx b) {   /* a * conj(b) */
    cplx r = { a.re * b.re + a.im * b.im, a.im * b.re - a.re * b.im };
    return r;
}

/* c[0..nc) = (a[0..na) * b[0..nb)) truncated to nc, exact linear size. */
static void polymul_fft(const double *a, int na, const double *b, int nb,
                        double *c, int nc, cplx *ws) {
    int conv = na + nb - 1;
    int p = ceil_log2((size_t)conv);
    size_t N = (size_t)1 << p;
    for (size_t i = 0; i < N; i++) {
        ws[i].re = (i < (size_t)na) ? a[i] : 0.0;
        ws[i].im = (i < (size_t)nb) ? b[i] : 0.0;
    }
    fft_c(ws, p, 0);
    for (size_t j = 0; j <= N / 2; j++) {
        size_t mj = (N - j) & (N - 1);
        cplx X, Y;
        split2(ws[j], ws[mj], &X, &Y);
        cplx P = cmul(X, Y);
        ws[j] = P;
        if (mj != j) { ws[mj].re = P.re; ws[mj].im = -P.im; }
    }
    fft_c(ws, p, 1);
    double inv = 1.0 / (double)N;
    for (int m = 0; m < nc; m++) c[m] = (m < conv) ? ws[m].re * inv : 0.0;
}

/* Pair correlate through the FFT:
 *   outL[m] = sum_t PR[t] g[m+t], outR[m] = sum_t PL[t] g[m+t], m < lo,
 * g treated as zero beyond lg. */
static void correlate_pair_fft(const double *g, int lg, const double *PL,
                               const double *PR, int lp, double *outL,
                               double *outR, int lo, cplx *ws, cplx *ws2) {
    int need = lo + lp - 1;
    int gl = lg < need ? lg : need;
    int p = ceil_log2((size_t)need);
    size_t N = (size_t)1 << p;
    for (size_t i = 0; i < N; i++) {
        ws[i].re = (i < (size_t)lp) ? PL[i] : 0.0;
        ws[i].im = (i < (size_t)lp) ? PR[i] : 0.0;
        ws2[i].re = (i < (size_t)gl) ? g[i] : 0.0;
        ws2[i].im = 0.0;
    }
    fft_c(ws, p, 0);
    fft_c(ws2, p, 0);
    for (size_t j = 0; j <= N / 2; j++) {
        size_t mj = (N - j) & (N - 1);
        cplx XL, XR;
        split2(ws[j], ws[mj], &XL, &XR);
        cplx G = ws2[j];
        cplx A = cmulc(G, XR);   /* -> outL */
        cplx Bv = cmulc(G, XL);  /* -> outR */
        /* pack U = A + i Bv (A, Bv hermitian) */
        cplx Uj = { A.re - Bv.im, A.im + Bv.re };
        cplx Um = { A.re + Bv.im, -A.im + Bv.re };  /* A* + i Bv* at -j */
        ws[j] = Uj;
        if (mj != j) ws[mj] = Um;
    }
    fft_c(ws, p, 1);
    double inv = 1.0 / (double)N;
    for (int m = 0; m < lo; m++) {
        outL[m] = ws[m].re * inv;
        outR[m] = ws[m].im * inv;
    }
}

/* Single correlate through the FFT: out[m] = sum_t P[t] g[m+t]. */
static void correlate_fft(const double *g, int lg, const double *P, int lp,
                          double *out, int lo, cplx *ws) {
    int need = lo + lp - 1;
    int gl = lg < need ? lg : need;
    int p = ceil_log2((size_t)need);
    size_t N = (size_t)1 << p;
    for (size_t i = 0; i < N; i++) {
        ws[i].re = (i < (size_t)gl) ? g[i] : 0.0;
        ws[i].im = (i < (size_t)lp) ? P[i] : 0.0;
    }
    fft_c(ws, p, 0);
    for (size_t j = 0; j <= N / 2; j++) {
        size_t mj = (N - j) & (N - 1);
        cplx G, X;
        split2(ws[j], ws[mj], &G, &X);
        cplx A = cmulc(G, X);
        ws[j] = A;
        if (mj != j) { ws[mj].re = A.re; ws[mj].im = -A.im; }
    }
    fft_c(ws, p, 1);
    double inv = 1.0 / (double)N;
    for (int m = 0; m < lo; m++) out[m] = ws[m].re * inv;
}

/* ------------------------------------------------------------------ */
/* Schoolbook kernels: src/cpu/icm.c:465-507                          */
/* ------------------------------------------------------------------ */

static void polymul_modk(const double *a, int na, const double *b, int nb,
                         double *c, int k) {
    memset(c, 0, (size_t)k * sizeof(double));
    for (int i = 0; i < na && i < k; i++) {
        double ai = a[i];
        if (ai == 0.0) continue;
        int jmax = nb;
        if (i + jmax > k) jmax = k - i;
        for (int j = 0; j < jmax; j++) c[i + j] += ai * b[j];
    }
}

static void correlate_school(const double *g, int len_g, const double *P,
                             int len_P, double *out, int len_out) {
    for (int m = 0; m < len_out; m++) {
        double sum = 0;
        int jmax = len_P;
        if (m + jmax > len_g) jmax = len_g - m;
        for (int j = 0; j < jmax; j++) sum += P[j] * g[m + j];
        out[m] = sum;
    }
}

/* ------------------------------------------------------------------ */
/* Naive engine: src/cpu/icm.c:1880-1914                              */
/* ------------------------------------------------------------------ */

static void engine_naive(int n, const double *a, const double *payout, int k,
                         double *inner, const unsigned char *active) {
    int pk = (k < n) ? k : n;
    double *P = (double *)calloc((size_t)n + 1, sizeof(double));
    double *Qv = (double *)malloc((size_t)n * sizeof(double));
    P[0] = 1;
    for (int j = 0; j < n; j++) {
        double aj = a[j], bj = 1 - aj;
        for (int m = (j + 1 < n ? j + 1 : n); m >= 1; m--)
            P[m] = aj * P[m] + bj * P[m - 1];
        P[0] *= aj;
    }
    for (int i = 0; i < n; i++) {
        if (active && !active[i]) { inner[i] = 0; continue; }
        double ai = a[i], bi = 1 - ai;
        if (ai > 0.5) {
            double ia = 1 / ai, c = -bi * ia;
            Qv[0] = P[0] * ia;
            for (int m = 1; m < n; m++) Qv[m] = c * Qv[m - 1] + P[m] * ia;
        } else if (ai > 1e-15) {
            double ib = 1 / bi, c = -ai * ib;
            Qv[n - 1] = P[n] * ib;
            for (int m = n - 2; m >= 0; m--) Qv[m] = c * Qv[m + 1] + P[m + 1] * ib;
        } else {
            for (int m = 0; m < n; m++) Qv[m] = P[m + 1];
        }
        double eq = 0;
        for (int m = 0; m < pk; m++) eq += payout[m] * Qv[m];
        inner[i] = eq;
    }
    free(P);
    free(Qv);
}

/* ------------------------------------------------------------------ */
/* Linear forward-backward (flat): src/cpu/icm.c:2032-2075            */
/* ------------------------------------------------------------------ */

static void engine_linear(int n, const double *a, const double *payout, int k,
                          double *inner, const unsigned char *active) {
    double *g_store = (double *)malloc((size_t)n * k * sizeof(double));
    double *R = (double *)malloc((size_t)k * sizeof(double));
    const double *g_prev = payout;
    for (int j = 0; j < n; j++) {
        double aj = a[j], bj = 1 - aj;
        double *g_cur = g_store + (size_t)j * k;
        for (int m = 0; m < k - 1; m++) g_cur[m] = aj * g_prev[m] + bj * g_prev[m + 1];
        g_cur[k - 1] = aj * g_prev[k - 1];
        g_prev = g_cur;
    }
    memset(R, 0, (size_t)k * sizeof(double));
    R[0] = 1.0;
    for (int j = n - 1; j >= 0; j--) {
        double aj = a[j], bj = 1 - aj;
        if (!active || active[j]) {
            const double *gb = (j > 0) ? (g_store + (size_t)(j - 1) * k) : payout;
            double eq = gb[0] * R[0];
            for (int m = k - 1; m >= 1; m--) {
                eq += gb[m] * R[m];
                R[m] = aj * R[m] + bj * R[m - 1];
            }
            R[0] = aj * R[0];
            inner[j] = eq;
        } else {
            for (int m = k - 1; m >= 1; m--) R[m] = aj * R[m] + bj * R[m - 1];
            R[0] = aj * R[0];
            inner[j] = 0;
        }
    }
    free(g_store);
    free(R);
}

/* ------------------------------------------------------------------ */
/* Hybrid engine                                                      */
/* ------------------------------------------------------------------ */

#define ORC_MAX_LEVELS 40

typedef struct {
    int N, L;
    int n_real[ORC_MAX_LEVELS], nn[ORC_MAX_LEVELS], psz[ORC_MAX_LEVELS];
    int below_sat[ORC_MAX_LEVELS], g_needed[ORC_MAX_LEVELS];
    size_t plev_off[ORC_MAX_LEVELS];
    size_t plev_total, max_g;
} Geo;

/* Tree geometry: src/cpu/icm.c:1205-1263 (k_pad = k). */
static void geo_make(Geo *t, int n_leaves, int leaf_degree, int k, int leaf_extract) {
    memset(t, 0, sizeof(*t));
    t->N = 1;
    while (t->N < n_leaves) t->N <<= 1;
    t->L = 0;
    { int tmp = t->N; while (tmp > 1) { tmp >>= 1; t->L++; } t->L++; }
    t->n_real[0] = n_leaves;
    for (int l = 1; l < t->L; l++) t->n_real[l] = (t->n_real[l - 1] + 1) / 2;
    size_t off = 0;
    for (int l = 0; l < t->L; l++) {
        t->nn[l] = t->N >> l;
        long d = (long)leaf_degree * (1L << (l + 1));
        t->psz[l] = (d > k) ? k : (int)d;
        t->plev_off[l] = off;
        size_t sz = (size_t)t->nn[l] * t->psz[l];
        off += sz;
        if (sz > t->max_g) t->max_g = sz;
    }
    t->plev_total = off;
    for (int l = 0; l < t->L; l++) {
        t->below_sat[l] = 0;
        if (l == 0) continue;
        int cps = t->psz[l - 1];
        if (t->psz[l] == 2 * cps && cps >= 2) t->below_sat[l] = 1;
    }
    t->g_needed[0] = (leaf_extract < t->psz[0]) ? leaf_extract : t->psz[0];
    for (int l = 1; l < t->L; l++) {
        int need = t->g_needed[l - 1] + t->psz[l - 1] - 1;
        t->g_needed[l] = (need < t->psz[l]) ? need : t->psz[l];
    }
}

typedef struct {
    Geo t;
    double *ws;        /* plev_total + 2 * max_g */
    cplx *f1, *f2;     /* FFT work, 2 * max conv each */
    int fft_min;
} Tree;

static int use_fft_for(const Tree *tr, int len) {
    return tr->fft_min > 0 && len >= tr->fft_min;
}

/* Tree build levels 1..L-2: src/cpu/icm.c:1605-1654. */
static void tree_build(Tree *tr) {
    Geo *t = &tr->t;
    for (int l = 1; l < t->L - 1; l++) {
        int cps = t->psz[l - 1], pps = t->psz[l];
        int p_eff = t->below_sat[l] ? cps / 2 + 1 : cps;
        double *cb = tr->ws + t->plev_off[l - 1];
        double *pb = tr->ws + t->plev_off[l];
        for (int j = 0; j < t->n_real[l]; j++) {
            double *Lc = cb + (size_t)(2 * j) * cps;
            double *out = pb + (size_t)j * pps;
            if (2 * j + 1 >= t->n_real[l - 1]) {
                int cp = (cps < pps) ? cps : pps;
                memcpy(out, Lc, (size_t)cp * sizeof(double));
                if (cp < pps) memset(out + cp, 0, (size_t)(pps - cp) * sizeof(double));
            } else {
                double *Rc = cb + (size_t)(2 * j + 1) * cps;
                if (use_fft_for(tr, p_eff))
                    polymul_fft(Lc, p_eff, Rc, p_eff, out, pps, tr->f1);
                else
                    polymul_modk(Lc, p_eff, Rc, p_eff, out, pps);
            }
        }
    }
}

/* Propagate: src/cpu/icm.c:1667-1819.  hot == NULL: all nodes hot. */
static double *tree_propagate(Tree *tr, int k, const double *payout,
                              unsigned char *const *hot) {
    Geo *t = &tr->t;
    int L = t->L, top = L - 1;
    double *g0 = tr->ws + t->plev_total;
    double *g1 = g0 + t->max_g;
    double *gp_base = g0;
    int root_gsz = t->psz[top];
    memset(gp_base, 0, (size_t)t->nn[top] * root_gsz * sizeof(double));
    int copy = (k < root_gsz) ? k : root_gsz;
    memcpy(gp_base, payout, (size_t)copy * sizeof(double));

    for (int l = top; l >= 1; l--) {
        int pgsz = t->psz[l], cgsz = t->psz[l - 1], cps = t->psz[l - 1];
        double *gc_base = (gp_base == g0) ? g1 : g0;
        double *cb = tr->ws + t->plev_off[l - 1];
        int is_below = t->below_sat[l];
        int p_eff = is_below ? cps / 2 + 1 : cps;
        int out_needed = t->g_needed[l - 1];
        int g_eff_needed = out_needed + p_eff - 1;
        int g_eff_max = is_below ? (cgsz + cps / 2) : pgsz;
        int g_eff = (g_eff_needed < g_eff_max) ? g_eff_needed : g_eff_max;
        int nrp = t->n_real[l], nrc = t->n_real[l - 1];
        const unsigned char *hc = hot ? hot[l - 1] : NULL;
        int fft = use_fft_for(tr, p_eff);

        for (int j = 0; j < nrp; j++) {
            double *gp = gp_base + (size_t)j * pgsz;
            double *gL = gc_base + (size_t)(2 * j) * cgsz;
            double *gR = gc_base + (size_t)(2 * j + 1) * cgsz;
            int right_exists = (2 * j + 1 < nrc);
            int left_hot = hc ? hc[2 * j] : 1;
            int right_hot = right_exists ? (hc ? hc[2 * j + 1] : 1) : 0;
            if (!right_exists) {
                if (left_hot) {
                    int cp = (out_needed < pgsz) ? out_needed : pgsz;
                    memcpy(gL, gp, (size_t)cp * sizeof(double));
                    if (cp < out_needed)
                        memset(gL + cp, 0, (size_t)(out_needed - cp) * sizeof(double));
                } else {
                    memset(gL, 0, (size_t)cgsz * sizeof(double));
                }
                continue;
            }
            double *PL = cb + (size_t)(2 * j) * cps;
            double *PR = cb + (size_t)(2 * j + 1) * cps;
            if (left_hot && right_hot) {
                if (fft)
                    correlate_pair_fft(gp, g_eff, PL, PR, p_eff, gL, gR, out_needed,
                                       tr->f1, tr->f2);
                else {
                    correlate_school(gp, g_eff, PR, p_eff, gL, out_needed);
                    correlate_school(gp, g_eff, PL, p_eff, gR, out_needed);
                }
            } else if (left_hot) {
                memset(gR, 0, (size_t)cgsz * sizeof(double));
                if (fft) correlate_fft(gp, g_eff, PR, p_eff, gL, out_needed, tr->f1);
                else correlate_school(gp, g_eff, PR, p_eff, gL, out_needed);
            } else if (right_hot) {
                memset(gL, 0, (size_t)cgsz * sizeof(double));
                if (fft) correlate_fft(gp, g_eff, PL, p_eff, gR, out_needed, tr->f1);
                else correlate_school(gp, g_eff, PL, p_eff, gR, out_needed);
            } else {
                memset(gL, 0, (size_t)cgsz * sizeof(double));
                memset(gR, 0, (size_t)cgsz * sizeof(double));
            }
        }
        gp_base = gc_base;
    }
    return gp_base;
}

typedef struct {
    int B, nblocks;
    double *block_prods;
    Tree tr;
    unsigned char *hot_data;
    unsigned char *hot[ORC_MAX_LEVELS];
} Hybrid;

static size_t max_conv_len(const Geo *t) {
    size_t m = 2;
    for (int l = 0; l < t->L; l++) {
        size_t v = 2 * (size_t)t->psz[l] + 2 * (size_t)(l > 0 ? t->psz[l - 1] : 0) + 4;
        if (v > m) m = v;
    }
    return m;
}

static void hybrid_init(Hybrid *h, int n, int k, int B, int fft_min) {
    memset(h, 0, sizeof(*h));
    h->B = B;
    h->nblocks = (n + B - 1) / B;
    geo_make(&h->tr.t, h->nblocks, B, k, B);
    h->tr.fft_min = fft_min;
    h->tr.ws = (double *)calloc(h->tr.t.plev_total + 2 * h->tr.t.max_g, sizeof(double));
    h->block_prods = (double *)calloc((size_t)h->tr.t.N * (B + 1), sizeof(double));
    size_t mc = (size_t)1 << ceil_log2(max_conv_len(&h->tr.t));
    h->tr.f1 = (cplx *)malloc(mc * sizeof(cplx));
    h->tr.f2 = (cplx *)malloc(mc * sizeof(cplx));
}

static void hybrid_free(Hybrid *h) {
    free(h->tr.ws);
    free(h->block_prods);
    free(h->tr.f1);
    free(h->tr.f2);
    free(h->hot_data);
}

/* Hot mask: src/cpu/icm.c:2430-2473 (active in engine order). */
static void hybrid_hot_mask(Hybrid *h, int n, const unsigned char *active) {
    Geo *t = &h->tr.t;
    size_t total = 0;
    for (int l = 0; l < t->L; l++) total += (size_t)t->nn[l];
    h->hot_data = (unsigned char *)calloc(total, 1);
    unsigned char *p = h->hot_data;
    for (int l = 0; l < t->L; l++) { h->hot[l] = p; p += t->nn[l]; }
    for (int b = 0; b < h->nblocks; b++) {
        int s = b * h->B, e = s + h->B;
        if (e > n) e = n;
        int hotb = 0;
        for (int j = s; j < e; j++) if (active[j]) { hotb = 1; break; }
        h->hot[0][b] = (unsigned char)hotb;
    }
    for (int l = 1; l < t->L; l++) {
        int nc = t->nn[l - 1], np = t->nn[l];
        for (int j = 0; j < np; j++) {
            int lft = (2 * j < nc) ? h->hot[l - 1][2 * j] : 0;
            int rgt = (2 * j + 1 < nc) ? h->hot[l - 1][2 * j + 1] : 0;
            h->hot[l][j] = (unsigned char)(lft | rgt);
        }
    }
}

/* Block build + tree + bidirectional divide: src/cpu/icm.c:2186-2303. */
static void engine_hybrid(Hybrid *h, int n, const double *a, const double *payout,
                          int k, double *inner, const unsigned char *active) {
    int B = h->B, nblocks = h->nblocks;
    Geo *t = &h->tr.t;
    int leaf_psz = t->psz[0];
    for (int b = 0; b < nblocks; b++) {
        int s = b * B, e = s + B;
        if (e > n) e = n;
        int bsize = e - s;
        double *P = h->block_prods + (size_t)b * (B + 1);
        memset(P, 0, (size_t)(B + 1) * sizeof(double));
        P[0] = 1.0;
        for (int j = s; j < e; j++) {
            double aj = a[j], bj = 1 - aj;
            for (int m = bsize; m >= 1; m--) P[m] = aj * P[m] + bj * P[m - 1];
            P[0] *= aj;
        }
        double *leaf = h->tr.ws + t->plev_off[0] + (size_t)b * leaf_psz;
        int cp = (B + 1 < leaf_psz) ? B + 1 : leaf_psz;
        memcpy(leaf, P, (size_t)cp * sizeof(double));
        if (cp < leaf_psz) memset(leaf + cp, 0, (size_t)(leaf_psz - cp) * sizeof(double));
    }
    for (int b = nblocks; b < t->N; b++) {
        double *leaf = h->tr.ws + t->plev_off[0] + (size_t)b * leaf_psz;
        memset(leaf, 0, (size_t)leaf_psz * sizeof(double));
        leaf[0] = 1.0;
    }
    tree_build(&h->tr);
    double *g_leaf = tree_propagate(&h->tr, k, payout, active ? h->hot : NULL);
    int g_need = t->g_needed[0];

    double *Qarr = (double *)malloc((size_t)(B + 1) * sizeof(double));
    for (int b = 0; b < nblocks; b++) {
        int s = b * B, e = s + B;
        if (e > n) e = n;
        int bsize = e - s;
        if (active) {
            int any = 0;
            for (int j = s; j < e; j++) if (active[j]) { any = 1; break; }
            if (!any) { for (int j = s; j < e; j++) inner[j] = 0; continue; }
        }
        const double *P_b = h->block_prods + (size_t)b * (B + 1);
        const double *g_b = g_leaf + (size_t)b * leaf_psz;
        int pk_g = g_need < bsize ? g_need : bsize;
        if (pk_g > k) pk_g = k;
        for (int jj = 0; jj < bsize; jj++) {
            if (active && !active[s + jj]) { inner[s + jj] = 0; continue; }
            double aj = a[s + jj], bj = 1 - aj, eq = 0;
            if (aj > 0.5) {
                double ia = 1.0 / aj, c = -bj * ia;
                double qv = P_b[0] * ia;
                eq = g_b[0] * qv;
                for (int m = 1; m < pk_g; m++) {
                    qv = c * qv + P_b[m] * ia;
                    eq += g_b[m] * qv;
                }
            } else if (aj > 1e-15) {
                double ib = 1.0 / bj, c = -aj * ib;
                double qv = P_b[bsize] * ib;
                Qarr[bsize - 1] = qv;
                for (int m = bsize - 2; m >= 0; m--) {
                    qv = c * qv + P_b[m + 1] * ib;
                    Qarr[m] = qv;
                }
                for (int m = 0; m < pk_g; m++) eq += g_b[m] * Qarr[m];
            } else {
                for (int m = 0; m < pk_g; m++) eq += g_b[m] * P_b[m + 1];
            }
            inner[s + jj] = eq;
        }
    }
    free(Qarr);
}

/* ------------------------------------------------------------------ */
/* Per-node entry                                                     */
/* ------------------------------------------------------------------ */

static int fft_log2_needed(int n, int k, int B) {
    (void)n;
    long v = 4L * (k > B ? k : B) + 8;
    return ceil_log2((size_t)v);
}

int orc_inner(int n, const double *a, const double *payout, int k,
              double *inner, int engine, int B, int fft_min,
              const unsigned char *active) {
    if (n < 1 || k < 1 || k > n || !a || !payout || !inner) return -1;
    if (engine == ORC_ENGINE_NAIVE) {
        engine_naive(n, a, payout, k, inner, active);
    } else if (engine == ORC_ENGINE_LINEAR) {
        engine_linear(n, a, payout, k, inner, active);
    } else if (engine == ORC_ENGINE_HYBRID) {
        if (B < 1) return -1;
        if (fft_min > 0) tw_prepare(fft_log2_needed(n, k, B));
        Hybrid h;
        hybrid_init(&h, n, k, B, fft_min);
        if (active) hybrid_hot_mask(&h, n, active);
        engine_hybrid(&h, n, a, payout, k, inner, active);
        hybrid_free(&h);
    } else {
        return -1;
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* Integration loop: src/cpu/icm.c:2523-2611                          */
/* ------------------------------------------------------------------ */

typedef struct { double S; int idx; } SI;
static int cmp_si(const void *x, const void *y) {
    const SI *a = (const SI *)x, *b = (const SI *)y;
    if (a->S > b->S) return -1;
    if (a->S < b->S) return 1;
    return (a->idx < b->idx) ? -1 : (a->idx > b->idx);
}

int orc_equity(int n, const double *S, int Q, const double *payout, int k,
               double *equity, const int *targets, int n_targets,
               int engine, int B, int fft_min, int q_lo, int q_hi,
               int nthreads) {
    if (n < 1 || k < 1 || k > n || Q < 1 || !S || !payout || !equity) return -1;
    if (engine == ORC_ENGINE_HYBRID && B < 1) return -1;
    if (q_lo < 0) q_lo = 0;
    if (q_hi > Q) q_hi = Q;

    /* Sorted order (hybrid engine sorts; the others run in input order, as
     * the reference's linear/naive engines do). */
    int *perm = (int *)malloc((size_t)n * sizeof(int));
    if (engine == ORC_ENGINE_HYBRID) {
        SI *si = (SI *)malloc((size_t)n * sizeof(SI));
        for (int i = 0; i < n; i++) { si[i].S = S[i]; si[i].idx = i; }
        qsort(si, (size_t)n, sizeof(SI), cmp_si);
        for (int i = 0; i < n; i++) perm[i] = si[i].idx;
        free(si);
    } else {
        for (int i = 0; i < n; i++) perm[i] = i;
    }
    double *Ss = (double *)malloc((size_t)n * sizeof(double));
    for (int i = 0; i < n; i++) Ss[i] = S[perm[i]];

    unsigned char *active = NULL;
    if (targets) {
        unsigned char *act_orig = (unsigned char *)calloc((size_t)n, 1);
        for (int i = 0; i < n_targets; i++)
            if (targets[i] >= 0 && targets[i] < n) act_orig[targets[i]] = 1;
        active = (unsigned char *)calloc((size_t)n, 1);
        for (int i = 0; i < n; i++) active[i] = act_orig[perm[i]];
        free(act_orig);
    }

    double Smax = 0;
    for (int i = 0; i < n; i++) if (S[i] > Smax) Smax = S[i];
    double *logv = (double *)malloc((size_t)Q * sizeof(double));
    double *wq = (double *)malloc((size_t)Q * sizeof(double));
    orc_make_nodes(Q, Smax, logv, wq);
    memset(equity, 0, (size_t)n * sizeof(double));

    if (engine == ORC_ENGINE_HYBRID && fft_min > 0) tw_prepare(fft_log2_needed(n, k, B));

    int nt = 1;
#ifdef _OPENMP
    nt = (nthreads > 0) ? nthreads : omp_get_max_threads();
#endif
    if (nt > (q_hi - q_lo)) nt = (q_hi - q_lo) > 0 ? (q_hi - q_lo) : 1;
    double *eq_t = (double *)calloc((size_t)nt * n, sizeof(double));

#ifdef _OPENMP
#pragma omp parallel num_threads(nt)
#endif
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        double *a = (double *)malloc((size_t)n * sizeof(double));
        double *inner = (double *)malloc((size_t)n * sizeof(double));
        double *eql = eq_t + (size_t)tid * n;
        Hybrid h;
        int have_h = 0;
        if (engine == ORC_ENGINE_HYBRID) {
            hybrid_init(&h, n, k, B, fft_min);
            if (active) hybrid_hot_mask(&h, n, active);
            have_h = 1;
        }
        int span = q_hi - q_lo;
        int per = (span + nt - 1) / nt;
        int qs = q_lo + tid * per, qe = qs + per;
        if (qe > q_hi) qe = q_hi;
        for (int q = qs; q < qe; q++) {
            if (wq[q] == 0) continue;
            double lv = logv[q];
            for (int j = 0; j < n; j++) {
                double arg = Ss[j] * lv;
                a[j] = (arg < -700) ? 0 : exp(arg);
            }
            if (engine == ORC_ENGINE_HYBRID)
                engine_hybrid(&h, n, a, payout, k, inner, active);
            else if (engine == ORC_ENGINE_LINEAR)
                engine_linear(n, a, payout, k, inner, active);
            else
                engine_naive(n, a, payout, k, inner, active);
            double inv_v = exp(-lv);
            for (int i = 0; i < n; i++) {
                double pw = wq[q] * Ss[i] * a[i] * inv_v;
                if (!isfinite(pw)) pw = 0;
                eql[i] += pw * inner[i];
            }
        }
        if (have_h) hybrid_free(&h);
        free(a);
        free(inner);
    }
    for (int t = 0; t < nt; t++)
        for (int i = 0; i < n; i++) equity[perm[i]] += eq_t[(size_t)t * n + i];
    if (active) {
        /* non-targets are exactly zero (icm.c:2531 zeroes, inactive skip) */
        for (int i = 0; i < n; i++) if (!active[i]) equity[perm[i]] = 0.0;
    }
    free(eq_t);
    free(logv);
    free(wq);
    free(Ss);
    free(perm);
    free(active);
    return 0;
}
