// Multigrid solve of the POISSON system (1d.f:3529-3543) of a substep, one
// workgroup per case.
//
// The V-cycle is a matrix-dependent ("ideal") two-level scheme applied
// recursively: the odd rows of a level are its fine (F) points and the even
// rows its coarse (C) points; the coarse operator is the Schur complement of
// the F rows (Galerkin with the operator-dependent interpolation
// P_F = -A_FF^-1 A_FC), the restriction is R = [I, -A_CF A_FF^-1], and after
// the coarse correction the F rows are relaxed exactly.  On a tridiagonal
// operator that cycle is exact in exact arithmetic (it is cyclic reduction
// written as a V-cycle), so its contraction per cycle is set by rounding
// alone.
//
// The POISSON rows are a nearly singular M-matrix: b = -(a + c) rounded, so a
// row sums to the rounding error of that addition, except the last row,
// which carries the Dirichlet coefficient of the implicit p = 0 above the top
// face.  Computing a coarse pivot as b - a c'/b' - c a'/b' cancels almost all
// of its digits.  The setup instead carries every level as (a, c, s, 1/b)
// with s the row sum -- exact on the finest level by error-free additions --
// and forms a coarse row as s' = s - (a/b_l) s_l - (c/b_u) s_u,
// a' = -(a/b_l) a_l, c' = -(c/b_u) c_u, b' = (s' - a') - c': every sum then
// adds terms of one sign (Grassmann-Taksar-Heyman elimination), so each
// level is accurate to a few ulps of its own entries, in FP64 and in the FP32
// copy the mixed-precision cycles read.  The coarsest level (<= MG_COARSE
// rows) is eliminated the same way in FP64.
//
// Every mode is defect correction x += V(g - A x) in FP64 with the residual
// in double-double (error-free products and sums, rounded once); the cycle V
// runs in
//   MG64:      FP64 (IOLDV=0, lica.in:8: the iteration starts from x = 0),
//   MIXED:     FP32 (the coarsest solve in FP64),
//   LEARNED:   FP32, started from the warm-start network's guess,
//   PROJECTED: FP32, started from the projection on the solution history.
// A solve ends when the normwise backward error
// max|g - A x| / (||A|| max|x| + max|g|) <= rtol AND the last correction
// max|dx| <= rtol max|x|.  The second test bounds the forward error: the
// cycle contracts the error by 1e-3 or more per cycle in FP32 (to rounding in
// FP64), so the error left after a correction of relative size rtol is far
// below rtol.  A starting point is never accepted without one cycle.
// Non-finite residuals or iterates never pass a test (their maxima read as
// infinity), a non-finite guess is replaced by zero, and a system with
// non-finite rows yields a non-finite solution, as the direct solve would.
//
// Memory: setup rows of the coarse levels, the pair weights of every level
// (FP64 or FP32) and the iterate sit in a per-case global scratch; the
// finest level's setup is fused with the first residual pass (one read of
// the rows for both); the cycle's
// vectors -- level 0's right-hand side r and correction e, and one vector per
// coarser level (its f on the way down, its x on the way up) -- live in LDS
// from the coarsest level up as far as the launch's dynamic LDS allows (the
// rest in the scratch).
#pragma once
#include "lica_kern_common.h"

#ifndef MG_COARSE
#define MG_COARSE 32       // levels stop at this many rows
#endif
#define MG_MAXLEV 20
#define MG_NRED 5         // values reduced at once (mg_correct), per wave of up to 16

// rows of level l for n fine rows
HOSTDEV int mg_rows(int n, int l)
{
    int m = n;
    for (int i = 0; i < l; i++) m = (m + 1) / 2;
    return m;
}
HOSTDEV int mg_levels(int n)
{
    int l = 1, m = n;
    while (m > MG_COARSE && l < MG_MAXLEV) { m = (m + 1) / 2; l++; }
    return l;
}
// rows of levels 0..l-1
HOSTDEV long long mg_rows_before(int n, int l)
{
    long long s = 0;
    int m = n;
    for (int i = 0; i < l; i++) { s += m; m = (m + 1) / 2; }
    return s;
}
HOSTDEV long long mg_rows_total(int n) { return mg_rows_before(n, mg_levels(n)); }

// Per-case scratch, in doubles (every block 32-byte aligned when the case's
// scratch is); a level's rows pair up as (C row 2J, F row 2J+1), one pair per
// row J of the next level, tot - n pairs in all:
//   4 (tot - n)  FP64 setup rows (a, c, s, 1/b) of levels 1..L-1
//   2 (tot - n)  C weights (w0, w1) of every pair (FP64, or FP32 in its first half)
//   4 (tot - n)  room for the F weights (w0, w1, w2) of every pair (likewise)
//   2 n          the iterate and its double buffer
//   n + tot      the cycle's vectors kept outside LDS (room for FP64)
// (the C and F weights are apart so that the downward pass, which reads
// only the C weights, and the upward pass, which reads only the F weights,
// each pull only the bytes they use)
HOSTDEV long long mg_ru4(long long x) { return (x + 3) / 4 * 4; }
HOSTDEV long long mg_scratch_doubles(int n)
{
    const long long tot = mg_rows_total(n), np = tot - n;
    return mg_ru4(4 * np + mg_ru4(2 * np) + 4 * np + mg_ru4(2LL * n) + (n + tot) + 8);
}
// LDS that holds every vector of a cycle in elements of tsize bytes
HOSTDEV long long mg_lds_bytes(int n, int tsize) { return (long long)tsize * (n + mg_rows_total(n)); }

struct MgSh {
    double red[16 * MG_NRED];                              // per-wave partials: 16 waves x MG_NRED values
    double cm[MG_COARSE], cid[MG_COARSE], cc[MG_COARSE];   // coarsest level: multipliers, 1/pivots, c
    double cy[MG_COARSE];                                  // its forward sweep
    void *lds;                                             // the cycle's vectors (dynamic LDS)
    long long lds_bytes;
    // level table of the current solve (filled once: the loops that give a
    // level's rows and offset would otherwise run before every pass)
    long long lv_pre[MG_MAXLEV + 1];                       // rows of the levels before l
    int lv_n[MG_MAXLEV];                                   // rows of level l
    int nlev;
};

// fill the level table for n fine rows (one thread, then a barrier)
DEV void mg_table(int n, MgSh &sh)
{
    SERIAL {
        const int L = mg_levels(n);
        long long pre = 0;
        int m = n;
        for (int l = 0; l < L; l++) {
            sh.lv_n[l] = m;
            sh.lv_pre[l] = pre;
            pre += m;
            m = (m + 1) / 2;
        }
        sh.lv_pre[L] = pre;
        sh.nlev = L;
    }
    SERIAL_END
}

// setup row of a level: off-diagonals, row sum, 1/diagonal; the weights of
// a pair of rows of a level:
//   C row i = 2J:    f_c(J) = f_i + w0 f_{i-1} + w1 f_{i+1}   (w0 = -a_i/b_{i-1}, w1 = -c_i/b_{i+1})
//   F row k = 2J+1:  x_k = w0 f_k + w1 x_{k-1} + w2 x_{k+1}   (w0 = 1/b_k, w1 = -a_k/b_k, w2 = -c_k/b_k)
#ifdef LICA_EMUL
struct MgSet {
    double a, c, s, ib;
};
template <class T>
struct MgCw {
    T w0, w1;
};
template <class T>
struct MgFw {
    T w0, w1, w2;
};
#else
struct alignas(32) MgSet {
    double a, c, s, ib;
};
template <class T>
struct alignas(2 * sizeof(T)) MgCw {
    T w0, w1;
};
// three weights, no padding: the upward pass reads 12 (FP32) or 24 bytes
// per pair instead of 16 or 32
template <class T>
struct MgFw {
    T w0, w1, w2;
};
#endif

// row i (0..n-1) of the finest level from the POISSON rows k = i+1 (a(1) and
// c(n) are not part of the system): the row sum exactly, by error-free
// additions, rounded once
// b(k) of the rows: b null means the POISSON rows' b = -(a + c), evaluated
// as the momentum stage evaluated it (the same value as the stored one,
// without reading it)
DEV double mg_b(const double *a, const double *b, const double *c, int k)
{
    return b ? b[k] : -(a[k] + c[k]);
}

DEV MgSet mg_row0(const double *a, const double *b, const double *c, int n, int i)
{
    const double ai = (i == 0) ? 0.0 : a[i + 1], bi = mg_b(a, b, c, i + 1), ci = (i == n - 1) ? 0.0 : c[i + 1];
    double t, e1, s, e2;
    two_sum(ai, ci, &t, &e1);
    two_sum(t, bi, &s, &e2);
    MgSet r;
    r.a = ai; r.c = ci; r.s = s + (e1 + e2); r.ib = 1.0 / bi;
    return r;
}

struct MgView {
    double *scr;
    int n;
    long long tot;
    MgSet *set;                 // setup rows of levels >= 1 (level l row i at lv_pre[l] - n + i)
    MgCw<double> *cw64;         // weights of the pairs of level l at lv_pre[l+1] - n + J
    MgFw<double> *fw64;
    MgCw<float> *cw32;          // the same memory: a solve uses one precision
    MgFw<float> *fw32;
    double *x0, *x1;            // iterate buffers
    double *gv;                 // global home of the cycle's vectors
};

DEV MgView mg_view(double *scr, int n, const MgSh &sh)
{
    MgView v;
    v.scr = scr;
    v.n = n;
    v.tot = sh.lv_pre[sh.nlev];
    const long long np = v.tot - n;
    double *q = scr;
    v.set = (MgSet *)q;             q += 4 * np;
    v.cw64 = (MgCw<double> *)q;
    v.cw32 = (MgCw<float> *)q;      q += mg_ru4(2 * np);
    v.fw64 = (MgFw<double> *)q;
    v.fw32 = (MgFw<float> *)q;      q += 4 * np;
    v.x0 = q;
    v.x1 = q + n;                   q += mg_ru4(2LL * n);
    v.gv = q;
    return v;
}

// The cycle's vectors in one index space of n + tot elements: level 0's r at
// [0, n), its e at [n, 2n), level l >= 1 at [n + lv_pre[l], n + lv_pre[l+1]).
// The tail of that space from Q = max(0, n + tot - LDS elements) lives in LDS;
// a vector that starts below Q stays in global memory.
template <class T>
DEV T *mg_vec(const MgView &V, const MgSh &sh, long long start)
{
    const long long cap = sh.lds_bytes / (long long)sizeof(T);
    const long long q0 = (V.n + V.tot > cap) ? V.n + V.tot - cap : 0;
    if (start >= q0) return (T *)sh.lds + (start - q0);
    return (T *)V.gv + start;
}
template <class T>
DEV T *mg_v(const MgView &V, const MgSh &sh, int l) { return mg_vec<T>(V, sh, V.n + sh.lv_pre[l]); }
template <class T>
DEV T *mg_r0(const MgView &V, const MgSh &sh) { return mg_vec<T>(V, sh, 0); }
template <class T>
DEV T *mg_e0(const MgView &V, const MgSh &sh) { return mg_vec<T>(V, sh, V.n); }

template <class T>
DEV MgCw<T> *mg_cw(const MgView &V, const MgSh &sh, int l)
{
    if constexpr (sizeof(T) == 8) return V.cw64 + (sh.lv_pre[l + 1] - V.n);
    else return V.cw32 + (sh.lv_pre[l + 1] - V.n);
}
template <class T>
DEV MgFw<T> *mg_fw(const MgView &V, const MgSh &sh, int l)
{
    if constexpr (sizeof(T) == 8) return V.fw64 + (sh.lv_pre[l + 1] - V.n);
    else return V.fw32 + (sh.lv_pre[l + 1] - V.n);
}

// One pair (C row i = 2J, F row i+1) of a level: the coarse row J in
// row-sum form and the pair's weights (rows of the level from row(k))
template <class T, class RowF>
DEV MgSet mg_pair(RowF row, int i, int nl, MgCw<T> &cw, MgFw<T> &fw, bool &has_f)
{
    const MgSet R = row(i);
    double sv = R.s, av = 0.0, cv = 0.0, w0 = 0.0, w1 = 0.0;
    if (i >= 1) {
        const MgSet Lo = row(i - 1);
        const double q = R.a * Lo.ib;
        av = -q * Lo.a;
        sv = sv - q * Lo.s;
        w0 = -q;
    }
    has_f = (i + 1 < nl);
    if (has_f) {
        const MgSet Up = row(i + 1);
        const double q = R.c * Up.ib;
        cv = -q * Up.c;
        sv = sv - q * Up.s;
        w1 = -q;
        fw.w0 = (T)Up.ib; fw.w1 = (T)(-Up.a * Up.ib); fw.w2 = (T)(-Up.c * Up.ib);
    }
    cw.w0 = (T)w0; cw.w1 = (T)w1;
    MgSet C;
    C.a = av; C.c = cv; C.s = sv; C.ib = 1.0 / ((sv - av) - cv);
    return C;
}

// Setup of the levels below level 1 (level 1's rows come from mg_first):
// level l+1's rows and level l's pair weights for l = 1..L-2, then the
// coarsest level's elimination factors.
template <class T>
DEV void mg_setup_rest(const double *a, const double *b, const double *c, const MgView &V, MgSh &sh)
{
    const int L = sh.nlev, n = V.n;
    for (int l = 1; l + 1 < L; l++) {
        const int nl = sh.lv_n[l], nc = sh.lv_n[l + 1];
        const MgSet *fine = V.set + (sh.lv_pre[l] - n);
        MgSet *crs = V.set + (sh.lv_pre[l + 1] - n);
        MgCw<T> *cw = mg_cw<T>(V, sh, l);
        MgFw<T> *fw = mg_fw<T>(V, sh, l);
        PFOR(J, 0, nc) {
            MgCw<T> w;
            MgFw<T> f;
            bool has_f;
            crs[J] = mg_pair<T>([&](int k) { return fine[k]; }, 2 * J, nl, w, f, has_f);
            cw[J] = w;
            if (has_f) fw[J] = f;
        }
        PFOR_END
    }
    // coarsest level: Thomas elimination in row-sum form, sigma_j the row sum
    // of row j after the rows below it are eliminated, pivot sigma_j - c_j
    SERIAL {
        const int l = L - 1, m = sh.lv_n[l];
        const MgSet *rows = (l == 0) ? nullptr : V.set + (sh.lv_pre[l] - n);
        double sig = 0.0, piv = 1.0;
        for (int j = 0; j < m; j++) {
            const MgSet r = rows ? rows[j] : mg_row0(a, b, c, n, j);
            const double mu = (j == 0) ? 0.0 : r.a / piv;
            sig = r.s - mu * sig;
            piv = sig - r.c;
            sh.cm[j] = mu;
            sh.cid[j] = 1.0 / piv;
            sh.cc[j] = r.c;
        }
    }
    SERIAL_END
}

// exact FP64 solve on the coarsest level: f in, x out (may alias)
template <class T>
DEV void mg_coarsest(const T *f, T *x, int m, MgSh &sh)
{
    SERIAL {
        double y = (double)f[0];
        sh.cy[0] = y;
        for (int j = 1; j < m; j++) {
            y = (double)f[j] - sh.cm[j] * y;
            sh.cy[j] = y;
        }
        double v = sh.cy[m - 1] * sh.cid[m - 1];
        x[m - 1] = (T)v;
        for (int j = m - 2; j >= 0; j--) {
            v = (sh.cy[j] - sh.cc[j] * v) * sh.cid[j];
            x[j] = (T)v;
        }
    }
    SERIAL_END
}

// one cycle on level 0's r; the correction in level 0's e
template <class T>
DEV void mg_cycle(const MgView &V, MgSh &sh)
{
    const int L = sh.nlev;
    T *r0 = mg_r0<T>(V, sh), *e0 = mg_e0<T>(V, sh);
    if (L == 1) {
        mg_coarsest<T>(r0, e0, sh.lv_n[0], sh);
        return;
    }
    for (int l = 0; l + 1 < L; l++) {
        const int nl = sh.lv_n[l], nc = sh.lv_n[l + 1];
        const T *fin = (l == 0) ? r0 : mg_v<T>(V, sh, l);
        T *fout = mg_v<T>(V, sh, l + 1);
        const MgCw<T> *cw = mg_cw<T>(V, sh, l);
        PFOR(J, 0, nc) {
            const int i = 2 * J;
            const MgCw<T> w = cw[J];
            T v = fin[i];
            if (i >= 1) v = v + w.w0 * fin[i - 1];
            if (i + 1 < nl) v = v + w.w1 * fin[i + 1];
            fout[J] = v;
        }
        PFOR_END
    }
    {
        T *vc = mg_v<T>(V, sh, L - 1);
        mg_coarsest<T>(vc, vc, sh.lv_n[L - 1], sh);
    }
    for (int l = L - 2; l >= 0; l--) {
        const int nl = sh.lv_n[l], nc = sh.lv_n[l + 1];
        const T *xc = mg_v<T>(V, sh, l + 1);
        const T *fx = (l == 0) ? r0 : mg_v<T>(V, sh, l);
        T *xo = (l == 0) ? e0 : mg_v<T>(V, sh, l);
        const MgFw<T> *fw = mg_fw<T>(V, sh, l);
        PFOR(J, 0, nc) {
            const T xl = xc[J];
            xo[2 * J] = xl;
            const int k = 2 * J + 1;
            if (k < nl) {
                const MgFw<T> w = fw[J];
                const T xr = (J + 1 < nc) ? xc[J + 1] : (T)0;
                xo[k] = w.w0 * fx[k] + w.w1 * xl + w.w2 * xr;
            }
        }
        PFOR_END
    }
}

DEV double mg_mag(double v) { return isfinite(v) ? fabs(v) : INFINITY; }

// x_new = x + e (e absent on the first pass; xa null: x = 0) into xb,
// r = g - A x_new in double-double into level 0's r; mx = {max|r|,
// max|x_new|, max|e|, max|g|, ||A||_inf} (non-finite values as infinity; the
// last two on the first pass only)
template <class T>
DEV void mg_correct(const double *a, const double *b, const double *c, const double *g, const MgView &V,
                    const double *xa, double *xb, bool first, MgSh &sh, double *mx)
{
    const int n = V.n;
    T *r0 = mg_r0<T>(V, sh);
    const T *e0 = mg_e0<T>(V, sh);
    auto xv = [&](int k) -> double { return xa ? xa[k] : 0.0; };
    auto xn = [&](int k) -> double { return first ? xv(k) : xv(k) + (double)e0[k]; };
    red_maxN<MG_NRED>(0, n, [&](int i, double *v) {
        const double ai = (i == 0) ? 0.0 : a[i + 1], bi = mg_b(a, b, c, i + 1), ci = (i == n - 1) ? 0.0 : c[i + 1];
        const double x = xn(i), xl = (i > 0) ? xn(i - 1) : 0.0, xr = (i < n - 1) ? xn(i + 1) : 0.0;
        const double r = dd_residual(g[i + 1], ai, xl, bi, x, ci, xr);
        xb[i] = x;
        r0[i] = (T)r;
        v[0] = mg_mag(r);
        v[1] = mg_mag(x);
        v[2] = first ? 0.0 : mg_mag((double)e0[i]);
        v[3] = first ? mg_mag(g[i + 1]) : 0.0;
        v[4] = first ? mg_mag(fabs(ai) + fabs(bi) + fabs(ci)) : 0.0;
    }, sh.red, mx);
}

// The finest level's pass of the setup together with the first residual
// (one read of a, b, c, g and the starting point; needs L >= 2): level 1's
// rows and level 0's pair weights, r = g - A x for x = x0 (rows 1..n; zero
// when x0 is null) in double-double into level 0's r, and the maxima of
// mg_correct's first pass (the first correction reads x0 where it is)
template <class T>
DEV void mg_first(const double *a, const double *b, const double *c, const double *g, const MgView &V,
                  const double *x0, MgSh &sh, double *mx)
{
    const int n = V.n, nc = sh.lv_n[1];
    MgSet *crs = V.set;
    MgCw<T> *cw = mg_cw<T>(V, sh, 0);
    MgFw<T> *fw = mg_fw<T>(V, sh, 0);
    T *r0 = mg_r0<T>(V, sh);
    auto xv = [&](int k) -> double { return x0 ? x0[k + 1] : 0.0; };
    red_maxN<MG_NRED>(0, nc, [&](int J, double *v) {
        MgCw<T> w;
        MgFw<T> f;
        bool has_f;
        crs[J] = mg_pair<T>([&](int k) { return mg_row0(a, b, c, n, k); }, 2 * J, n, w, f, has_f);
        cw[J] = w;
        if (has_f) fw[J] = f;
        for (int q = 0; q < 5; q++) v[q] = 0.0;
        for (int k = 2 * J; k <= 2 * J + 1 && k < n; k++) {
            const double ai = (k == 0) ? 0.0 : a[k + 1], bi = mg_b(a, b, c, k + 1), ci = (k == n - 1) ? 0.0 : c[k + 1];
            const double x = xv(k), xl = (k > 0) ? xv(k - 1) : 0.0, xr = (k < n - 1) ? xv(k + 1) : 0.0;
            const double r = dd_residual(g[k + 1], ai, xl, bi, x, ci, xr);
            r0[k] = (T)r;
            v[0] = fmax(v[0], mg_mag(r));
            v[1] = fmax(v[1], mg_mag(x));
            v[3] = fmax(v[3], mg_mag(g[k + 1]));
            v[4] = fmax(v[4], mg_mag(fabs(ai) + fabs(bi) + fabs(ci)));
        }
    }, sh.red, mx);
}

DEV double mg_berr(double rmax, double xmax, double gmax, double anorm)
{
    const double den = anorm * xmax + gmax;
    return (den > 0.0) ? rmax / den : rmax;
}

template <class T>
DEV int mg_run(const double *a, const double *b, const double *c, const double *g, const MgView &V, double rtol,
               int maxc, const double *x0, MgSh &sh, double **cur, double *res_out, bool *nonfinite_rows)
{
    double *xa = V.x0, *xb = V.x1;
    double mx[MG_NRED];
    // the starting point, rows 0..n-1 (null: zero): the first correction
    // reads it where it is
    const double *xs = x0 ? x0 + 1 : nullptr;
    if (sh.nlev >= 2) {
        mg_first<T>(a, b, c, g, V, x0, sh, mx);
        mg_setup_rest<T>(a, b, c, V, sh);
    } else {
        mg_setup_rest<T>(a, b, c, V, sh);
        mg_correct<T>(a, b, c, g, V, xs, xb, true, sh, mx);
    }
    const double gmax = mx[3], anorm = mx[4];
    *nonfinite_rows = !(gmax < INFINITY) || !(anorm < INFINITY);
    if (*nonfinite_rows) { *cur = xb; *res_out = NAN; return 0; }
    if (x0 && !(mx[1] < INFINITY && mx[0] < INFINITY)) {
        // a non-finite guess is dropped: start from zero
        xs = nullptr;
        mg_correct<T>(a, b, c, g, V, xs, xb, true, sh, mx);
    }
    // the corrections read the iterate xc and write the other buffer
    const double *xc = xs;
    double res = mg_berr(mx[0], mx[1], gmax, anorm);
    int cyc = 0;
    bool conv = false;
    while (cyc < maxc) {
        mg_cycle<T>(V, sh);
        mg_correct<T>(a, b, c, g, V, xc, xa, false, sh, mx);
        xc = xa;
        { double *t = xa; xa = xb; xb = t; }
        cyc++;
        res = mg_berr(mx[0], mx[1], gmax, anorm);
        if (!(mx[1] < INFINITY)) break;
        if (res <= rtol && mx[2] <= rtol * mx[1]) { conv = true; break; }
    }
    if (cyc == 0) {
        // no cycle allowed: the starting point is the result
        PFOR(i, 0, V.n) { xb[i] = xs ? xs[i] : 0.0; }
        PFOR_END
    }
    *cur = xb;   // the last correction's buffer
    *res_out = res;
    return conv ? cyc : -1;
}

// Solve rows 1..n of a, b, c, g (a(1), c(n) ignored) into p(1..n); b null:
// b = -(a + c) (the POISSON rows).
// mode 2: FP64 cycles; 3-5: FP32 cycles (x0, rows 1..n, holds the initial
// guess for modes 4 and 5).  Returns the cycles, or -1 when maxc cycles did
// not meet the stopping test; *res_out the final backward error.  Rows with
// non-finite entries give a non-finite p and 0 cycles.
DEV int mg_solve(const double *a, const double *b, const double *c, const double *g, double *p, int n,
                 double *scr, int mode, double rtol, int maxc, const double *x0, MgSh &sh, double *res_out)
{
    const bool guess = (LICA_MG_WARM(mode) && x0);
    mg_table(n, sh);
    const MgView V = mg_view(scr, n, sh);
    double *cur;
    bool bad = false;
    int cyc;
    if (mode == LICA_MG_FP64_MODE)
        cyc = mg_run<double>(a, b, c, g, V, rtol, maxc, guess ? x0 : nullptr, sh, &cur, res_out, &bad);
    else
        cyc = mg_run<float>(a, b, c, g, V, rtol, maxc, guess ? x0 : nullptr, sh, &cur, res_out, &bad);
    PFOR(i, 0, n) { p[i + 1] = bad ? NAN : cur[i]; }
    PFOR_END
    return cyc;
}
