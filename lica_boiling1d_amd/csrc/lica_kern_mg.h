// Multigrid solve of the POISSON system (1d.f:3529-3543) of a substep, one
// workgroup per case: V(2,2)-cycles with red-black Gauss-Seidel smoothing,
// Galerkin coarse operators P^T A P (linear interpolation, coarse points at
// the even fine rows, zero beyond the top row as in the reference's implicit
// p = 0 above the last face) and an exact FP64 elimination on the coarsest
// level.  Every mode is FP64 defect correction x += V(g - A x): the residual,
// the update and the stopping test are FP64, the V-cycle runs on the residual
// equation from a zero start in
//   MG64:    FP64 (IOLDV=0, lica.in:8: the iteration starts from x = 0),
//   MIXED:   FP32 (matrices, smoother, transfers; the coarsest solve in FP64),
//   LEARNED: FP32, the iteration started from the warm-start network's guess,
//   PROJECTED: FP32, started from the projection on the solution history,
// until the normwise backward error max|g - A x| / (||A|| max|x| + max|g|)
// reaches rtol.  The solution replaces P(1:N) as the direct solve does.
//
// Memory: the iterate, the fine matrix and the matrices of every level sit in
// a per-case global scratch (read-only during the cycles), one packed row
// (a, b, c, 1/b) per unknown so that a row is one wide load; x and f of the
// levels live in LDS from the coarsest level up as far as the launch's
// dynamic LDS allows (the rest in the scratch), so that the passes of a cycle
// -- each one barrier -- do not wait on global stores.
#pragma once
#include "lica_kern_common.h"

#ifndef MG_COARSE
#define MG_COARSE 32       // levels stop at this many rows
#endif
#define MG_MAXLEV 20
#define MG_NU 2            // pre- and post-smoothing sweeps

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
// scratch is):
//   4 tot      FP64 rows (a, b, c, 1/b) of every level, level 0 first
//   2 tot      FP32 rows of every level
//   4n         g, iterate, its double buffer (and one spare row)
//   2 tot      x, f of the levels kept outside LDS (room for FP64)
HOSTDEV long long mg_scratch_doubles(int n)
{
    const long long tot = mg_rows_total(n);
    return (4LL * tot + 2LL * tot + 4LL * n + 2LL * tot + 8 + 3) / 4 * 4;
}
// LDS that holds x, f of every level in elements of tsize bytes
HOSTDEV long long mg_lds_bytes(int n, int tsize) { return 2LL * tsize * mg_rows_total(n); }

struct MgSh {
    double red[64];                                        // reductions of up to four values
    double cg[MG_COARSE], cib[MG_COARSE], ca[MG_COARSE];   // coarsest-level elimination factors
    double cy[MG_COARSE];                                  // its forward sweep
    void *lds;                                             // x, f of the coarse levels
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

// one packed matrix row
#ifdef LICA_EMUL
template <class T>
struct Row4 {
    T a, b, c, ib;
};
#else
template <class T>
struct alignas(4 * sizeof(T)) Row4 {
    T a, b, c, ib;
};
#endif

template <class T>
struct MgLev {
    Row4<T> *m;   // rows (a, b, c, 1/b); written only while the levels are built
    T *x, *f;
    int n;
};
struct MgFine {
    Row4<double> *m;
    double *g, *x0, *x1;
    int n;
};

DEV MgFine mg_fine(double *scr, int n, const MgSh &sh)
{
    const long long tot = sh.lv_pre[sh.nlev];
    MgFine F;
    F.n = n;
    F.m = (Row4<double> *)scr;
    double *v = scr + 6 * tot;
    F.g = v; F.x0 = v + n; F.x1 = v + 2 * n;
    return F;
}

// rows of level l (FP64 or FP32); x, f in LDS when levels l..L-1 fit
template <class T>
DEV MgLev<T> mg_lev(double *scr, int n, int l, const MgSh &sh)
{
    const int m = sh.lv_n[l];
    const long long pre = sh.lv_pre[l], tot = sh.lv_pre[sh.nlev];
    MgLev<T> v;
    v.n = m;
    if constexpr (sizeof(T) == 8) v.m = (Row4<T> *)scr + pre;
    else v.m = (Row4<T> *)(scr + 4 * tot) + pre;
    const long long suffix = tot - pre;   // rows of levels l..L-1
    T *q;
    if ((long long)(2 * sizeof(T)) * suffix <= sh.lds_bytes) q = (T *)sh.lds + 2 * (suffix - m);
    else q = (T *)(scr + 6 * tot + 4LL * n) + 2 * pre;
    v.x = q;
    v.f = q + m;
    return v;
}

// interpolation weight of coarse point J at fine row k (fine rows 0..nf-1,
// coarse J <-> fine 2J; odd fine rows average their two coarse neighbours)
DEV double mg_pw(int k, int J, int nf, int nc)
{
    if (k < 0 || k >= nf || J < 0 || J >= nc) return 0.0;
    if (k == 2 * J) return 1.0;
    if (k == 2 * J - 1 || k == 2 * J + 1) return 0.5;
    return 0.0;
}

// Galerkin operator of level l+1 from level l (FP64), with 1/b, and the FP32
// copy of level l+1 when the cycles run in FP32
DEV void mg_coarsen(const MgLev<double> &F, const MgLev<double> &C, Row4<float> *c32, bool f32)
{
    const int nf = F.n, nc = C.n;
    PFOR(I, 0, nc) {
        double acc[3] = {0.0, 0.0, 0.0};
        for (int t = 0; t < 3; t++) {
            const int i = 2 * I - 1 + t;
            const double wi = mg_pw(i, I, nf, nc);
            if (wi == 0.0) continue;
            const Row4<double> ri = F.m[i];
            const double ai = ri.a, ci = ri.c, bi = ri.b;
            for (int d = -1; d <= 1; d++) {
                const int J = I + d;
                double apj = ai * mg_pw(i - 1, J, nf, nc) + bi * mg_pw(i, J, nf, nc) + ci * mg_pw(i + 1, J, nf, nc);
                acc[d + 1] += wi * apj;
            }
        }
        const double va = (I == 0) ? 0.0 : acc[0], vc = (I == nc - 1) ? 0.0 : acc[2], vb = acc[1];
        Row4<double> r;
        r.a = va; r.b = vb; r.c = vc; r.ib = 1.0 / vb;
        C.m[I] = r;
        if (f32) {
            Row4<float> q;
            q.a = (float)va; q.b = (float)vb; q.c = (float)vc; q.ib = 1.0f / (float)vb;
            c32[I] = q;
        }
    }
    PFOR_END
}

// red-black Gauss-Seidel; zero_start: x is taken as zero before the first
// half-sweep (its stale values are never read)
template <class T>
DEV void mg_smooth(const MgLev<T> &L, int sweeps, bool zero_start)
{
    const int n = L.n;
    for (int s = 0; s < sweeps; s++) {
        for (int color = 0; color < 2; color++) {
            const bool z = zero_start && s == 0 && color == 0;
            PFOR(h, 0, (n - color + 1) / 2) {
                const int i = 2 * h + color;
                const Row4<T> r = L.m[i];
                T v = L.f[i];
                if (!z) {
                    if (i > 0) v = v - r.a * L.x[i - 1];
                    if (i < n - 1) v = v - r.c * L.x[i + 1];
                }
                L.x[i] = v * r.ib;
            }
            PFOR_END
        }
    }
}

template <class T>
DEV T mg_res_at(const MgLev<T> &F, int i)
{
    const Row4<T> r = F.m[i];
    T v = F.f[i] - r.b * F.x[i];
    if (i > 0) v = v - r.a * F.x[i - 1];
    if (i < F.n - 1) v = v - r.c * F.x[i + 1];
    return v;
}

// f_c = P^T (f - A x) with the residual evaluated on the fly
template <class T>
DEV void mg_restrict(const MgLev<T> &F, const MgLev<T> &C)
{
    PFOR(J, 0, C.n) {
        const int i = 2 * J;
        T v = mg_res_at(F, i);
        if (i - 1 >= 0) v = v + (T)0.5 * mg_res_at(F, i - 1);
        if (i + 1 < F.n) v = v + (T)0.5 * mg_res_at(F, i + 1);
        C.f[J] = v;
    }
    PFOR_END
}

// x_f += P x_c
template <class T>
DEV void mg_prolong(const MgLev<T> &F, const MgLev<T> &C)
{
    PFOR(i, 0, F.n) {
        const int J = i / 2;
        T e;
        if (i % 2 == 0) e = C.x[J];
        else e = (T)0.5 * (C.x[J] + ((J + 1 < C.n) ? C.x[J + 1] : (T)0.0));
        F.x[i] = F.x[i] + e;
    }
    PFOR_END
}

// elimination factors of the coarsest level (FP64, once per solve)
DEV void mg_coarsest_factor(const MgLev<double> &L, MgSh &sh)
{
    SERIAL {
        const int n = L.n;
        double ib = 1.0 / L.m[0].b;
        sh.cib[0] = ib;
        sh.ca[0] = 0.0;
        sh.cg[0] = 0.0;
        for (int j = 1; j < n; j++) {
            const Row4<double> r = L.m[j];
            const double g = L.m[j - 1].c * ib;
            ib = 1.0 / (r.b - r.a * g);
            sh.cg[j] = g;
            sh.cib[j] = ib;
            sh.ca[j] = r.a;
        }
    }
    SERIAL_END
}

// exact solve on the coarsest level in FP64 (one thread)
template <class T>
DEV void mg_coarsest(const MgLev<T> &L, MgSh &sh)
{
    SERIAL {
        const int n = L.n;
        double y = (double)L.f[0] * sh.cib[0];
        sh.cy[0] = y;
        for (int j = 1; j < n; j++) {
            y = ((double)L.f[j] - sh.ca[j] * y) * sh.cib[j];
            sh.cy[j] = y;
        }
        double x = sh.cy[n - 1];
        L.x[n - 1] = (T)x;
        for (int j = n - 2; j >= 0; j--) {
            x = sh.cy[j] - sh.cg[j + 1] * x;
            L.x[j] = (T)x;
        }
    }
    SERIAL_END
}

// one V-cycle on the residual equation held in level 0's f; result in its x
template <class T>
DEV void mg_vcycle(double *scr, int n, MgSh &sh)
{
    const int L = sh.nlev;
    for (int l = 0; l + 1 < L; l++) {
        MgLev<T> F = mg_lev<T>(scr, n, l, sh), C = mg_lev<T>(scr, n, l + 1, sh);
        mg_smooth<T>(F, MG_NU, true);
        mg_restrict<T>(F, C);
    }
    mg_coarsest<T>(mg_lev<T>(scr, n, L - 1, sh), sh);
    for (int l = L - 2; l >= 0; l--) {
        MgLev<T> F = mg_lev<T>(scr, n, l, sh), C = mg_lev<T>(scr, n, l + 1, sh);
        mg_prolong<T>(F, C);
        mg_smooth<T>(F, MG_NU, false);
    }
}

// xn = x + e (e = Z.x, absent on the first pass) into the other buffer,
// r = g - A xn in FP64, Z.f = r in T; *rmax = max|r|, *xmax = max|xn|
template <class T>
DEV void mg_update(const MgFine &F, const MgLev<T> &Z, bool first, const double *xa, double *xb, MgSh &sh,
                   double *rmax, double *xmax)
{
    const int n = F.n;
    auto xn = [&](int k) -> double { return first ? xa[k] : xa[k] + (double)Z.x[k]; };
    red_max2(0, n, [&](int i, double &ra, double &xa_) {
        const Row4<double> r = F.m[i];
        const double xi = xn(i);
        double v = F.g[i] - r.b * xi;
        if (i > 0) v = v - r.a * xn(i - 1);
        if (i < n - 1) v = v - r.c * xn(i + 1);
        xb[i] = xi;
        Z.f[i] = (T)v;
        ra = fabs(v);
        xa_ = fabs(xi);
    }, sh.red, rmax, xmax);
}

// residual of the starting point; when it already meets rtol the solve ends
// there, before any level is built
template <class T>
DEV double mg_start(const MgFine &F, int n, double fmax, double anorm, bool warm, MgSh &sh, double **cur, double **oth,
                    double *scr)
{
    MgLev<T> Z = mg_lev<T>(scr, n, 0, sh);
    double rmax, xmax;
    mg_update<T>(F, Z, true, *cur, *oth, sh, &rmax, &xmax);
    { double *t = *cur; *cur = *oth; *oth = t; }
    if (warm && rmax > fmax) {
        // a starting guess with a larger residual than zero's is dropped
        PFOR(i, 0, n) { (*cur)[i] = 0.0; }
        PFOR_END
        mg_update<T>(F, Z, true, *cur, *oth, sh, &rmax, &xmax);
        { double *t = *cur; *cur = *oth; *oth = t; }
    }
    const double den = anorm * xmax + fmax;
    return (den > 0.0) ? rmax / den : rmax;
}

// V-cycles until rtol; *cur is the iterate on return (the update alternates
// the two buffers)
template <class T>
DEV int mg_iterate(const MgFine &F, double *scr, int n, double rtol, int maxc, double fmax, double anorm, MgSh &sh,
                   double **cur, double **oth, double res, double *res_out)
{
    MgLev<T> Z = mg_lev<T>(scr, n, 0, sh);
    double rmax, xmax;
    int cyc = 0;
    while (res > rtol && cyc < maxc) {
        mg_vcycle<T>(scr, n, sh);
        mg_update<T>(F, Z, false, *cur, *oth, sh, &rmax, &xmax);
        { double *t = *cur; *cur = *oth; *oth = t; }
        cyc++;
        const double den = anorm * xmax + fmax;
        res = (den > 0.0) ? rmax / den : rmax;
    }
    *res_out = res;
    return (res > rtol) ? -1 : cyc;
}

// Solve rows 1..n of a, b, c, g (a(1), c(n) ignored) into p(1..n).
// mode 2: FP64 cycles; 3-5: FP32 cycles (x0 holds the initial guess for
// modes 4 and 5).  The levels are built only when the starting point misses
// rtol.  Returns the cycles, or -1 when maxc cycles did not reach rtol; *res
// the final backward error.
DEV int mg_solve(const double *a, const double *b, const double *c, const double *g, double *p, int n,
                 double *scr, int mode, double rtol, int maxc, const double *x0, MgSh &sh, double *res_out)
{
    const bool guess = (LICA_MG_WARM(mode) && x0);
    if (guess) {
        // a warm start usually meets rtol as it is: its backward error from the
        // rows themselves, before anything is packed (the maxima of the packed
        // path below, so the same decision and, when it is taken, p = x0)
        double mx[4];   // max|g|, max(|a|+|b|+|c|), max|r|, max|x0|
        red_maxN<4>(0, n, [&](int i, double *v) {
            const double ai = (i == 0) ? 0.0 : a[i + 1], bi = b[i + 1], ci = (i == n - 1) ? 0.0 : c[i + 1];
            const double xi = x0[i + 1];
            double r = g[i + 1] - bi * xi;
            if (i > 0) r = r - ai * x0[i];
            if (i < n - 1) r = r - ci * x0[i + 2];
            v[0] = fabs(g[i + 1]);
            v[1] = fabs(ai) + fabs(bi) + fabs(ci);
            v[2] = fabs(r);
            v[3] = fabs(xi);
        }, sh.red, mx);
        const double den = mx[1] * mx[3] + mx[0];
        const double res0 = (den > 0.0) ? mx[2] / den : mx[2];
        if (!(mx[2] > mx[0]) && !(res0 > rtol)) {
            PFOR(i, 0, n) { p[i + 1] = x0[i + 1]; }
            PFOR_END
            *res_out = res0;
            return 0;
        }
    }
    mg_table(n, sh);
    const int L = sh.nlev;
    const bool f32 = (mode != LICA_MG_FP64_MODE);
    MgFine F = mg_fine(scr, n, sh);
    Row4<float> *m32 = mg_lev<float>(scr, n, 0, sh).m;
    double fmax, anorm;
    red_max2(0, n, [&](int i, double &fm, double &an) {
        const double ai = (i == 0) ? 0.0 : a[i + 1], bi = b[i + 1], ci = (i == n - 1) ? 0.0 : c[i + 1];
        Row4<double> r;
        r.a = ai; r.b = bi; r.c = ci; r.ib = 1.0 / bi;
        F.m[i] = r;
        F.g[i] = g[i + 1];
        F.x0[i] = guess ? x0[i + 1] : 0.0;
        if (f32) {
            Row4<float> q;
            q.a = (float)ai; q.b = (float)bi; q.c = (float)ci; q.ib = 1.0f / (float)bi;
            m32[i] = q;
        }
        fm = fabs(g[i + 1]);
        an = fabs(ai) + fabs(bi) + fabs(ci);
    }, sh.red, &fmax, &anorm);
    double *cur = F.x0, *oth = F.x1, res;
    if (f32) res = mg_start<float>(F, n, fmax, anorm, guess, sh, &cur, &oth, scr);
    else res = mg_start<double>(F, n, fmax, anorm, guess, sh, &cur, &oth, scr);
    int cyc = 0;
    if (res > rtol) {
        for (int l = 0; l + 1 < L; l++)
            mg_coarsen(mg_lev<double>(scr, n, l, sh), mg_lev<double>(scr, n, l + 1, sh),
                       mg_lev<float>(scr, n, l + 1, sh).m, f32);
        mg_coarsest_factor(mg_lev<double>(scr, n, L - 1, sh), sh);
        if (f32) cyc = mg_iterate<float>(F, scr, n, rtol, maxc, fmax, anorm, sh, &cur, &oth, res, &res);
        else cyc = mg_iterate<double>(F, scr, n, rtol, maxc, fmax, anorm, sh, &cur, &oth, res, &res);
    }
    PFOR(i, 0, n) { p[i + 1] = cur[i]; }
    PFOR_END
    *res_out = res;
    return (res > rtol) ? -1 : cyc;
}
