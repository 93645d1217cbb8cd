// Multigrid solve of the POISSON system (1d.f:3529-3543) of a substep, one
// workgroup per case: V(2,2)-cycles with red-black Gauss-Seidel smoothing,
// Galerkin coarse operators P^T A P (linear interpolation, coarse points at
// the even fine rows, zero beyond the top row as in the reference's implicit
// p = 0 above the last face) and an exact elimination on the coarsest level.
//   MG64:    FP64 cycles from a zero initial guess (IOLDV=0, lica.in:8) until
//            the backward error max|g - A p| / (||A|| max|p| + max|g|) <= rtol;
//   MIXED:   FP64 defect correction: the residual and the update are FP64, the
//            correction comes from one FP32 V-cycle (FP32 matrices, smoother,
//            transfers; the coarsest solve in FP64);
//   LEARNED: MIXED from the initial guess of the convolutional network.
// The solution replaces P(1:N) as the direct solve does.
#pragma once
#include "lica_kern_common.h"

#define MG_COARSE 32       // levels stop at this many rows
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
// doubles of the per-case scratch: FP64 a,b,c,x,f,r per level, then the
// FP32 a,b,c,x,f,r per level (packed two to a double)
HOSTDEV long long mg_scratch_doubles(int n)
{
    long long tot = 0;
    const int L = mg_levels(n);
    for (int l = 0; l < L; l++) tot += 6LL * mg_rows(n, l) + 3LL * mg_rows(n, l) + 8;
    return tot;
}

template <class T>
struct MgLevel {
    T *a, *b, *c, *x, *f, *r;
    int n;
};

// level l of the FP64 (T=double) or FP32 (T=float) hierarchy
template <class T>
DEV MgLevel<T> mg_level(double *base, int n, int l)
{
    const int L = mg_levels(n);
    long long off64 = 0, off32 = 0;
    for (int i = 0; i < L; i++) off32 += 6LL * mg_rows(n, i);
    long long o64 = 0, o32 = 0;
    for (int i = 0; i < l; i++) { o64 += 6LL * mg_rows(n, i); o32 += 6LL * mg_rows(n, i); }
    const int m = mg_rows(n, l);
    MgLevel<T> v;
    v.n = m;
    if constexpr (sizeof(T) == 8) {
        T *p = (T *)(base + o64);
        v.a = p; v.b = p + m; v.c = p + 2 * m; v.x = p + 3 * m; v.f = p + 4 * m; v.r = p + 5 * m;
    } else {
        T *p = (T *)(base + off32) + o32;
        v.a = p; v.b = p + m; v.c = p + 2 * m; v.x = p + 3 * m; v.f = p + 4 * m; v.r = p + 5 * m;
    }
    (void)off64;
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

// Galerkin coarse operator of level l+1 from level l (FP64)
DEV void mg_coarsen(const MgLevel<double> &F, const MgLevel<double> &C)
{
    const int nf = F.n, nc = C.n;
    PFOR(I, 0, nc) {
        double acc[3] = {0.0, 0.0, 0.0};
        for (int t = 0; t < 3; t++) {
            const int i = 2 * I - 1 + t;
            const double wi = mg_pw(i, I, nf, nc);
            if (wi == 0.0) continue;
            const double ai = (i == 0) ? 0.0 : F.a[i], ci = (i == nf - 1) ? 0.0 : F.c[i], bi = F.b[i];
            for (int d = -1; d <= 1; d++) {
                const int J = I + d;
                double apj = ai * mg_pw(i - 1, J, nf, nc) + bi * mg_pw(i, J, nf, nc) + ci * mg_pw(i + 1, J, nf, nc);
                acc[d + 1] += wi * apj;
            }
        }
        C.a[I] = (I == 0) ? 0.0 : acc[0];
        C.b[I] = acc[1];
        C.c[I] = (I == nc - 1) ? 0.0 : acc[2];
    }
    PFOR_END
}

template <class T>
DEV void mg_smooth(const MgLevel<T> &L, int sweeps)
{
    const int n = L.n;
    for (int s = 0; s < sweeps; s++) {
        for (int color = 0; color < 2; color++) {
            PFOR(h, 0, (n - color + 1) / 2) {
                const int i = 2 * h + color;
                T v = L.f[i];
                if (i > 0) v = v - L.a[i] * L.x[i - 1];
                if (i < n - 1) v = v - L.c[i] * L.x[i + 1];
                L.x[i] = v / L.b[i];
            }
            PFOR_END
        }
    }
}

template <class T>
DEV void mg_residual(const MgLevel<T> &L)
{
    const int n = L.n;
    PFOR(i, 0, n) {
        T v = L.f[i] - L.b[i] * L.x[i];
        if (i > 0) v = v - L.a[i] * L.x[i - 1];
        if (i < n - 1) v = v - L.c[i] * L.x[i + 1];
        L.r[i] = v;
    }
    PFOR_END
}

// f_c = P^T r_f, x_c = 0
template <class T>
DEV void mg_restrict(const MgLevel<T> &F, const MgLevel<T> &C)
{
    PFOR(J, 0, C.n) {
        const int i = 2 * J;
        T v = F.r[i];
        if (i - 1 >= 0) v = v + (T)0.5 * F.r[i - 1];
        if (i + 1 < F.n) v = v + (T)0.5 * F.r[i + 1];
        C.f[J] = v;
        C.x[J] = (T)0.0;
    }
    PFOR_END
}

// x_f += P x_c
template <class T>
DEV void mg_prolong(const MgLevel<T> &F, const MgLevel<T> &C)
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

// exact solve on the coarsest level (FP64 elimination by one thread)
template <class T>
DEV void mg_coarsest(const MgLevel<T> &L, double *ws)
{
    SERIAL {
        const int n = L.n;
        double *gam = ws;
        double bet = 1.0 / (double)L.b[0];
        double prev = (double)L.f[0] * bet;
        ws[n] = prev;
        for (int j = 1; j < n; j++) {
            gam[j] = (double)L.c[j - 1] * bet;
            bet = 1.0 / ((double)L.b[j] - (double)L.a[j] * gam[j]);
            prev = ((double)L.f[j] - (double)L.a[j] * prev) * bet;
            ws[n + j] = prev;
        }
        for (int j = n - 2; j >= 0; j--) ws[n + j] = ws[n + j] - gam[j + 1] * ws[n + j + 1];
        for (int j = 0; j < n; j++) L.x[j] = (T)ws[n + j];
    }
    SERIAL_END
}

template <class T>
DEV void mg_vcycle(double *base, int n, double *ws)
{
    const int L = mg_levels(n);
    for (int l = 0; l + 1 < L; l++) {
        MgLevel<T> F = mg_level<T>(base, n, l), C = mg_level<T>(base, n, l + 1);
        mg_smooth<T>(F, MG_NU);
        mg_residual<T>(F);
        mg_restrict<T>(F, C);
    }
    mg_coarsest<T>(mg_level<T>(base, n, L - 1), ws);
    for (int l = L - 2; l >= 0; l--) {
        MgLevel<T> F = mg_level<T>(base, n, l), C = mg_level<T>(base, n, l + 1);
        mg_prolong<T>(F, C);
        mg_smooth<T>(F, MG_NU);
    }
}

// normwise backward error max|g - A p| / (||A|| max|p| + max|g|) on the FP64
// fine level (residual recomputed).  max|g - A p| / max|g| alone has a
// rounding floor of about eps ||A|| max|p| / max|g|, which for the POISSON rows
// of a substep reaches 1e-10 (the reference's direct solve lands there too),
// so the stopping test is on the backward error.
DEV double mg_relres(const MgLevel<double> &F, double fmax, double anorm, double *red)
{
    mg_residual<double>(F);
    double rmax = red_max(0, F.n, 0.0, [&](int i) { return fabs(F.r[i]); }, red);
    double xmax = red_max(0, F.n, 0.0, [&](int i) { return fabs(F.x[i]); }, red);
    double den = anorm * xmax + fmax;
    return (den > 0.0) ? rmax / den : rmax;
}

// Solve rows 1..n of a, b, c, g (a(1), c(n) ignored) into p(1..n).
// mode 2: FP64 cycles, 3/4: FP32 corrections in FP64 defect correction
// (x0 holds the initial guess for mode 4).  Returns cycles; *res the final
// relative residual; cycles = -1 when maxc cycles did not reach rtol.
DEV int mg_solve(const double *a, const double *b, const double *c, const double *g, double *p, int n,
                 double *base, int mode, double rtol, int maxc, const double *x0, double *ws, double *red,
                 double *res_out)
{
    const int L = mg_levels(n);
    MgLevel<double> F = mg_level<double>(base, n, 0);
    PFOR(i, 0, n) {
        F.a[i] = (i == 0) ? 0.0 : a[i + 1];
        F.b[i] = b[i + 1];
        F.c[i] = (i == n - 1) ? 0.0 : c[i + 1];
        F.f[i] = g[i + 1];
        F.x[i] = (mode == LICA_MG_LEARNED_MODE && x0) ? x0[i + 1] : 0.0;
    }
    PFOR_END
    for (int l = 0; l + 1 < L; l++) mg_coarsen(mg_level<double>(base, n, l), mg_level<double>(base, n, l + 1));
    if (mode != LICA_MG_FP64_MODE) {
        for (int l = 0; l < L; l++) {
            MgLevel<double> D = mg_level<double>(base, n, l);
            MgLevel<float> S = mg_level<float>(base, n, l);
            PFOR(i, 0, D.n) { S.a[i] = (float)D.a[i]; S.b[i] = (float)D.b[i]; S.c[i] = (float)D.c[i]; }
            PFOR_END
        }
    }
    const double fmax = red_max(0, n, 0.0, [&](int i) { return fabs(F.f[i]); }, red);
    const double anorm = red_max(0, n, 0.0, [&](int i) { return fabs(F.a[i]) + fabs(F.b[i]) + fabs(F.c[i]); }, red);
    double res = mg_relres(F, fmax, anorm, red);
    int cyc = 0;
    while (res > rtol && cyc < maxc) {
        if (mode == LICA_MG_FP64_MODE) {
            mg_vcycle<double>(base, n, ws);
        } else {
            // one FP32 correction of the FP64 residual held in F.r
            MgLevel<float> S = mg_level<float>(base, n, 0);
            PFOR(i, 0, n) { S.f[i] = (float)F.r[i]; S.x[i] = 0.0f; }
            PFOR_END
            mg_vcycle<float>(base, n, ws);
            PFOR(i, 0, n) { F.x[i] = F.x[i] + (double)S.x[i]; }
            PFOR_END
        }
        cyc++;
        res = mg_relres(F, fmax, anorm, red);
    }
    PFOR(i, 0, n) { p[i + 1] = F.x[i]; }
    PFOR_END
    *res_out = res;
    return (res > rtol) ? -1 : cyc;
}
