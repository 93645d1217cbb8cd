// HIP kernels of the per-step path (gfx950).  One workgroup per case.
//
// Kernel sequence of a time step: k_dt (DETERMINE_DT, 1d.f:201), then per
// RK3 substep (1d.f:220-246):
//   k_cont_energy    CAL_CONTINUITY, CAL_CONT_BC,
//                    energy RHS + LHSPS/LHSPS2 solve           (block)
//   k_temp_grad      TEMP_GRAD incl. TEMP_EXTP, TF_GRAD, VOL_TOT_MFDOT
//                    (steps after the first)                   (one wave)
//   k_momentum       CAL_CONTINUITY2, RHSZ, LHSW, BC, FSM_CHOI, DIVGS and
//                    the POISSON matrix                        (block)
//   k_poisson        POISSON solve (1d.f:3492-3578) or multigrid (block)
//   k_ucalc          UCALC, p-=p(1), CONVRGE1, VEL_INTP's UT/UO  (block)
//   k_transport      TRANSPORT + GRID_COUPLING(0)              (one wave)
// The narrow band, the refined level-set window and the temperature
// extrapolation window live in LDS; NS fields stream from HBM.
#include "lica_par.h"
#include "lica_types.h"
#include <type_traits>
#pragma once

#ifdef LICA_EMUL
#include <cstdio>
#endif

#define WMAX 192       // refined window (z-planes) held in LDS
#define NSWMAX 128     // NS window of the temperature extrapolation

struct CaseP {
    const InstParam *P;
    double *ns, *rf;
    int *nsi, *rfi;
    DEV double *A(int f) const { return ns + (size_t)f * P->ln + 2; }
    DEV double *R(int f) const { return rf + (size_t)f * P->lf - P->m3l; }
    DEV int *AI(int f) const { return nsi + (size_t)f * P->lni + 2; }
    DEV int *RI(int f) const { return rfi + (size_t)f * P->lfi; }
};

DEV CaseP case_ptrs(const BatchDev &bd, int inst)
{
    CaseP c;
    c.P = &bd.par[inst];
    c.ns = bd.ns + c.P->ns_off;
    c.rf = bd.rf + c.P->rf_off;
    c.nsi = bd.nsi + c.P->nsi_off;
    c.rfi = bd.rfi + c.P->rfi_off;
    return c;
}

// ---------------------------------------------------------------------
// scalar building blocks (two_phase.f:2448-2500, 2869-2920)

DEV double d_hjweno(double a, double b, double c, double d)
{
    double eps = 1.0e-6;
    double s1 = 13.0 * ((a - b) * (a - b)) + 3.0 * ((a - 3.0 * b) * (a - 3.0 * b));
    double s2 = 13.0 * ((b - c) * (b - c)) + 3.0 * ((b + c) * (b + c));
    double s3 = 13.0 * ((c - d) * (c - d)) + 3.0 * ((3.0 * c - d) * (3.0 * c - d));
    double a1 = 1.0 / ((eps + s1) * (eps + s1));
    double a2 = 6.0 / ((eps + s2) * (eps + s2));
    double a3 = 3.0 / ((eps + s3) * (eps + s3));
    double w1 = a1 / (a1 + a2 + a3);
    double w3 = a3 / (a1 + a2 + a3);
    return 1.0 / 3.0 * w1 * (a - 2.0 * b + c) + 1.0 / 6.0 * (w3 - 0.5) * (b - 2.0 * c + d);
}

DEV double d_godunov(double asign, double adlphip, double adlphim)
{
    double aplus = asign * adlphip, aminus = asign * adlphim, r = 0.0;
    if (aplus >= 0.0 && aminus >= 0.0) r = adlphim;
    else if (aplus <= 0.0 && aminus <= 0.0) r = adlphip;
    else if (aplus > 0.0 && aminus < 0.0) r = 0.0;
    else if (aplus < 0.0 && aminus > 0.0) r = (aplus + aminus > 0.0) ? adlphim : adlphip;
    return r;
}

// liquid fraction of one refined cell of an x/y-flat column; -1 = bug branch
DEV double d_psif_cell(double phi, double d3)
{
    const double eps = 1.0e-8;
    double d1 = 0.0, d2 = 0.0;
    double e1 = fmax(fmax(d1, d2), d3);
    double e3 = fmin(fmin(d1, d2), d3);
    double e2 = d1 + d2 + d3 - e1 - e3;
    double t;
    if (e1 > eps && e2 > eps && e3 > eps) {
        double a = fabs(phi);
        double f1 = fmax(0.5 * (e1 + e2 + e3) - a, 0.0), f2 = fmax(0.5 * (e1 + e2 - e3) - a, 0.0);
        double f3 = fmax(0.5 * (e1 - e2 + e3) - a, 0.0), f4 = fmax(0.5 * (-e1 + e2 + e3) - a, 0.0);
        double f5 = fmax(0.5 * (e1 - e2 - e3) - a, 0.0);
        t = (f1 * f1 * f1 - f2 * f2 * f2 - f3 * f3 * f3 - f4 * f4 * f4 + f5 * f5 * f5) / (6.0 * e1 * e2 * e3);
    } else if (e1 > eps && e2 > eps && e3 <= eps) {
        double a = fabs(phi);
        double f1 = fmax(0.5 * (e1 + e2 + e3) - a, 0.0), f3 = fmax(0.5 * (e1 - e2 + e3) - a, 0.0);
        t = (f1 * f1 - f3 * f3) / (2.0 * e1 * e2);
    } else if (e1 > eps && e2 <= eps && e3 <= eps) {
        double f1 = fmax(0.5 * (e1 + e2 + e3) - fabs(phi), 0.0);
        t = f1 / e1;
    } else if (e1 <= eps && e2 <= eps && e3 <= eps && phi != 0.0) {
        t = 0.0;
    } else if (e1 <= eps && e2 <= eps && e3 <= eps && phi == 0.0) {
        t = 0.5;
    } else {
        return -1.0;
    }
    return (phi <= 0.0) ? t : 1.0 - t;
}

DEV double snap5(double v) { return (v < 1.0e-5) ? 0.0 : ((v > 0.99999) ? 1.0 : v); }
DEV double snap3(double v) { return (v < 1.0e-3) ? 0.0 : ((v > 0.999) ? 1.0 : v); }

// box average of a refined field over [z1,z2] (INTER_PSI* family with the
// x/y loops of the single NS column collapsed); zf(k) and f(k) are accessors
template <class Z, class F>
DEV double box_avg_acc(Z zf, double ssdzf, F f, double z1, double z2, int k1, int k2)
{
    double sum = 0.0, volf = 0.0;
    for (int kk = k1; kk <= k2; kk++) {
        double zl = zf(kk), zh = zf(kk + 1);
        if (zl <= z1 && zh >= z2) { sum = sum + f(kk); volf = volf + 1.0; break; }
        else if (zl < z1 && zh > z1) { sum = sum + f(kk) * (zh - z1) * ssdzf; volf = volf + (zh - z1) * ssdzf; }
        else if (zl >= z1 && zh <= z2) { sum = sum + f(kk); volf = volf + 1.0; }
        else if (zl < z2 && zh > z2) { sum = sum + f(kk) * (z2 - zl) * ssdzf; volf = volf + (z2 - zl) * ssdzf; }
        else if (zl >= z2) break;
    }
    return sum / volf;
}

// A window [lo, hi] of refined z nodes and of up to two refined fields staged
// in LDS, so that the dependent loads of the box averages near the interface
// are LDS loads; indices outside the window read global memory.
#define RWMAX 384
struct RefWin {
    double zf[RWMAX + 1];
    double f0[RWMAX], f1[RWMAX];
    int lo, hi;           // staged refined cells; hi < lo: nothing staged
};

// stage [lo, hi] (clamped to [klo, khi]) of zf and of f0/f1 (either may be null)
DEV void stage_refwin(RefWin &w, int lo, int hi, int klo, int khi, const double *zf, const double *f0,
                      const double *f1)
{
    if (lo < klo) lo = klo;
    if (hi > khi) hi = khi;
    if (hi - lo + 1 > RWMAX || hi < lo) { lo = 1; hi = 0; }
    PFOR(i, 0, hi - lo + 2) {
        w.zf[i] = zf[lo + i];
        if (i <= hi - lo) {
            if (f0) w.f0[i] = f0[lo + i];
            if (f1) w.f1[i] = f1[lo + i];
        }
    }
    PFOR_END
    SERIAL { w.lo = lo; w.hi = hi; }
    SERIAL_END
}

template <class F>
DEV double box_avg(const double *zf, double ssdzf, F f, double z1, double z2, int k1, int k2)
{
    double sum = 0.0, volf = 0.0;
    for (int kk = k1; kk <= k2; kk++) {
        double zl = zf[kk], zh = zf[kk + 1];
        if (zl <= z1 && zh >= z2) { sum = sum + f(kk); volf = volf + 1.0; break; }
        else if (zl < z1 && zh > z1) { sum = sum + f(kk) * (zh - z1) * ssdzf; volf = volf + (zh - z1) * ssdzf; }
        else if (zl >= z1 && zh <= z2) { sum = sum + f(kk); volf = volf + 1.0; }
        else if (zl < z2 && zh > z2) { sum = sum + f(kk) * (z2 - zl) * ssdzf; volf = volf + (z2 - zl) * ssdzf; }
        else if (zl >= z2) break;
    }
    return sum / volf;
}

// ---------------------------------------------------------------------
// Block-level tridiagonal solve of rows l1..l2: a(j)x(j-1)+b(j)x(j)+c(j)x(j+1)=g(j),
// a(l1) and c(l2) ignored; the solution overwrites g.
//   mode 0: TRDIAG3 (1d.f:3783-3812) on one thread, the reference's order;
//   mode 1: partitioned elimination.  Every thread reduces a chunk of rows
//           to its two end rows (Gaussian elimination inside the chunk, the
//           chunk's rows held in registers when they fit), the end-row
//           system -- tridiagonal in the order s0,e0,s1,e1,... -- is reduced
//           the same way level by level in LDS down to TRI_SERIAL rows,
//           which one thread eliminates, and the levels are substituted
//           back.  The Schur complements of a diagonally dominant matrix
//           stay so, and elimination (unlike cyclic reduction) keeps
//           TRDIAG3's accuracy on the Laplacian-like rows of POISSON.

#define TRI_RPT 8        // rows per thread held in registers at the first level
#define TRI_LCHUNK 8     // rows per chunk of the LDS levels (at most)
#define TRI_SERIAL 16    // end-row systems this small are eliminated by one thread

// rows of the next LDS level of an end-row system of m rows: two end rows
// of each of its ceil(m / TRI_LCHUNK) chunks
HOSTDEV int tri_next(int m) { return 2 * ((m + TRI_LCHUNK - 1) / TRI_LCHUNK); }

#ifndef TRI_CHUNK_ROWS
#define TRI_CHUNK_ROWS 8
#endif
// chunks of the partitioned solve of nrow >= 4 rows: TRI_CHUNK_ROWS rows
// each while that needs at most 1024 chunks, else TRI_RPT rows each (at
// least two chunks).  The partition depends on the row count alone, so a
// case's result does not depend on the workgroup size; a workgroup with
// fewer threads than chunks eliminates serially instead (the host gives the
// kernels that solve systems tri_chunks(nmax) threads or more).
HOSTDEV int tri_chunks(int nrow)
{
    int nc = (nrow + TRI_CHUNK_ROWS - 1) / TRI_CHUNK_ROWS;
    if (nc > 1024) nc = (nrow + TRI_RPT - 1) / TRI_RPT;
    if (nc < 2) nc = 2;
    if (nc > nrow / 2) nc = nrow / 2;
    return nc;
}

// rows of the LDS workspace (4 arrays) the end-row levels of nthreads chunks need
HOSTDEV int tri_ws_rows(int nthreads)
{
    int m = 2 * nthreads, tot = 0;
    while (m > TRI_SERIAL) {
        tot += m;
        m = tri_next(m);
    }
    return tot + m;
}

struct TriSh {
    double *ws;    // LDS: ra | rb | rc | rg, ws_rows each
    int ws_rows;
};

DEV void chunk_bounds(int nrow, int nc, int ch, int *s, int *L)
{
    const int base = nrow / nc, rem = nrow % nc;
    *s = ch * base + (ch < rem ? ch : rem);
    *L = base + (ch < rem ? 1 : 0);
}

// reduce rows s..s+L-1 (L >= 2) in place; r = start row (a,b,c,g), end row (a,b,c,g)
// (the four arrays never overlap: the compiler may keep a row it has just
// written in registers and load the next one early)
DEV void chunk_reduce(double *__restrict__ a, double *__restrict__ b, double *__restrict__ c, double *__restrict__ g,
                      int s, int L, double *__restrict__ r)
{
    for (int j = 2; j <= L - 1; j++) {
        int i = s + j;
        double q = a[i] / b[i - 1];
        a[i] = -q * a[i - 1];
        b[i] = b[i] - q * c[i - 1];
        g[i] = g[i] - q * g[i - 1];
    }
    if (L >= 3) {
        int i = s + L - 2;
        double inv = 1.0 / b[i];
        a[i] = a[i] * inv;
        c[i] = c[i] * inv;
        g[i] = g[i] * inv;
        for (int j = L - 3; j >= 1; j--) {
            i = s + j;
            double cc = c[i];
            inv = 1.0 / b[i];
            double ah = (a[i] - cc * a[i + 1]) * inv;
            double ch2 = -(cc * c[i + 1]) * inv;
            double gh = (g[i] - cc * g[i + 1]) * inv;
            a[i] = ah; c[i] = ch2; g[i] = gh;
        }
        double cs = c[s];
        r[1] = b[s] - cs * a[s + 1];
        r[2] = -(cs * c[s + 1]);
        r[3] = g[s] - cs * g[s + 1];
    } else {
        r[1] = b[s]; r[2] = c[s]; r[3] = g[s];
    }
    r[0] = a[s];
    const int e = s + L - 1;
    r[4] = a[e]; r[5] = b[e]; r[6] = c[e]; r[7] = g[e];
}

DEV void chunk_recover(const double *__restrict__ a, const double *__restrict__ c, double *__restrict__ g, int s, int L,
                       double xs, double xe)
{
    for (int j = 1; j <= L - 2; j++) {
        int i = s + j;
        g[i] = g[i] - a[i] * xs - c[i] * xe;
    }
    g[s] = xs;
    g[s + L - 1] = xe;
}

// compile-time loop: f(std::integral_constant<int, J>) for J = 0..N-1, so that
// register arrays are only ever indexed by constants
template <int J, int N, class F>
DEV void static_for_impl(F &f)
{
    if constexpr (J < N) {
        f(std::integral_constant<int, J>{});
        static_for_impl<J + 1, N>(f);
    }
}
template <int N, class F>
DEV void static_for(F f) { static_for_impl<0, N>(f); }

// chunk_reduce / chunk_recover on a chunk held in registers (L <= TRI_RPT):
// the same operations in the same order, with every index a constant
DEV void chunk_reduce_reg(double (&a)[TRI_RPT], double (&b)[TRI_RPT], double (&c)[TRI_RPT], double (&g)[TRI_RPT],
                          int L, double (&r)[8])
{
    static_for<TRI_RPT>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        if constexpr (j >= 2) {
            if (j <= L - 1) {
                double q = a[j] / b[j - 1];
                a[j] = -q * a[j - 1];
                b[j] = b[j] - q * c[j - 1];
                g[j] = g[j] - q * g[j - 1];
            }
        }
    });
    if (L >= 3) {
        static_for<TRI_RPT>([&](auto jc) {
            constexpr int j = TRI_RPT - 1 - decltype(jc)::value;   // TRI_RPT-1 .. 0
            if constexpr (j >= 1 && j <= TRI_RPT - 2) {
                if (j == L - 2) {
                    double inv = 1.0 / b[j];
                    a[j] = a[j] * inv;
                    c[j] = c[j] * inv;
                    g[j] = g[j] * inv;
                } else if (j <= L - 3) {
                    double cc = c[j];
                    double inv = 1.0 / b[j];
                    double ah = (a[j] - cc * a[j + 1]) * inv;
                    double ch2 = -(cc * c[j + 1]) * inv;
                    double gh = (g[j] - cc * g[j + 1]) * inv;
                    a[j] = ah; c[j] = ch2; g[j] = gh;
                }
            }
        });
        double cs = c[0];
        r[1] = b[0] - cs * a[1];
        r[2] = -(cs * c[1]);
        r[3] = g[0] - cs * g[1];
    } else {
        r[1] = b[0]; r[2] = c[0]; r[3] = g[0];
    }
    r[0] = a[0];
    static_for<TRI_RPT>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        if (j == L - 1) { r[4] = a[j]; r[5] = b[j]; r[6] = c[j]; r[7] = g[j]; }
    });
}

DEV void chunk_recover_reg(const double (&a)[TRI_RPT], const double (&c)[TRI_RPT], double (&g)[TRI_RPT], int L,
                           double xs, double xe)
{
    static_for<TRI_RPT>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        double v = g[j];
        if constexpr (j >= 1 && j <= TRI_RPT - 2) {
            double t = v - a[j] * xs - c[j] * xe;
            v = (j <= L - 2) ? t : v;
        }
        if constexpr (j >= 1) v = (j == L - 1) ? xe : v;
        if constexpr (j == 0) v = xs;
        g[j] = v;
    });
}

// end rows of chunk ch of nc into rows 2ch, 2ch+1 of the next level
DEV void put_end_rows(double *ra, double *rb, double *rc, double *rg, int ch, int nc, const double *r)
{
    ra[2 * ch] = (ch == 0) ? 0.0 : r[0];
    rb[2 * ch] = r[1];
    rc[2 * ch] = r[2];
    rg[2 * ch] = r[3];
    ra[2 * ch + 1] = r[4];
    rb[2 * ch + 1] = r[5];
    rc[2 * ch + 1] = (ch == nc - 1) ? 0.0 : r[6];
    rg[2 * ch + 1] = r[7];
}

// offset and size of level l of the end-row systems that start with m rows
DEV void lds_level(int m, int l, int *off, int *rows)
{
    int o = 0, cur = m;
    for (int i = 0; i < l; i++) { o += cur; cur = tri_next(cur); }
    *off = o;
    *rows = cur;
}

// the end-row system of m rows at row 0 of the LDS workspace; solution in
// rg.  Each level splits its rows into ceil(rows / TRI_LCHUNK) chunks of at
// most TRI_LCHUNK rows.
DEV void lds_tri_solve(TriSh &ts, int m)
{
    double *RA = ts.ws, *RB = ts.ws + ts.ws_rows, *RC = ts.ws + 2 * ts.ws_rows, *RG = ts.ws + 3 * ts.ws_rows;
    int nlev = 0;
    for (int cur = m; cur > TRI_SERIAL; cur = tri_next(cur)) nlev++;
    for (int l = 0; l < nlev; l++) {
        int bo, rows;
        lds_level(m, l, &bo, &rows);
        const int nc = (rows + TRI_LCHUNK - 1) / TRI_LCHUNK, no = bo + rows;
        PFOR(ch, 0, nc) {
            int s, L;
            chunk_bounds(rows, nc, ch, &s, &L);
            double r[8];
            chunk_reduce(RA + bo, RB + bo, RC + bo, RG + bo, s, L, r);
            put_end_rows(RA + no, RB + no, RC + no, RG + no, ch, nc, r);
        }
        PFOR_END
    }
    int o, cur;
    lds_level(m, nlev, &o, &cur);
    SERIAL {
        double *__restrict__ ra = RA + o, *__restrict__ rb = RB + o, *__restrict__ rc = RC + o,
               *__restrict__ rg = RG + o;
        double bet = 1.0 / rb[0];
        rg[0] = rg[0] * bet;
        for (int j = 1; j < cur; j++) {
            double gam = rc[j - 1] * bet;
            bet = 1.0 / (rb[j] - ra[j] * gam);
            rg[j] = (rg[j] - ra[j] * rg[j - 1]) * bet;
            ra[j] = gam;
        }
        for (int j = cur - 2; j >= 0; j--) rg[j] = rg[j] - ra[j + 1] * rg[j + 1];
    }
    SERIAL_END
    for (int l = nlev - 1; l >= 0; l--) {
        int bo, rows;
        lds_level(m, l, &bo, &rows);
        const int nc = (rows + TRI_LCHUNK - 1) / TRI_LCHUNK, no = bo + rows;
        PFOR(ch, 0, nc) {
            int s, L;
            chunk_bounds(rows, nc, ch, &s, &L);
            chunk_recover(RA + bo, RC + bo, RG + bo, s, L, RG[no + 2 * ch], RG[no + 2 * ch + 1]);
        }
        PFOR_END
    }
}

// g - (a*xm + b*x + c*xp) with error-free products (fma) and sums, rounded once
DEV void two_sum(double a, double b, double *s, double *e)
{
    double t = a + b;
    double bb = t - a;
    *e = (a - (t - bb)) + (b - bb);
    *s = t;
}
DEV double dd_residual(double g, double a, double xm, double b, double x, double c, double xp)
{
    double p1 = a * xm, e1 = fma(a, xm, -p1);
    double p2 = b * x, e2 = fma(b, x, -p2);
    double p3 = c * xp, e3 = fma(c, xp, -p3);
    double hi, lo, t;
    two_sum(g, -p1, &hi, &lo);
    two_sum(hi, -p2, &hi, &t);
    lo = lo + t;
    two_sum(hi, -p3, &hi, &t);
    lo = lo + t;
    lo = lo - e1 - e2 - e3;
    return hi + lo;
}

#define TRI_NREF 2    // refinement steps of the partitioned solve

// TRDIAG3 (1d.f:3783-3812) on one thread, pivots stored in a
DEV void tri_thomas(double *a, const double *b, const double *c, double *g, int l1, int l2)
{
    SERIAL {
        double bet = 1.0 / b[l1];
        g[l1] = g[l1] * bet;
        for (int j = l1 + 1; j <= l2; j++) {
            double gam = c[j - 1] * bet;
            bet = 1.0 / (b[j] - a[j] * gam);
            g[j] = (g[j] - a[j] * g[j - 1]) * bet;
            a[j] = gam;
        }
        for (int j = l2 - 1; j >= l1; j--) g[j] = g[j] - a[j + 1] * g[j + 1];
    }
    SERIAL_END
}

#ifndef LICA_EMUL
// one partitioned pass of the register path: rhs (registers) -> solution
DEV void part_pass_reg(const double *A, const double *B, const double *C, double (&rhs)[TRI_RPT], bool act, int ch,
                       int nc, int L, int s, TriSh &ts)
{
    double *RA = ts.ws, *RB = ts.ws + ts.ws_rows, *RC = ts.ws + 2 * ts.ws_rows, *RG = ts.ws + 3 * ts.ws_rows;
    double xa[TRI_RPT], xb[TRI_RPT], xc[TRI_RPT];
    if (act) {
        static_for<TRI_RPT>([&](auto jc) {
            constexpr int j = decltype(jc)::value;
            xa[j] = (j < L) ? A[s + j] : 0.0;
            xb[j] = (j < L) ? B[s + j] : 1.0;
            xc[j] = (j < L) ? C[s + j] : 0.0;
        });
        double r[8];
        chunk_reduce_reg(xa, xb, xc, rhs, L, r);
        put_end_rows(RA, RB, RC, RG, ch, nc, r);
    }
    BARRIER();
    lds_tri_solve(ts, 2 * nc);
    if (act) chunk_recover_reg(xa, xc, rhs, L, RG[2 * ch], RG[2 * ch + 1]);
    BARRIER();
}
#endif

// NREF_MAX bounds the refinement steps at compile time: a kernel that never
// refines (0) carries none of the refinement's registers
template <int NREF_MAX = TRI_NREF>
DEV void tri_solve(double *a, double *b, double *c, double *g, int l1, int l2, int mode, TriSh &ts)
{
    const int nrow = l2 - l1 + 1;
    if (mode == 0 || nrow < 4) {
        tri_thomas(a, b, c, g, l1, l2);
        return;
    }
    const int nc = tri_chunks(nrow);
    if (nc > NTHREADS) {
        // fewer threads than chunks: the serial elimination (exact reference order)
        tri_thomas(a, b, c, g, l1, l2);
        return;
    }
    double *RA = ts.ws;
    const double *A = a + l1, *B = b + l1, *C = c + l1;
    double *G = g + l1;
    // The partitioned elimination is backward stable but, on the nearly
    // singular POISSON rows, its forward error grows with the chunk count;
    // in mode 1 TRI_NREF refinement steps with residuals in double-double
    // bring the solution back to (or below) TRDIAG3's error.  Mode 2 (the
    // diagonally dominant energy and momentum rows when multigrid takes the
    // POISSON rows) does without them.  a, b and c are left untouched; g
    // receives the solution.
    const int nref = (mode == 2) ? 0 : (NREF_MAX < TRI_NREF ? NREF_MAX : TRI_NREF);
#ifndef LICA_EMUL
    const int ch = (int)threadIdx.x;
    const bool act = ch < nc;
    int s = 0, L = 0;
    double x[TRI_RPT];
    if (act) {
        chunk_bounds(nrow, nc, ch, &s, &L);
        static_for<TRI_RPT>([&](auto jc) {
            constexpr int j = decltype(jc)::value;
            x[j] = (j < L) ? G[s + j] : 0.0;
        });
    }
    part_pass_reg(A, B, C, x, act, ch, nc, L, s, ts);
    for (int it = 0; it < nref; it++) {
        if (act) {
            RA[2 * ch] = x[0];
            static_for<TRI_RPT>([&](auto jc) {
                constexpr int j = decltype(jc)::value;
                if (j == L - 1) RA[2 * ch + 1] = x[j];
            });
        }
        BARRIER();
        double d[TRI_RPT];
        if (act) {
            const double xl = (ch > 0) ? RA[2 * ch - 1] : 0.0, xr = (ch < nc - 1) ? RA[2 * ch + 2] : 0.0;
            static_for<TRI_RPT>([&](auto jc) {
                constexpr int j = decltype(jc)::value;
                if (j < L) {
                    const int i = s + j;
                    double xm = xl, xp = xr;
                    if constexpr (j > 0) xm = x[j - 1];
                    if constexpr (j + 1 < TRI_RPT) { if (j != L - 1) xp = x[j + 1]; }
                    const double am = (i == 0) ? 0.0 : A[i];
                    const double cp = (i == nrow - 1) ? 0.0 : C[i];
                    d[j] = dd_residual(G[i], am, xm, B[i], x[j], cp, xp);
                } else {
                    d[j] = 0.0;
                }
            });
        }
        BARRIER();
        part_pass_reg(A, B, C, d, act, ch, nc, L, s, ts);
        if (act) {
            static_for<TRI_RPT>([&](auto jc) {
                constexpr int j = decltype(jc)::value;
                x[j] = x[j] + d[j];
            });
        }
    }
    if (act) {
        static_for<TRI_RPT>([&](auto jc) {
            constexpr int j = decltype(jc)::value;
            if (j < L) G[s + j] = x[j];
        });
    }
    BARRIER();
#else
    // the same passes over copies of the rows (CPU logic-check build)
    double *RB = ts.ws + ts.ws_rows, *RC = ts.ws + 2 * ts.ws_rows, *RG = ts.ws + 3 * ts.ws_rows;
    static double *wa = nullptr, *wb, *wc, *wx, *wd;
    static int cap = 0;
    if (nrow > cap) {
        delete[] wa;
        cap = nrow;
        wa = new double[5 * (size_t)cap];
        wb = wa + cap; wc = wb + cap; wx = wc + cap; wd = wx + cap;
    }
    auto pass = [&](double *rhs) {
        for (int i = 0; i < nrow; i++) { wa[i] = A[i]; wb[i] = B[i]; wc[i] = C[i]; }
        for (int ch = 0; ch < nc; ch++) {
            int s, L;
            chunk_bounds(nrow, nc, ch, &s, &L);
            double r[8];
            chunk_reduce(wa, wb, wc, rhs, s, L, r);
            put_end_rows(RA, RB, RC, RG, ch, nc, r);
        }
        lds_tri_solve(ts, 2 * nc);
        for (int ch = 0; ch < nc; ch++) {
            int s, L;
            chunk_bounds(nrow, nc, ch, &s, &L);
            chunk_recover(wa, wc, rhs, s, L, RG[2 * ch], RG[2 * ch + 1]);
        }
    };
    for (int i = 0; i < nrow; i++) wx[i] = G[i];
    pass(wx);
    for (int it = 0; it < nref; it++) {
        for (int i = 0; i < nrow; i++) {
            const double xm = (i == 0) ? 0.0 : wx[i - 1], xp = (i == nrow - 1) ? 0.0 : wx[i + 1];
            const double am = (i == 0) ? 0.0 : A[i], cp = (i == nrow - 1) ? 0.0 : C[i];
            wd[i] = dd_residual(G[i], am, xm, B[i], wx[i], cp, xp);
        }
        pass(wd);
        for (int i = 0; i < nrow; i++) wx[i] = wx[i] + wd[i];
    }
    for (int i = 0; i < nrow; i++) G[i] = wx[i];
#endif
}

// A tridiagonal solve whose rows come from a function instead of arrays:
// row(k, a, b, c, g) gives row k of l1..l2 (called once per row), out(k, x)
// receives its solution (after every row was read).  In mode 2 with one
// chunk per thread nothing goes through global memory: the rows are
// evaluated in the coalesced order, CR = NTHREADS / TRI_RPT chunks at a time
// (at most NTHREADS rows, one per thread), staged in LDS as slot j of chunk
// ch at j * CR + ch (so that a chunk's thread reads its rows without bank
// conflicts), and moved into the chunks' registers; the solve is then
// tri_solve's register path, operation for operation.  (Evaluating each
// chunk's rows in its own thread reads the rows' inputs lane-strided, 64 B
// apart, and the lines leave the caches before the next row of the chunk
// comes back to them: three to four times the bytes.)  The staging uses the
// LDS workspace, so row() must not read what the workspace overlays.
// Otherwise the rows go to the arrays a, b, c, g and tri_solve runs on them.
template <int NREF_MAX = TRI_NREF, class RowF, class OutF>
DEV void tri_solve_rows(double *a, double *b, double *c, double *g, int l1, int l2, int mode, TriSh &ts, RowF row,
                        OutF out)
{
    const int nrow = l2 - l1 + 1;
    const int nc = (nrow >= 4) ? tri_chunks(nrow) : 0;
    const bool fits = nc <= NTHREADS;
#ifndef LICA_EMUL
    const bool regs = (mode == 2 && nrow >= 4 && fits);
#else
    const bool regs = false;
#endif
    if (!regs) {
        PFOR(k, l1, l2 + 1) { row(k, a[k], b[k], c[k], g[k]); }
        PFOR_END
        tri_solve<NREF_MAX>(a, b, c, g, l1, l2, mode, ts);
        PFOR(k, l1, l2 + 1) { out(k, g[k]); }
        PFOR_END
        return;
    }
#ifndef LICA_EMUL
    double *RA = ts.ws, *RB = ts.ws + ts.ws_rows, *RC = ts.ws + 2 * ts.ws_rows, *RG = ts.ws + 3 * ts.ws_rows;
    const int ch = (int)threadIdx.x;
    const bool act = ch < nc;
    int s = 0, L = 0;
    double xa[TRI_RPT], xb[TRI_RPT], xc[TRI_RPT], x[TRI_RPT];
    if (act) chunk_bounds(nrow, nc, ch, &s, &L);
    {
        const int CR = NTHREADS / TRI_RPT;
        double *SA = ts.ws, *SB = SA + TRI_RPT * CR, *SC = SB + TRI_RPT * CR, *SG = SC + TRI_RPT * CR;
        // chunks 0..rem-1 hold base+1 rows, the others base (chunk_bounds)
        const int base = nrow / nc, rem = nrow % nc, big = rem * (base + 1);
        for (int c0 = 0; c0 < nc; c0 += CR) {
            const int c1 = (c0 + CR < nc) ? c0 + CR : nc;
            int r0, r1, ll;
            chunk_bounds(nrow, nc, c0, &r0, &ll);
            chunk_bounds(nrow, nc, c1 - 1, &r1, &ll);
            r1 += ll;
            const int i = r0 + (int)threadIdx.x;
            if (i < r1) {
                int cc, j;
                if (i < big) { cc = i / (base + 1); j = i - cc * (base + 1); }
                else { cc = rem + (i - big) / base; j = (i - big) - (cc - rem) * base; }
                double ra, rb, rc, rg;
                row(l1 + i, ra, rb, rc, rg);
                const int q = j * CR + (cc - c0);
                SA[q] = ra; SB[q] = rb; SC[q] = rc; SG[q] = rg;
            }
            BARRIER();
            if (ch >= c0 && ch < c1) {
                static_for<TRI_RPT>([&](auto jc) {
                    constexpr int j = decltype(jc)::value;
                    const int q = j * CR + (ch - c0);
                    if (j < L) {
                        xa[j] = SA[q]; xb[j] = SB[q]; xc[j] = SC[q]; x[j] = SG[q];
                    } else {
                        xa[j] = 0.0; xb[j] = 1.0; xc[j] = 0.0; x[j] = 0.0;
                    }
                });
            }
            BARRIER();
        }
    }
    if (act) {
        double r[8];
        chunk_reduce_reg(xa, xb, xc, x, L, r);
        put_end_rows(RA, RB, RC, RG, ch, nc, r);
    }
    BARRIER();
    lds_tri_solve(ts, 2 * nc);
    if (act) {
        chunk_recover_reg(xa, xc, x, L, RG[2 * ch], RG[2 * ch + 1]);
        static_for<TRI_RPT>([&](auto jc) {
            constexpr int j = decltype(jc)::value;
            if (j < L) out(l1 + s + j, x[j]);
        });
    }
    BARRIER();
#endif
}

// ---------------------------------------------------------------------
// Region classes of CAL_CONTINUITY / CAL_CONTINUITY2 (1d.f:2037-2056):
// 2 within two cells of a jump of PSI_CN, 1 pure liquid, 0 otherwise.
DEV int region_of(const double *pc, int n, int k)
{
    int j0 = k - 2 < 1 ? 1 : k - 2, j1 = k + 2 > n ? n : k + 2;
    for (int j = j0; j <= j1; j++)
        if ((pc[j + 1] - pc[j] != 0.0) || (pc[j] - pc[j - 1] != 0.0)) return 2;
    return (pc[k] == 1.0) ? 1 : 0;
}

// lowest and highest j in [0, n] with PSI_CN(j+1) != PSI_CN(j) (*jhi < 0:
// none), in one pass; a cell outside [jlo - 3, jhi + 3] is in region 0/1 and
// so is its upper neighbour
DEV void jump_range(const double *pc, int n, double *red, int *jlo, int *jhi)
{
    double u, v;
    red_max2(0, n + 1, [&](int j, double &a, double &b) {
        const bool c = pc[j + 1] != pc[j];
        a = c ? 1.0e9 - (double)j : 0.0;
        b = c ? (double)j + 1.0 : 0.0;
    }, red, &u, &v);
    *jhi = (int)v - 1;
    *jlo = (v > 0.0) ? (int)(1.0e9 - u) : (1 << 30);
}

// The workgroup's LDS: a few reductions' partials in the static part, and
// in the dynamic part (blk_dyn_bytes) the refined window of the box
// averages, which the partitioned solve's workspace overlays: a kernel
// stages the window again after its solve if it reads it there.  The
// overlay keeps a 512-thread workgroup of the momentum stage at 44 KB
// (three per CU) instead of 53 KB (two).
// The per-wave partials of the workgroup reductions: up to 16 waves (1024
// threads) x up to three values (red_max2, red_maxN<3>).
#define BLK_RED_SLOTS (16 * 3)
struct BlkSh {
    double red[BLK_RED_SLOTS];
    double v[16];
    int iv[16];
    TriSh tri;          // ws at the start of the dynamic LDS (tri_ws_rows(blockDim) x 4 doubles)
    RefWin *win;        // refined window of the box averages (dynamic LDS, under the solve's workspace)
};
// dynamic LDS of a BlkSh kernel with `threads` per workgroup (tri: it runs
// partitioned solves)
HOSTDEV long long blk_dyn_bytes(int threads, int tri)
{
    const long long win = ((long long)sizeof(RefWin) + 15) / 16 * 16;
    const long long ws = tri ? 32LL * tri_ws_rows(threads) : 0;
    return ws > win ? ws : win;
}

// DETERMINE_DT (1d.f:683-817) at the start of a step, with the per-step
// reset of RK3ZO (1d.f:206-218); the CFL maximum |W(k)+W(k+1)|*SSDZ(k) was
// recorded after the last projection of the previous step
DEV void body_dt(const BatchDev &bd, int inst, BlkSh &sh)
{
    InstState &S = bd.st[inst];
    if (S.status) return;
    CaseP cp = case_ptrs(bd, inst);
    const InstParam &P = *cp.P;
    const int n = P.n, n3 = P.n3;
    const double *PCN = cp.A(NS_PSI_CN), *SDZ = cp.A(NS_SDZ);
    // DETERMINE_DT: capillary limit over interface cells, CFL from the
    // max |W(k)+W(k+1)|*SSDZ(k) recorded after the last projection
    double dtsrc;
    if (P.time_st == 0.0) dtsrc = 1000000.0;
    else {
        dtsrc = red_min(1, n + 1, 1000.0, [&](int k) {
            if (fabs(PCN[k] - 0.5) < 0.5) {
                double dx1 = fmin(P.sdy1, SDZ[k]);
                double t1 = 0.5 * dx1 / sqrt(P.time_st / dx1 + P.time_gv * dx1);
                double dx2 = fmax(P.sdy1, SDZ[k]);
                double t2 = 0.5 * dx2 / sqrt(P.time_st / dx2 + P.time_gv * dx2);
                return fmin(t1, t2);
            }
            return 1000.0;
        }, sh.red);
    }
    SERIAL {
        S.m += 1;
        double dt_old = S.dt;
        double cflm = 0.5 * S.c3max * S.dt;
        S.cflm = cflm;
        double cfl_surf = S.dt / dtsrc;
        if (P.idtopt == 1) {
            double cfl_sum = 0.5 * (cflm + sqrt(cflm * cflm + 4.0 * cfl_surf));
            S.dt = S.dt * P.cflmax / cfl_sum;
        }
        if (S.dt / dt_old > 2.0) S.dt = dt_old * 1.2;
    }
    SERIAL_END
    double *RK3ZO = cp.A(NS_RK3ZO);
    PFOR(k, 0, n3 + 1) { RK3ZO[k] = 0.0; }
    PFOR_END
}

// ---------------------------------------------------------------------
// k_cont_energy: CAL_CONTINUITY (1d.f:1994-2345), CAL_CONT_BC (2961-3003),
// energy stage of PASSIVE_SCALAR (3630-3707) with CAL_RHSPS/CAL_RHSPS2 and
// LHSPS (4038-4497).
DEV void body_cont_energy(const BatchDev &bd, int inst, BlkSh &sh)
{
    InstState &S = bd.st[inst];
    if (S.status) return;
    CaseP cp = case_ptrs(bd, inst);
    const InstParam &P = *cp.P;
    const int n = P.n, n3 = P.n3, msub = bd.msub;
    double *__restrict__ W = cp.A(NS_W), *__restrict__ T = cp.A(NS_T), *__restrict__ PZN = cp.A(NS_PSI_ZN), *__restrict__ PCN = cp.A(NS_PSI_CN);
    double *__restrict__ DFZ = cp.A(NS_DF_Z), *__restrict__ VFZ = cp.A(NS_VF_Z), *__restrict__ DFC = cp.A(NS_DF_C);
    double *__restrict__ PZ = cp.A(NS_PSI_Z), *__restrict__ PC = cp.A(NS_PSI_C), *__restrict__ RHSPS = cp.A(NS_RHSPS);
    double *__restrict__ TA = cp.A(NS_TA), *__restrict__ TB = cp.A(NS_TB), *__restrict__ TC = cp.A(NS_TC), *__restrict__ TG = cp.A(NS_TG);
    const double *__restrict__ ZP = cp.A(NS_ZP), *__restrict__ Z = cp.A(NS_Z), *__restrict__ SSDZ = cp.A(NS_SSDZ), *__restrict__ VVDZ = cp.A(NS_VVDZ);
    const double *__restrict__ AKUV = cp.A(NS_AKUV), *__restrict__ CKUV = cp.A(NS_CKUV);
    const double *__restrict__ ZF = cp.R(RF_ZF), *__restrict__ PSIF = cp.R(RF_PSIF);
    const int *__restrict__ KCOU1 = cp.AI(NSI_KCOU1), *__restrict__ KCOU2 = cp.AI(NSI_KCOU2);
    const int *__restrict__ KCOUMP1 = cp.AI(NSI_KCOUMP1), *__restrict__ KCOUMP2 = cp.AI(NSI_KCOUMP2);

    // substep constants (1d.f:221-225), evaluated by every thread
    const double dt = S.dt;
    const long long m = S.m;
    const double alpha = 0.5 * (P.gamma[msub - 1] + P.ro[msub - 1]);
    const double dtconst = 2.0 * alpha * dt;
    SERIAL {
        S.msub = msub;
        S.alpha = alpha;
        S.acoef = alpha * dt;
        S.dtconst = dtconst;
        S.dtconsti = 1.0 / dtconst;
    }
    SERIAL_END

    // the donor boxes of region-2 cells lie within a few NS cells of the
    // jumps of PSI_CN: stage that stretch of ZF and PSIF in LDS
    int jlo, jhi;
    jump_range(PCN, n, sh.red, &jlo, &jhi);
    {
        int lo = 1, hi = 0;
        if (jhi >= 0) {
            const int ka = jlo - 5 < -1 ? -1 : jlo - 5, kb = jhi + 6 > n3 + 2 ? n3 + 2 : jhi + 6;
            lo = KCOU1[ka] - 2;
            hi = KCOU2[kb] + 2;
        }
        stage_refwin(*sh.win, lo, hi, P.m3l, P.n3f + P.r, ZF, PSIF, nullptr);
    }
    const RefWin &win = *sh.win;
    auto zfa = [&](int kk) { return (kk >= win.lo && kk <= win.hi + 1) ? win.zf[kk - win.lo] : ZF[kk]; };
    auto psa = [&](int kk) { return (kk >= win.lo && kk <= win.hi) ? win.f0[kk - win.lo] : PSIF[kk]; };
    // PSIFACE_Z / PSIFACE_C (1d.f:2349-2549) and the face densities
    auto psiface_z = [&](int k, double uu) {
        double z1, z2; int k1, k2;
        if (uu >= 0.0) { z1 = ZP[k] - uu * dtconst; z2 = ZP[k]; k1 = KCOUMP2[k - 1]; k2 = KCOUMP1[k]; }
        else { z1 = ZP[k]; z2 = ZP[k] - uu * dtconst; k1 = KCOUMP2[k]; k2 = KCOUMP1[k + 1]; }
        return box_avg_acc(zfa, P.ssdzf, psa, z1, z2, k1, k2);
    };
    auto psiface_c = [&](int k, double uu) {
        double z1, z2; int k1, k2;
        if (uu >= 0.0) { z1 = Z[k + 1] - uu * dtconst; z2 = Z[k + 1]; k1 = KCOU2[k]; k2 = KCOU1[k + 1]; }
        else { z1 = Z[k + 1]; z2 = Z[k + 1] - uu * dtconst; k1 = KCOU2[k + 1]; k2 = KCOU1[k + 2]; }
        return box_avg_acc(zfa, P.ssdzf, psa, z1, z2, k1, k2);
    };
    // face value computed by a region-2 iteration for the z face at ZP(f)
    auto zface2 = [&](int f, double *vf) {
        double wt = 0.5 * (W[f + 1] + W[f]);
        double ps;
        if (fabs(wt) <= 1.0e-8) ps = 0.5 * (PZN[f] + PZN[f + 1]);
        else ps = psiface_z(f, wt);
        *vf = P.vism + (P.visp - P.vism) * ps;
        return P.denm + P.den_diff * ps;
    };
    auto zown = [&](int f, int rg, double *vf) {
        if (rg == 0) { *vf = P.vism; return P.denm; }
        if (rg == 1) { *vf = P.visp; return P.denp; }
        return zface2(f, vf);
    };
    auto cface2 = [&](int f) {
        double wt = W[f + 1];
        double ps;
        if (fabs(wt) <= 1.0e-8) ps = PCN[f + 1];
        else ps = psiface_c(f, wt);
        return P.denm + P.den_diff * ps;
    };
    auto cown = [&](int f, int rg) {
        if (rg == 0) return P.denm;
        if (rg == 1) return P.denp;
        return cface2(f);
    };

    auto region = [&](int k) { return region_of(PCN, n, k); };
    // CONTI_Z / CONTI_C in closed form: the serial loops leave on each face
    // the value written last, i.e. the next cell's constant when that cell
    // is region 0/1 and otherwise the value the face's own cell computed.
    auto conti_cell = [&](int k) {
        const int rk = region(k);
        const int rn = (k < n) ? region(k + 1) : -1;
        double vf, df;
        if (k == n) df = zown(k, rk, &vf);
        else if (rn == 0) { df = P.denm; vf = P.vism; }
        else if (rn == 1) { df = P.denp; vf = P.visp; }
        else if (k >= 2) df = zown(k, rk, &vf);
        else df = zface2(1, &vf);
        DFZ[k] = df;
        VFZ[k] = vf;
        if (k >= 2) {
            double pz;
            if (rk != 2) pz = PZN[k];
            else {
                double vv, own = zface2(k, &vv);
                double prev = (k - 1 >= 2) ? zown(k - 1, region(k - 1), &vv) : zface2(1, &vv);
                double wt = 0.5 * (W[k + 1] + W[k]);
                double wb = 0.5 * (W[k] + W[k - 1]);
                double flux = (own * wt - prev * wb) * VVDZ[k];
                double den_zn = P.denm + P.den_diff * PZN[k];
                double den_z = den_zn - dtconst * flux;
                pz = (den_z - P.denm) / P.den_diff;
            }
            if (pz > 1.0) pz = 1.0;
            else if (pz < 0.0) pz = 0.0;
            PZ[k] = pz;
        }
        // CONTI_C
        double dc;
        if (k == n) dc = cown(k, rk);
        else if (rn == 0) dc = P.denm;
        else if (rn == 1) dc = P.denp;
        else dc = cown(k, rk);
        DFC[k] = dc;
        if (k == 1) {
            double d0;
            if (rk == 0) d0 = P.denm;
            else if (rk == 1) d0 = P.denp;
            else d0 = cface2(0);
            DFC[0] = d0;
        }
        double pc;
        if (rk != 2) pc = PCN[k];
        else {
            double own = cface2(k);
            double prev = (k >= 2) ? cown(k - 1, region(k - 1)) : cface2(0);
            double flux = (own * W[k + 1] - prev * W[k]) * SSDZ[k];
            double den_cn = P.denm + P.den_diff * PCN[k];
            double den_c = den_cn - dtconst * (flux);
            pc = (den_c - P.denm) / P.den_diff;
        }
        if (pc > 1.0) pc = 1.0;
        else if (pc < 0.0) pc = 0.0;
        PC[k] = pc;
    };
    // Cells 2..n-1 away from every jump (k outside [jlo - 3, jhi + 3]): they
    // and their upper neighbours are region 0/1, so the faces take the upper
    // cell's constants and PSI_Z/PSI_C keep PSI_ZN/PSI_CN (clamped) -- four
    // cells per trip, their loads in flight together.  The cells within
    // reach of a jump and the two ends take the full rules.
    const int sa = (jhi >= 0) ? (jlo - 3 < 1 ? 1 : jlo - 3) : (1 << 30);
    const int sb = (jhi >= 0) ? (jhi + 3 > n ? n : jhi + 3) : -1;
    PLOOP {
        constexpr int U = 4;
        for (int k0 = 2 + PTID; k0 <= n - 1; k0 += U * PSTRIDE) {
            double pzn[U], pcn[U], pcn1[U];
#pragma unroll
            for (int j = 0; j < U; j++) {
                const int kc = (k0 + j * PSTRIDE < n - 1) ? k0 + j * PSTRIDE : n - 1;
                pzn[j] = PZN[kc];
                pcn[j] = PCN[kc];
                pcn1[j] = PCN[kc + 1];
            }
#pragma unroll
            for (int j = 0; j < U; j++) {
                const int k = k0 + j * PSTRIDE;
                if (k > n - 1 || (k >= sa && k <= sb)) continue;
                const bool liq = (pcn1[j] == 1.0);   // region of cell k + 1
                DFZ[k] = liq ? P.denp : P.denm;
                VFZ[k] = liq ? P.visp : P.vism;
                double pz = pzn[j];
                if (pz > 1.0) pz = 1.0;
                else if (pz < 0.0) pz = 0.0;
                PZ[k] = pz;
                DFC[k] = liq ? P.denp : P.denm;
                double pc = pcn[j];
                if (pc > 1.0) pc = 1.0;
                else if (pc < 0.0) pc = 0.0;
                PC[k] = pc;
            }
        }
    }
    PFOR(i, 0, (sb >= sa ? sb - sa + 1 : 0) + 2) {
        const int k = (i == 0) ? 1 : (i == 1) ? n : sa + i - 2;
        if (!(i == 1 && n == 1) && !(i >= 2 && (k == 1 || k == n))) conti_cell(k);
    }
    PFOR_END
    SERIAL {
        PZ[1] = PZ[2]; PZ[n3] = PZ[n];
        PC[1] = PC[2]; PC[n3] = PC[n];
        DFZ[0] = 0.0; DFZ[n3] = 0.0; DFC[n3] = 0.0;
    }
    SERIAL_END

    // energy RHS (CAL_RHSPS/CAL_RHSPS2 + PASSIVE_SCALAR scaling) and LHSPS
    const double g = P.gamma[msub - 1], ro = P.ro[msub - 1];
    const double den_diffi = (P.den_diff == 0.0) ? 1.0 : 1.0 / P.den_diff;
    const double acops = alpha * dt / (P.re * P.prm);
    auto lhsps_row = [&](int k, double &ta, double &tb, double &tc, double &tg) {
        double rhs;
        if (m != 1) {
            rhs = (g * T[k] * (W[k + 1] - W[k]) * VVDZ[k]) * dt;
            double densc, denscf;
            if (PCN[k] > 0.5) { densc = P.denscm + P.densc_diff; denscf = densc; }
            else { densc = P.denscm; denscf = densc; }
            rhs = ((densc - denscf) * T[k] + rhs) / denscf;
        } else {
            rhs = ((g * 0.0 + ro * 0.0) + 2.0 * alpha * 0.0 + 2.0 * alpha * (0.0 - 0.0)) * dt;
            double densc = P.denscm + P.densc_diff * PCN[k];
            double denscf = P.denscm + P.densc_diff * PC[k];
            rhs = ((densc - denscf) * T[k] + rhs) / denscf;
        }
        double denscf = P.denscm + P.densc_diff * PC[k];
        double coef = acops / denscf;
        double pt = (DFC[k - 1] - P.denm) * den_diffi;
        double tcb = P.tcm + P.tc_diff * pt;
        pt = (DFC[k] - P.denm) * den_diffi;
        double tct = P.tcm + P.tc_diff * pt;
        double ak = coef * tcb * AKUV[k];
        double ck = coef * tct * CKUV[k];
        double bk = 1.0 - (ak + ck);
        double pcn = PCN[k];
        if (fabs(pcn - 0.5) != 0.5) {
            if (pcn <= 0.5) ck = 0.0;
            else if (pcn > 0.5) ak = 0.0;
        }
        ta = ak; tb = bk; tc = ck; tg = rhs;
    };
    tri_solve_rows(TA, TB, TC, TG, 1, n, bd.tri_mode, sh.tri, lhsps_row, [&](int k, double x) {
        RHSPS[k] = x;
        T[k] = x + T[k];
    });
    SERIAL {
        T[0] = 0.0;
        T[n] = 1.0;
        T[n3] = T[n];
    }
    SERIAL_END
}

// ---------------------------------------------------------------------
// Band context of the wave kernels (narrow band + refined window in LDS)
static_assert(LICA_MLAYER < 256 && LICA_MAXP <= 256, "band entry tables hold layer and slot in bytes");

struct BandSh {
    int bn[LICA_MLAYER + 2];
    int bk[LICA_MLAYER + 2][LICA_MAXP];
    int off[LICA_MLAYER + 2];           // entries of layer n: [off[n], off[n+1])
    int ent_k[LICA_MAXENT];
    unsigned char ent_n[LICA_MAXENT], ent_q[LICA_MAXENT];   // layer (<= LICA_MLAYER), slot (< LICA_MAXP)
    int wlo, whi, dlo;                  // window [wlo,whi] of z-planes, work base dlo=wlo-4
    int status;
    double phw[WMAX], gl[3], gu[4];
    double phs[WMAX];
    double wf[WMAX + 8];
    double pm1[LICA_MAXENT], pm2[LICA_MAXENT], mfd[LICA_MAXENT];
    double tf0[WMAX], tf1[WMAX];
    int mask_s[WMAX];
    double red[BLK_RED_SLOTS];
    double v[24];
    int iv[24];
    int mcell[2 * LICA_MAXENT];
    // per-plane tables of the window for the register-resident stages
    int lay[WMAX + 8], slot[WMAX + 8], dmk[WMAX + 8];
};

struct Fill {           // CAL_PSIF_MTP column fill outside the band (two_phase.f:3011-3075)
    int knum, kc_low;
    double tmp1, tmp_outer;
    int khigh;          // highest sign-change plane (0 when knum == 0)
};

DEV double phi_get(const BandSh &B, int k, int n3f)
{
    if (k <= 0) return B.gl[k + 2];
    if (k >= n3f) return B.gu[k - n3f];
    if (k < B.wlo || k > B.whi) return 0.0;
    return B.phw[k - B.wlo];
}
DEV void phi_set(BandSh &B, int k, int n3f, double v)
{
    if (k <= 0) { B.gl[k + 2] = v; return; }
    if (k >= n3f) { B.gu[k - n3f] = v; return; }
    if (k < B.wlo || k > B.whi) { B.status = ST_WINDOW; return; }
    B.phw[k - B.wlo] = v;
}
