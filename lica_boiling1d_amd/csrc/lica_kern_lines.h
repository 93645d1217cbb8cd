// Line solvers of the 3-D code, which have no live caller on the 1-D column:
// TRDIAG1/2/3 (1d.f:3721-3812), the periodic TRDIAG1P/2P/3P (3815-3985)
// and the pentadiagonal PTDIAG1A (3988-4034).  The arrays are column-major
// m1 x m2 (element (i, j), 1-based, at (i-1) + (j-1)*m1), as the reference's
// A(M2,M1) / A(M1,M2) dummies; the systems run along j, one per line i.
//
// One thread solves one line.  Consecutive lines are consecutive addresses,
// so each row's loads and stores are coalesced across the wave, and every
// byte is touched once per pass over the line; a line's recurrences stay in
// registers (the reference keeps BET, PN, PDENO in per-line vectors).  Each
// line repeats the reference's operations in its order, so the results are
// the serial loops' bit for bit.
#pragma once
#include "lica_par.h"

// TRDIAG1/2/3 on line i, rows l1..l2: UU the solution, gam the elimination's
// multipliers (work, rows l1+1..l2); overwrite_r (TRDIAG3) also leaves UU in
// R(i, l1..l2-1).  R and UU may be the same array.
DEV void line_trdiag(const double *__restrict__ a, const double *__restrict__ b, const double *__restrict__ c,
                     double *r, double *uu, double *__restrict__ gam, long long m1, int i, int l1, int l2,
                     int overwrite_r)
{
    auto at = [&](int j) { return (size_t)(i - 1) + (size_t)(j - 1) * (size_t)m1; };
    double bet = 1.0 / b[at(l1)];
    double u = r[at(l1)] * bet;
    uu[at(l1)] = u;
    for (int j = l1 + 1; j <= l2; j++) {
        const double g = c[at(j - 1)] * bet;
        gam[at(j)] = g;
        bet = 1.0 / (b[at(j)] - a[at(j)] * g);
        u = (r[at(j)] - a[at(j)] * u) * bet;
        uu[at(j)] = u;
    }
    for (int j = l2 - 1; j >= l1; j--) {
        u = uu[at(j)] - gam[at(j + 1)] * u;
        uu[at(j)] = u;
        if (overwrite_r) r[at(j)] = u;
    }
}

// TRDIAG1P/2P/3P on line k, rows j1..j2 (j2 > j1): the row j1 couples to
// u(j2) through A(j1), the row j2 to u(j1) through C(j2).  F in and out;
// q, s, qe work (the forward multipliers, the fill-in column and the
// homogeneous part of the back substitution).
DEV void line_trdiag_periodic(const double *__restrict__ a, const double *__restrict__ b,
                              const double *__restrict__ c, double *__restrict__ f, double *__restrict__ q,
                              double *__restrict__ s, double *__restrict__ qe, long long m1, int k, int j1, int j2)
{
    auto at = [&](int j) { return (size_t)(k - 1) + (size_t)(j - 1) * (size_t)m1; };
    const double binv = 1.0 / b[at(j1)];
    double qv = -c[at(j1)] * binv;
    double sv = -a[at(j1)] * binv;
    q[at(j1)] = qv;
    s[at(j1)] = sv;
    const double fn = f[at(j2)];
    double fv = f[at(j1)] * binv;
    f[at(j1)] = fv;
    // forward elimination: q, s and f of row j from those of row j-1
    for (int j = j1 + 1; j <= j2; j++) {
        const double aj = a[at(j)];
        const double pn = 1.0 / (b[at(j)] + aj * qv);
        qv = -c[at(j)] * pn;
        sv = -aj * sv * pn;
        fv = (f[at(j)] - aj * fv) * pn;
        q[at(j)] = qv;
        s[at(j)] = sv;
        f[at(j)] = fv;
    }
    // backward pass from S(j2) = 1, QE(j2) = 0
    double sn = 1.0, qn = 0.0, s_top = 0.0, qe_top = 0.0;
    for (int j = j2 - 1; j >= j1; j--) {
        const double qj = q[at(j)];
        sn = s[at(j)] + qj * sn;
        qn = f[at(j)] + qj * qn;
        s[at(j)] = sn;
        qe[at(j)] = qn;
        if (j == j2 - 1) { s_top = sn; qe_top = qn; }
    }
    const double ca = c[at(j2)], aa = a[at(j2)];
    const double fl = (fn - ca * qn - aa * qe_top) / (ca * sn + aa * s_top + b[at(j2)]);
    f[at(j2)] = fl;
    for (int j = j2 - 1; j >= j1; j--) f[at(j)] = fl * s[at(j)] + qe[at(j)];
}

// PTDIAG1A on line i, rows j1..jn (jn > j1): E u(j-2) + D u(j-1) + A u(j) +
// B u(j+1) + C u(j+2) = F; O, Q, R receive the elimination's coefficients
// (outputs of the reference routine), UU the solution.
DEV void line_ptdiag(const double *__restrict__ a, const double *__restrict__ b, const double *__restrict__ c,
                     const double *__restrict__ d, const double *__restrict__ e, const double *__restrict__ f,
                     double *__restrict__ uu, double *__restrict__ o, double *__restrict__ q, double *__restrict__ r,
                     long long m1, int i, int j1, int jn)
{
    auto at = [&](int j) { return (size_t)(i - 1) + (size_t)(j - 1) * (size_t)m1; };
    const int j2 = j1 + 1, jnm = jn - 1;
    const double a1 = a[at(j1)];
    double o1 = -b[at(j1)] / a1, q1 = -c[at(j1)] / a1, r1 = f[at(j1)] / a1;
    o[at(j1)] = o1;
    q[at(j1)] = q1;
    r[at(j1)] = r1;
    const double d2 = d[at(j2)];
    const double p2 = 1.0 / (a[at(j2)] + d2 * o1);
    double o2 = -(b[at(j2)] + d2 * q1) * p2, q2 = -c[at(j2)] * p2, r2 = (f[at(j2)] - d2 * r1) * p2;
    o[at(j2)] = o2;
    q[at(j2)] = q2;
    r[at(j2)] = r2;
    // rows j >= j1 + 2 from rows j-2 (o1, q1, r1) and j-1 (o2, q2, r2);
    // t = D(j) + E(j) O(j-2), the factor the reference evaluates three times
    for (int j = j2 + 1; j <= jn; j++) {
        const double ej = e[at(j)];
        const double t = d[at(j)] + ej * o1;
        const double p = 1.0 / (a[at(j)] + ej * q1 + t * o2);
        const double on = -(b[at(j)] + t * q2) * p;
        const double qv = -c[at(j)] * p;
        const double rv = (f[at(j)] - ej * r1 - t * r2) * p;
        o[at(j)] = on;
        q[at(j)] = qv;
        r[at(j)] = rv;
        o1 = o2; q1 = q2; r1 = r2;
        o2 = on; q2 = qv; r2 = rv;
    }
    // back substitution: u(j) = O(j) u(j+1) + Q(j) u(j+2) + R(j)
    double u2 = r2;   // UU(jn) = R(jn)
    uu[at(jn)] = u2;
    double u1 = o[at(jnm)] * u2 + r[at(jnm)];
    uu[at(jnm)] = u1;
    for (int j = jnm - 1; j >= j1; j--) {
        const double u = o[at(j)] * u1 + q[at(j)] * u2 + r[at(j)];
        uu[at(j)] = u;
        u2 = u1;
        u1 = u;
    }
}
