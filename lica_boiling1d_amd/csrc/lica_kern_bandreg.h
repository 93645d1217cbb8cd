// Register-resident stages of the narrow band (one wave per case): the SSP-RK3
// stages of TRANSPORT_RK3 (two_phase.f:1133-1495) and the pseudo-time loop of
// REINITIALIZATION (two_phase.f:2033-2444).  The window of z-planes
// [dlo, dlo + 64*CPL) is spread over the lanes (plane dlo + CPL*lane + s in
// slot s); stencil neighbours come from lane shuffles, so a stage needs no
// LDS traffic or barrier.  Every plane update evaluates the same expressions,
// in the same order, as the LDS formulation (reinit_rk3 / transport_rk3), from
// the previous stage's values (Jacobi order).  The CPU logic-check build keeps
// the whole window in one lane.
#pragma once
#include "lica_kern_common.h"

#ifdef LICA_EMUL
template <int CPL>
DEV int bcell(int s) { return s; }
// out[s] = a[s + D] (zero beyond the window)
template <int CPL, int D>
DEV void xshift(const double *a, double *out)
{
    for (int s = 0; s < CPL; s++) {
        int t = s + D;
        out[s] = (t >= 0 && t < CPL) ? a[t] : 0.0;
    }
}
DEV double wave_max(double m) { return m; }
DEV double wave_min(double m) { return m; }
#else
template <int CPL>
DEV int bcell(int s) { return CPL * (int)threadIdx.x + s; }
constexpr int floor_div(int a, int b) { return (a >= 0) ? a / b : -((-a + b - 1) / b); }
template <int CPL, int D>
DEV void xshift(const double *a, double *out)
{
    static_for<CPL>([&](auto sc) {
        constexpr int s = decltype(sc)::value;
        constexpr int t = s + D;
        constexpr int lo = floor_div(t, CPL);
        constexpr int sp = t - lo * CPL;
        double v = a[sp];
        if constexpr (lo == 1) v = lane_dn1(v);
        else if constexpr (lo > 1) v = __shfl_down(v, lo, 64);
        if constexpr (lo == -1) v = lane_up1(v);
        else if constexpr (lo < -1) v = __shfl_up(v, -lo, 64);
        out[s] = v;
    });
}
DEV double wave_max(double m) { return wave_red_max(m); }
DEV double wave_min(double m) { return wave_red_min(m); }
#endif

DEV double weno_m5(double v1, double v2, double v3, double v4, double v5)
{
    double wz = d_hjweno(v2 - v1, v3 - v2, v4 - v3, v5 - v4);
    return 1.0 / 12.0 * (-v2 + 7.0 * v3 + 7.0 * v4 - v5) - wz;
}
DEV double weno_p5(double v1, double v2, double v3, double v4, double v5)
{
    double wz = d_hjweno(v1 - v2, v2 - v3, v3 - v4, v4 - v5);
    return 1.0 / 12.0 * (-v5 + 7.0 * v4 + 7.0 * v3 - v2) + wz;
}

// per-plane tables of the window (index plane - dlo), filled by band_tables
// layer: band layer of the plane (0 outside the band); slot: its list slot;
// dmask(nn1): DPHI_Z writes the plane (layers 1..nn1 and the faces around
// layer nn1, two_phase.f:1165-1232)
DEV void band_tables(BandSh &B, int nn1)
{
    PFOR(i, 0, WMAX + 8) { B.lay[i] = 0; B.slot[i] = 0; B.dmk[i] = 0; }
    PFOR_END
    PFOR(e, 0, B.off[LICA_MLAYER + 1]) {
        int i = B.ent_k[e] - B.dlo;
        B.lay[i] = B.ent_n[e];
        B.slot[i] = B.ent_q[e];
        if (B.ent_n[e] <= nn1) B.dmk[i] = 1;
    }
    PFOR_END
    PFOR(e, B.off[nn1], B.off[nn1 + 1]) {
        int i = B.ent_k[e] - B.dlo;
        B.dmk[i - 3] = 1; B.dmk[i - 2] = 1; B.dmk[i - 1] = 1; B.dmk[i + 1] = 1; B.dmk[i + 2] = 1;
    }
    PFOR_END
}

template <int CPL>
DEV void load_window(const BandSh &B, const InstParam &P, double *ph)
{
    for (int s = 0; s < CPL; s++) {
        int p = B.dlo + bcell<CPL>(s);
        ph[s] = (p >= -2 && p <= P.n3f + 3) ? phi_get(B, p, P.n3f) : 0.0;
    }
}

// write the band planes (layers 1..nmax) back to the LDS window
template <int CPL>
DEV void store_window(BandSh &B, const InstParam &P, const double *ph, int nmax)
{
    for (int s = 0; s < CPL; s++) {
        int i = bcell<CPL>(s);
        if (i < WMAX + 8) {
            int nl = B.lay[i];
            if (nl >= 1 && nl <= nmax) phi_set(B, B.dlo + i, P.n3f, ph[s]);
        }
    }
}

// DPHI_Z for the window: (phi(k+1)-phi(k))/dz on masked planes, else zero
template <int CPL>
DEV void dphi_regs(const double *ph, const double *up1, const int *dm, double ssdzf, double *dp)
{
    for (int s = 0; s < CPL; s++) dp[s] = dm[s] ? (up1[s] - ph[s]) * ssdzf : 0.0;
}

// REINITIALIZATION (two_phase.f:2033-2170) in registers; B.lay/B.dmk must
// hold band_tables(nn1)
template <int CPL>
DEV void reinit_regs(BandSh &B, const InstParam &P, int nn1, int nn2, int ilimit)
{
    double ph[CPL], po[CPL], rk[CPL];
    int nl[CPL], dm[CPL];
    load_window<CPL>(B, P, ph);
    for (int s = 0; s < CPL; s++) {
        int i = bcell<CPL>(s);
        bool in = i < WMAX + 8;
        nl[s] = in ? B.lay[i] : 0;
        dm[s] = in ? B.dmk[i] : 0;
        po[s] = ph[s];
    }
    const double sm_grid = fmin(fmin(P.sdxf, P.sdyf), P.sdzf);
    const double redtl = 0.5 * sm_grid;
    const double mg = fmin(fmin(P.sdxf, P.sdyf), P.sdzf);
    const double eps = mg * mg;
    const double wxm = weno_flat_m(), wxp = weno_flat_p();
    auto stage = [&]() {
        double up1[CPL], dn1[CPL], dp[CPL], dm3[CPL], dm2[CPL], dm1[CPL], dp1[CPL], dp2[CPL], nv[CPL];
        xshift<CPL, 1>(ph, up1);
        xshift<CPL, -1>(ph, dn1);
        dphi_regs<CPL>(ph, up1, dm, P.ssdzf, dp);
        xshift<CPL, -3>(dp, dm3);
        xshift<CPL, -2>(dp, dm2);
        xshift<CPL, -1>(dp, dm1);
        xshift<CPL, 1>(dp, dp1);
        xshift<CPL, 2>(dp, dp2);
        for (int s = 0; s < CPL; s++) {
            nv[s] = ph[s];
            if (nl[s] >= 1 && nl[s] <= nn2) {
                double v = ph[s];
                double dax = 0.5 * (0.0 - 0.0) * (1.0 / P.sdxf);
                double daz = 0.5 * (up1[s] - dn1[s]) * P.ssdzf;
                double dalphi = dax * dax + dax * dax + daz * daz;
                double asign = po[s] / sqrt(po[s] * po[s] + dalphi * eps);
                double ax, ay, az;
                if (nl[s] <= nn1) {
                    double azm = weno_m5(dm3[s], dm2[s], dm1[s], dp[s], dp1[s]);
                    double azp = weno_p5(dp2[s], dp1[s], dp[s], dm1[s], dm2[s]);
                    ax = d_godunov(asign, wxp, wxm);
                    ay = d_godunov(asign, wxp, wxm);
                    az = d_godunov(asign, azp, azm);
                } else {
                    double axm = 0.0, axp = 0.0;
                    double azm = (v - dn1[s]) * P.ssdzf;
                    double azp = (up1[s] - v) * P.ssdzf;
                    ax = d_godunov(asign, axp, axm);
                    ay = d_godunov(asign, axp, axm);
                    az = d_godunov(asign, azp, azm);
                }
                double drhs = asign * (1.0 - sqrt(ax * ax + ay * ay + az * az));
                nv[s] = v + redtl * drhs;
            }
        }
        for (int s = 0; s < CPL; s++) ph[s] = nv[s];
    };
    for (int it = 1; it <= ilimit; it++) {
        for (int s = 0; s < CPL; s++) rk[s] = ph[s];
        stage();
        stage();
        for (int s = 0; s < CPL; s++)
            if (nl[s] >= 1 && nl[s] <= nn2) ph[s] = 0.75 * rk[s] + 0.25 * ph[s];
        stage();
        for (int s = 0; s < CPL; s++)
            if (nl[s] >= 1 && nl[s] <= nn2) ph[s] = 1.0 / 3.0 * rk[s] + 2.0 / 3.0 * ph[s];
    }
    store_window<CPL>(B, P, ph, nn2);
}

// the three SSP-RK3 stages of one TRANSPORT subcycle (two_phase.f:855-880)
// in registers; B.lay/B.slot/B.dmk must hold band_tables(NN1).  ALPHI and
// MFLVS are the entry-time records, by list slot.
template <int CPL>
DEV void transport_stages_regs(BandSh &B, const InstParam &P, const InstState &S, double dtlvs)
{
    double ph[CPL], rk[CPL], u1[CPL], u2[CPL], mf[CPL];
    int nl[CPL], dm[CPL], liq[CPL];
    load_window<CPL>(B, P, ph);
    for (int s = 0; s < CPL; s++) {
        int i = bcell<CPL>(s);
        bool in = i < WMAX + 8;
        nl[s] = in ? B.lay[i] : 0;
        dm[s] = in ? B.dmk[i] : 0;
        u1[s] = in ? B.wf[i] : 0.0;
        u2[s] = (i + 1 < WMAX + 8) ? B.wf[i + 1] : 0.0;
        int q = in ? B.slot[i] : 0;
        int l = nl[s];
        mf[s] = 0.0;
        liq[s] = 0;
        if (l >= 1 && l <= LICA_NN1) {
            mf[s] = (q < S.bn[l]) ? S.mflvs[l][q] : 0.0;
            liq[s] = S.alphi[l][q] > 0.0;
        }
        rk[s] = ph[s];
    }
    const double wxm = weno_flat_m();
    auto stage = [&]() {
        double up1[CPL], dp[CPL], dm3[CPL], dm2[CPL], dm1[CPL], dp1[CPL], dp2[CPL], nv[CPL];
        xshift<CPL, 1>(ph, up1);
        dphi_regs<CPL>(ph, up1, dm, P.ssdzf, dp);
        xshift<CPL, -3>(dp, dm3);
        xshift<CPL, -2>(dp, dm2);
        xshift<CPL, -1>(dp, dm1);
        xshift<CPL, 1>(dp, dp1);
        xshift<CPL, 2>(dp, dp2);
        for (int s = 0; s < CPL; s++) {
            nv[s] = ph[s];
            if (nl[s] >= 1 && nl[s] <= LICA_NN1) {
                double ax = 0.0 * wxm, ay = 0.0 * wxm, az;
                double uu1 = u1[s], uu2 = u2[s], uu3 = 0.5 * (uu1 + uu2);
                if (uu1 >= 0.0 && uu2 >= 0.0 && uu3 >= 0.0) {
                    az = uu3 * weno_m5(dm3[s], dm2[s], dm1[s], dp[s], dp1[s]);
                } else if (uu1 <= 0.0 && uu2 <= 0.0 && uu3 <= 0.0) {
                    az = uu3 * weno_p5(dp2[s], dp1[s], dp[s], dm1[s], dm2[s]);
                } else {
                    double mm = weno_m5(dm3[s], dm2[s], dm1[s], dp[s], dp1[s]);
                    double pp = weno_p5(dp2[s], dp1[s], dp[s], dm1[s], dm2[s]);
                    double umax = fmax(fmax(fabs(uu1), fabs(uu2)), fabs(uu3));
                    az = 0.5 * (uu3 * (pp + mm) - umax * (pp - mm));
                }
                double v = ph[s] - dtlvs * (ax + ay + az);
                if (liq[s]) v = v - dtlvs * mf[s] / P.denr;
                else v = v - dtlvs * mf[s];
                nv[s] = v;
            }
        }
        for (int s = 0; s < CPL; s++) ph[s] = nv[s];
    };
    stage();
    stage();
    for (int s = 0; s < CPL; s++)
        if (nl[s] >= 1 && nl[s] <= LICA_NN1) ph[s] = 0.75 * rk[s] + 0.25 * ph[s];
    stage();
    for (int s = 0; s < CPL; s++)
        if (nl[s] >= 1 && nl[s] <= LICA_NN1) ph[s] = 1.0 / 3.0 * rk[s] + 2.0 / 3.0 * ph[s];
    store_window<CPL>(B, P, ph, LICA_NN1);
}

// dispatch on the window width (planes dlo .. whi+4)
DEV void reinit_band(BandSh &B, const InstParam &P, InstState &S, int nn1, int nn2, int ilimit)
{
    band_tables(B, nn1);
    const int nw = B.whi + 4 - B.dlo + 1;
#ifdef LICA_EMUL
    (void)nw;
    reinit_regs<WMAX + 8>(B, P, nn1, nn2, ilimit);
#else
    // the first wave holds the window (a second wave of the workgroup waits)
    if (threadIdx.x < 64) {
        if (nw <= 64) reinit_regs<1>(B, P, nn1, nn2, ilimit);
        else if (nw <= 128) reinit_regs<2>(B, P, nn1, nn2, ilimit);
        else reinit_regs<4>(B, P, nn1, nn2, ilimit);
    }
#endif
    BARRIER();
    SERIAL { S.n_reinit += 1.0; }
    SERIAL_END
}

DEV void transport_stages(BandSh &B, const InstParam &P, const InstState &S, double dtlvs)
{
    band_tables(B, LICA_NN1);
    const int nw = B.whi + 4 - B.dlo + 1;
#ifdef LICA_EMUL
    (void)nw;
    transport_stages_regs<WMAX + 8>(B, P, S, dtlvs);
#else
    if (nw <= 64) transport_stages_regs<1>(B, P, S, dtlvs);
    else if (nw <= 128) transport_stages_regs<2>(B, P, S, dtlvs);
    else transport_stages_regs<4>(B, P, S, dtlvs);
#endif
    BARRIER();
}
