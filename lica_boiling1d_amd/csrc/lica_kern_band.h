// Narrow-band level-set work on one wave per case: TRANSPORT and
// GRID_COUPLING(0) (two_phase.f:701-1130, 4863-5118) and TEMP_GRAD
// (chphase.f:55-368).  The refined column outside the band is zero
// (ALPHI_EXT is rebuilt from the band at every call), so only a window of
// z-planes around the band is held, in LDS.
#pragma once
#include "lica_kern_common.h"

// entry list (layers 1..11 in list order) from B.bn/B.bk; window bounds
DEV void band_index(BandSh &B, const InstParam &P, int set_window)
{
    int e = 0, kmin = 1 << 30, kmax = -1;
    for (int nl = 1; nl <= LICA_MLAYER; nl++) {
        B.off[nl] = e;
        for (int q = 0; q < B.bn[nl]; q++) {
            int k = B.bk[nl][q];
            B.ent_k[e] = k; B.ent_n[e] = nl; B.ent_q[e] = q;
            e++;
            if (k < kmin) kmin = k;
            if (k > kmax) kmax = k;
        }
    }
    B.off[LICA_MLAYER + 1] = e;
    if (set_window) {
        if (kmax < 0) { kmin = 1; kmax = 1; }
        B.wlo = kmin - 12 < 1 ? 1 : kmin - 12;
        B.whi = kmax + 12 > P.n3fm ? P.n3fm : kmax + 12;
        B.dlo = B.wlo - 4;
        if (B.whi - B.wlo + 1 > WMAX) B.status = ST_WINDOW;
    } else {
        if (kmax >= 0 && (kmin - 3 < B.wlo - 3 || kmax + 3 > B.whi + 3 || kmin < B.wlo || kmax > B.whi)) B.status = ST_WINDOW;
    }
}

DEV double wf_get(const BandSh &B, int k) { return B.wf[k - B.dlo]; }

// ALPHI_BC z part (two_phase.f:565-600) and the z copy (623-634); copies
// outside ALPHI_EXT(-2:N3F+3) are dropped.  Call inside SERIAL.
DEV void alphi_bc(BandSh &B, const InstParam &P)
{
    const int n3f = P.n3f, n3fm = P.n3fm;
    const double sd = P.sdzf;
    double a1 = phi_get(B, 1, n3f);
    if (fabs(a1) < 1.0e-8) { B.gl[2] = 0.0; B.gl[1] = 0.0; B.gl[0] = 0.0; }
    else if (a1 > 0.0) { B.gl[2] = a1 + sd; B.gl[1] = B.gl[2] + sd; B.gl[0] = B.gl[1] + sd; }
    else if (a1 < 0.0) { B.gl[2] = a1 - sd; B.gl[1] = B.gl[2] - sd; B.gl[0] = B.gl[1] - sd; }
    double an = phi_get(B, n3fm, n3f);
    if (fabs(an) < 1.0e-8) { for (int k = 0; k <= 3; k++) B.gu[k] = 0.0; }
    else if (an > 0.0) { B.gu[0] = an + sd; for (int k = 1; k <= 3; k++) B.gu[k] = B.gu[k - 1] + sd; }
    else if (an < 0.0) { B.gu[0] = an - sd; for (int k = 1; k <= 3; k++) B.gu[k] = B.gu[k - 1] - sd; }
    for (int k = 0; k <= P.r; k++) {
        if (-k >= -2) phi_set(B, -k, n3f, phi_get(B, n3fm - k, n3f));
        if (n3f + k <= n3f + 3) phi_set(B, n3f + k, n3f, phi_get(B, 1 + k, n3f));
    }
}

DEV double weno_flat_m() { return 1.0 / 12.0 * (-0.0 + 7.0 * 0.0 + 7.0 * 0.0 - 0.0) - d_hjweno(0.0, 0.0, 0.0, 0.0); }
DEV double weno_flat_p() { return 1.0 / 12.0 * (-0.0 + 7.0 * 0.0 + 7.0 * 0.0 - 0.0) + d_hjweno(0.0, 0.0, 0.0, 0.0); }

// CAL_PSIF_MTP band formula (two_phase.f:2867-2936) for layers 1..nn1 into
// out[entry], plus the window view pmw/tmk used by the implicit column.
DEV void psif_mtp_band(BandSh &B, const InstParam &P, int nn1, double *out)
{
    const int n3f = P.n3f;
    PFOR(e, B.off[1], B.off[nn1 + 1]) {
        int k = B.ent_k[e];
        double d3 = 0.5 * fabs(phi_get(B, k + 1, n3f) - phi_get(B, k - 1, n3f));
        double v = d_psif_cell(phi_get(B, k, n3f), d3);
        if (v == -1.0) B.status = ST_PSIF_BUG;
        out[e] = snap5(v);
    }
    PFOR_END
}

// column fill parameters (call inside SERIAL): sign changes are band planes
DEV Fill fill_params(const BandSh &B, const InstParam &P)
{
    Fill f;
    f.knum = 0; f.kc_low = 0; f.tmp1 = 1.0; f.tmp_outer = 1.0;
    int khigh = 0;
    int kt = B.whi < P.n3fm ? B.whi : P.n3fm;
    for (int k = kt; k >= B.wlo && k >= 1; k--) {
        if (phi_get(B, k, P.n3f) * phi_get(B, k + 1, P.n3f) < 0.0) {
            if (f.knum == 0) khigh = k;
            f.knum++;
            f.kc_low = k;
        }
    }
    if (f.knum > 0) {
        f.tmp1 = (phi_get(B, khigh, P.n3f) < 0.0) ? 0.0 : 1.0;
        f.tmp_outer = (phi_get(B, f.kc_low + 1, P.n3f) < 0.0) ? 0.0 : 1.0;
    }
    f.khigh = khigh;
    return f;
}

// fill_params with the scan as workgroup reductions (call from every
// thread; every thread gets the result): the count of sign-change planes,
// the highest (the downward scan's first) and the lowest (its last)
DEV Fill fill_params_all(BandSh &B, const InstParam &P)
{
    const int kt = B.whi < P.n3fm ? B.whi : P.n3fm, kb = B.wlo < 1 ? 1 : B.wlo;
    double hi, lo, cnt[1];
    red_max2(kb, kt + 1, [&](int k, double &u, double &v) {
        const bool c = phi_get(B, k, P.n3f) * phi_get(B, k + 1, P.n3f) < 0.0;
        u = c ? (double)k : 0.0;
        v = c ? (double)(kt + 1 - k) : 0.0;
    }, B.red, &hi, &lo);
    red_sumN<1>(kb, kt + 1, [&](int k, double *v) {
        v[0] = (phi_get(B, k, P.n3f) * phi_get(B, k + 1, P.n3f) < 0.0) ? 1.0 : 0.0;
    }, B.red, cnt);
    Fill f;
    f.knum = (int)cnt[0];
    f.kc_low = 0; f.tmp1 = 1.0; f.tmp_outer = 1.0; f.khigh = 0;
    if (f.knum > 0) {
        f.khigh = (int)hi;
        f.kc_low = kt + 1 - (int)lo;
        f.tmp1 = (phi_get(B, f.khigh, P.n3f) < 0.0) ? 0.0 : 1.0;
        f.tmp_outer = (phi_get(B, f.kc_low + 1, P.n3f) < 0.0) ? 0.0 : 1.0;
    }
    return f;
}

DEV double fill_val(const Fill &f, int k) { return (f.knum == 0) ? 1.0 : ((k <= f.kc_low) ? f.tmp1 : f.tmp_outer); }

// window view of a T-band psif_mtp (values by plane) for layers 1..nn1
DEV void psif_view(BandSh &B, int nn1, const double *vals, double *pmw, int *tmk)
{
    PFOR(i, 0, WMAX) { tmk[i] = 0; }
    PFOR_END
    PFOR(e, B.off[1], B.off[nn1 + 1]) { int k = B.ent_k[e]; pmw[k - B.wlo] = vals[e]; tmk[k - B.wlo] = 1; }
    PFOR_END
}

DEV double psif_implicit(const BandSh &B, const Fill &f, const double *pmw, const int *tmk, int k)
{
    if (k >= B.wlo && k <= B.whi && tmk[k - B.wlo]) return pmw[k - B.wlo];
    return fill_val(f, k);
}

// serial sum over k=1..N3FM of (1-psif_mtp(k))*DVOL in ascending order
// (LVSINIT 427-433, GLOBAL_MASS_CORRECTION 6252-6259, VOL_TOT_MFDOT
// 1444-1451); runs of empty cells starting from zero use the stored
// partial sums vpref(m) of m DVOL terms.  Call inside SERIAL.
DEV double vol_implicit(const BandSh &B, const InstParam &P, const double *vpref, const Fill &f,
                        const double *pmw, const int *tmk)
{
    const double dvol = P.sdzf;
    double s = 0.0;
    auto run = [&](int k0, int k1, double psi) {
        if (k1 < k0) return;
        if (psi == 1.0) return;
        if (psi == 0.0 && s == 0.0) { s = vpref[k1 - k0 + 1]; return; }
        for (int k = k0; k <= k1; k++) s = s + (1.0 - psi) * dvol;
    };
    auto runs = [&](int k0, int k1) {
        if (k1 < k0) return;
        if (f.knum == 0) { run(k0, k1, 1.0); return; }
        int a1 = k1 < f.kc_low ? k1 : f.kc_low;
        run(k0, a1, f.tmp1);
        int b0 = k0 > f.kc_low + 1 ? k0 : f.kc_low + 1;
        run(b0, k1, f.tmp_outer);
    };
    int wl = B.wlo < 1 ? 1 : B.wlo, wh = B.whi > P.n3fm ? P.n3fm : B.whi;
    runs(1, wl - 1);
    for (int k = wl; k <= wh; k++) s = s + (1.0 - psif_implicit(B, f, pmw, tmk, k)) * dvol;
    runs(wh + 1, P.n3fm);
    return s;
}

// BAND_GENERATION (two_phase.f:1622-2030) on z-planes.  Call inside SERIAL.
DEV void band_generation(BandSh &B, const InstParam &P, InstState &S, int (*nbk)[LICA_MAXP], int *nbn)
{
    const int n3f = P.n3f;
    for (int i = 0; i < WMAX; i++) B.mask_s[i] = 0;
    for (int nl = 0; nl <= LICA_MLAYER; nl++) nbn[nl] = 0;
    for (int nl = 1; nl <= LICA_NN2; nl++)
        for (int q = 0; q < B.bn[nl]; q++) {
            int k = B.bk[nl][q];
            double a = phi_get(B, k, n3f);
            if (a * phi_get(B, k + 1, n3f) < 0.0 || a * phi_get(B, k - 1, n3f) < 0.0) {
                if (nbn[1] >= LICA_MAXP) { B.status = ST_BAND_CAPACITY; return; }
                nbk[1][nbn[1]++] = k;
                B.mask_s[k - B.wlo] = 1;
            } else {
                B.mask_s[k - B.wlo] = 2;
            }
        }
    for (int nl = 2; nl <= LICA_MLAYER; nl++)
        for (int q = 0; q < nbn[nl - 1]; q++) {
            int k = nbk[nl - 1][q];
            for (int k2 = k - 1; k2 <= k + 1; k2++) {
                if (k2 == 1 || k2 == n3f) continue;
                if (k2 < B.wlo || k2 > B.whi) { B.status = ST_WINDOW; return; }
                int ms = B.mask_s[k2 - B.wlo];
                if (ms != 1) {
                    if (ms == 0 || nl >= LICA_NN2) {
                        double dz = P.sdzf;
                        double dist = sqrt(0.0 * 0.0 + 0.0 * 0.0 + dz * dz);
                        double pk = phi_get(B, k, n3f);
                        if (pk >= 0.0) phi_set(B, k2, n3f, pk + dist);
                        else phi_set(B, k2, n3f, pk - dist);
                    }
                    if (nbn[nl] >= LICA_MAXP) { B.status = ST_BAND_CAPACITY; return; }
                    nbk[nl][nbn[nl]++] = k2;
                    B.mask_s[k2 - B.wlo] = 1;
                }
            }
        }
    for (int nl = 1; nl <= LICA_MLAYER; nl++) {
        if ((long long)nbn[nl] * P.r * P.r > P.mf_band) { B.status = ST_BAND_OVERFLOW; return; }
    }
    for (int nl = 1; nl <= LICA_MLAYER; nl++) {
        B.bn[nl] = nbn[nl];
        for (int q = 0; q < nbn[nl]; q++) B.bk[nl][q] = nbk[nl][q];
    }
    S.n_bandgen += 1.0;
    band_index(B, P, 0);
}

// VEL_INTP's refined velocity (two_phase.f:5618-5663) from UT=(U+UO)/2
// with the x/y Lagrange weights summed exactly (VEL_R3 5990-6043)
DEV void vel_intp(BandSh &B, const CaseP &cp)
{
    const InstParam &P = *cp.P;
    const int n = P.n, n3 = P.n3;
    const double *WT = cp.A(NS_WT), *ZF = cp.R(RF_ZF), *ZP = cp.A(NS_ZP), *VDZ = cp.A(NS_VDZ);
    const int *KCV = cp.RI(RFI_KCOU_VEL);
    PFOR(e, B.off[1], B.off[LICA_NN2 + 1]) {
        int k = B.ent_k[e], nl = B.ent_n[e];
        int kk = KCV[k];
        int km1 = kk - 1, kp1 = kk + 1, kp2 = kp1 + 1;
        double zz = ZF[k], z2 = ZP[kk], z1, z3, z4;
        z1 = z2 - VDZ[kk];
        if (kk == n - 1) { z3 = z2 + VDZ[kp1]; z4 = z3 + 2.0 * VDZ[n3]; }
        else if (kk == n) { z3 = z2 + 2.0 * VDZ[n3]; z4 = z3 + VDZ[n]; }
        else { z3 = z2 + VDZ[kp1]; z4 = z3 + VDZ[kp2]; }
        double cz1 = ((zz - z2) * (zz - z3) * (zz - z4)) / ((z1 - z2) * (z1 - z3) * (z1 - z4));
        double cz2 = ((zz - z1) * (zz - z3) * (zz - z4)) / ((z2 - z1) * (z2 - z3) * (z2 - z4));
        double cz3 = ((zz - z1) * (zz - z2) * (zz - z4)) / ((z3 - z1) * (z3 - z2) * (z3 - z4));
        double cz4 = ((zz - z1) * (zz - z2) * (zz - z3)) / ((z4 - z1) * (z4 - z2) * (z4 - z3));
        double az1 = 0.5 * (WT[km1] + WT[km1 + 1]);
        double az2 = 0.5 * (WT[kk] + WT[kk + 1]);
        double az3 = 0.5 * (WT[kp1] + WT[kp1 + 1]);
        double az4 = 0.5 * (WT[kp2] + WT[kp2 + 1]);
        double wi = cz1 * az1 + cz2 * az2 + cz3 * az3 + cz4 * az4;
        B.wf[k - B.dlo] = (nl < LICA_NN2 - 1) ? wi : wi * 0.5;
    }
    PFOR_END
    PFOR(e, B.off[LICA_MLAYER], B.off[LICA_MLAYER + 1]) { B.wf[B.ent_k[e] - B.dlo] = 0.0; }
    PFOR_END
}

#include "lica_kern_bandreg.h"

// GLOBAL_MASS_CORRECTION (two_phase.f:6196-6318)
DEV void global_mass_correction(BandSh &B, const CaseP &cp, InstState &S, int nn1)
{
    const InstParam &P = *cp.P;
    const double *VPREF = cp.R(RF_VPREFIX);
    const int n3f = P.n3f;
    SERIAL {
        S.vol_tot_ori = S.vol_tot_iphs;
        B.v[10] = 0.0; B.v[11] = 0.0;   // vol_tot, vol_in
    }
    SERIAL_END
    for (int l = 1; l <= 15; l++) {
        psif_mtp_band(B, P, nn1, B.pm2);
        if (l == 1) {
            psif_view(B, nn1, B.pm2, B.tf0, B.mask_s);   // tf0/mask_s reused as scratch
        }
        SERIAL {
            double recor = fmin(fmin(P.sdxf, P.sdyf), P.sdzf);
            double dvol = P.sdzf;
            recor = recor * 1.5;
            double vol_band = 0.0, vol_tot;
            if (l == 1) {
                Fill f = fill_params(B, P);
                vol_tot = vol_implicit(B, P, VPREF, f, B.tf0, B.mask_s);
                for (int e = B.off[1]; e < B.off[nn1 + 1]; e++) vol_band = vol_band + (1.0 - B.pm2[e]) * dvol;
                B.v[11] = vol_tot - vol_band;
            } else {
                for (int e = B.off[1]; e < B.off[nn1 + 1]; e++) vol_band = vol_band + (1.0 - B.pm2[e]) * dvol;
                vol_tot = B.v[11] + vol_band;
            }
            B.v[10] = vol_tot;
            S.n_masscor += 1.0;
            if (fabs(vol_tot - S.vol_tot_ori) >= 1.0e-5) {
                B.v[12] = recor * (-(S.vol_tot_ori - vol_tot) / S.vol_tot_ori);
                B.iv[10] = 0;
            } else {
                B.iv[10] = 1;
            }
        }
        SERIAL_END
        if (B.iv[10]) break;
        const double cor = B.v[12];
        PFOR(e, B.off[1], B.off[nn1 + 1]) {
            int k = B.ent_k[e];
            phi_set(B, k, n3f, phi_get(B, k, n3f) + cor);
        }
        PFOR_END
    }
    SERIAL {
        if (fabs(B.v[10] - S.vol_tot_ori) >= 1.0e-5) S.n_mass_warn += 1.0;
    }
    SERIAL_END
}

// ---------------------------------------------------------------------
struct TransportSh {
    BandSh B;
    int nbk[LICA_MLAYER + 1][LICA_MAXP];
    int nbn[LICA_MLAYER + 1];
};

// the band of a case from its state into LDS: counts, then the cells (in
// parallel), then the entry tables and the window (call from every thread)
DEV void band_load(BandSh &B, const InstParam &P, const InstState &S)
{
    SERIAL {
        B.status = 0;
        for (int nl = 0; nl <= LICA_MLAYER + 1; nl++) B.bn[nl] = (nl >= 1 && nl <= LICA_MLAYER) ? S.bn[nl] : 0;
        for (int i = 0; i < 3; i++) B.gl[i] = 0.0;
        for (int i = 0; i < 4; i++) B.gu[i] = 0.0;
    }
    SERIAL_END
    PFOR(j, 0, LICA_MLAYER * LICA_MAXP) {
        const int nl = 1 + j / LICA_MAXP, q = j % LICA_MAXP;
        if (q < B.bn[nl]) B.bk[nl][q] = S.bk[nl][q];
    }
    PFOR_END
    SERIAL { band_index(B, P, 1); }
    SERIAL_END
}

DEV void body_transport(const BatchDev &bd, int inst, TransportSh &T)
{
    InstState &S = bd.st[inst];
    if (S.status) return;
    BandSh &B = T.B;
    CaseP cp = case_ptrs(bd, inst);
    const InstParam &P = *cp.P;
    const int n = P.n, n3 = P.n3, n3f = P.n3f, n3fm = P.n3fm;
    band_load(B, P, S);
    if (B.status) { SERIAL { S.status = B.status; } SERIAL_END return; }
    PFOR(i, 0, WMAX) { B.phw[i] = 0.0; }
    PFOR_END
    PFOR(i, 0, WMAX + 8) { B.wf[i] = 0.0; }
    PFOR_END
    PFOR(e, 0, B.off[LICA_MLAYER + 1]) { B.phw[B.ent_k[e] - B.wlo] = S.alphi[B.ent_n[e]][B.ent_q[e]]; }
    PFOR_END
    SERIAL { alphi_bc(B, P); }
    SERIAL_END
    vel_intp(B, cp);

    const double dtconst = S.dtconst;
    double time_tran = 0.0;
    int done = 0;
    for (int isub = 1; isub <= 10; isub++) {
        double cflm = red_max(B.off[1], B.off[LICA_NN1 + 1], 0.0, [&](int e) {
            int k = B.ent_k[e];
            double cflx = fabs(0.5 * (0.0 + 0.0)) * (1.0 / P.sdxf);
            double cflz = fabs(0.5 * (wf_get(B, k - 1) + wf_get(B, k))) * P.ssdzf;
            return cflx + cflx + cflz;
        }, B.red);
        SERIAL {
            double dtlvs;
            if (cflm == 0.0) {
                double uvw = 0.0;
                for (int e = B.off[1]; e < B.off[LICA_NN1 + 1]; e++) uvw = uvw + (0.0 + 0.0 + wf_get(B, B.ent_k[e]));
                if (uvw == 0.0) dtlvs = dtconst / 10.0;
                else dtlvs = 1.0 / cflm;
            } else {
                dtlvs = 1.0 / cflm;
            }
            if (time_tran + dtlvs >= dtconst) dtlvs = dtconst - time_tran;
            B.v[0] = dtlvs;
            B.v[1] = time_tran + dtlvs;
        }
        SERIAL_END
        const double dtlvs = B.v[0];
        time_tran = B.v[1];
        transport_stages(B, P, S, dtlvs);
        // REIN_TRIGGER (two_phase.f:646-697)
        auto dsum = [&](int e) {
            int k = B.ent_k[e];
            double dx = 0.5 * (0.0 - 0.0) * (1.0 / P.sdxf);
            double dz = 0.5 * (phi_get(B, k + 1, n3f) - phi_get(B, k - 1, n3f)) * P.ssdzf;
            return dx * dx + dx * dx + dz * dz;
        };
        double mx = red_max(B.off[1], B.off[LICA_NN1 + 1], 0.0, dsum, B.red);
        double mn = red_min(B.off[1], B.off[LICA_NN1 + 1], 1.0, dsum, B.red);
        double dmax = sqrt(mx), dmin = sqrt(mn);
        int trig = (dmax > 2.0 || dmin < 1.0e-4) ? 1 : 0;
        int imass = trig;
        reinit_band(B, P, S, LICA_NN1, LICA_NN2, LICA_NN2);
        if (S.m % 50 == 0) imass = 1;
        if (trig || imass) {
            if (imass) global_mass_correction(B, cp, S, LICA_NN1);
            SERIAL { alphi_bc(B, P); band_generation(B, P, S, T.nbk, T.nbn); }
            SERIAL_END
            if (B.status) { SERIAL { S.status = B.status; } SERIAL_END return; }
            SERIAL { alphi_bc(B, P); }
            SERIAL_END
            reinit_band(B, P, S, LICA_NN1, LICA_NN2, LICA_NN2);
        }
        SERIAL { alphi_bc(B, P); }
        SERIAL_END
        if (time_tran == dtconst) { done = 1; break; }
        if (trig) vel_intp(B, cp);
    }
    // CAL_PSIF(0) for LLVS=1 (two_phase.f:2561-2617, snaps 2715-2719, BCs)
    double *PSIF = cp.R(RF_PSIF);
    psif_mtp_band(B, P, LICA_NN1, B.pm1);
    PFOR(e, B.off[1], B.off[LICA_NN1 + 1]) { PSIF[B.ent_k[e]] = snap5(B.pm1[e]); }
    PFOR_END
    SERIAL {
        if (!done) S.n_subcycle_warn += 1.0;
        if (B.status) S.status = B.status;
        for (int k = 0; k <= P.r; k++) { PSIF[-k] = PSIF[1]; PSIF[n3f + k] = PSIF[n3fm]; }
        for (int nl = 1; nl <= LICA_MLAYER; nl++) {
            S.bn[nl] = B.bn[nl];
            for (int q = 0; q < B.bn[nl]; q++) S.bk[nl][q] = B.bk[nl][q];
        }
    }
    SERIAL_END
    PFOR(e, 0, B.off[LICA_MLAYER + 1]) { S.alphi[B.ent_n[e]][B.ent_q[e]] = phi_get(B, B.ent_k[e], n3f); }
    PFOR_END
    // GRID_COUPLING(0) (two_phase.f:4863-5118): NS cells under the T-band
    double *PZN = cp.A(NS_PSI_ZN), *PCN = cp.A(NS_PSI_CN);
    const double *ZP = cp.A(NS_ZP), *Z = cp.A(NS_Z), *ZF = cp.R(RF_ZF);
    const int *KCOU1 = cp.AI(NSI_KCOU1), *KCOU2 = cp.AI(NSI_KCOU2);
    const int *KCOUMP1 = cp.AI(NSI_KCOUMP1), *KCOUMP2 = cp.AI(NSI_KCOUMP2);
    const int *KCMV = cp.RI(RFI_KCOUMP_VEL);
    // each NS cell under the T-band once: the entry that meets it first
    // (the coupled values depend only on the cell)
    const int e0 = B.off[1], e1 = B.off[LICA_NN1 + 1];
    int *xk = B.mcell;
    PFOR(e, e0, e1) { xk[e - e0] = KCMV[B.ent_k[e]]; }
    PFOR_END
    PFOR(e, e0, e1) {
        const int k = xk[e - e0];
        bool first = true;
        for (int j = 0; j < e - e0 && first; j++) first = (xk[j] != k);
        if (first) {
            auto fpsif = [&](int kk) { return PSIF[kk]; };
            if (k >= 2) PZN[k] = snap3(box_avg(ZF, P.ssdzf, fpsif, ZP[k - 1], ZP[k], KCOUMP2[k - 1], KCOUMP1[k]));
            PCN[k] = snap3(box_avg(ZF, P.ssdzf, fpsif, Z[k], Z[k + 1], KCOU2[k], KCOU1[k + 1]));
        }
    }
    PFOR_END
    SERIAL {
        PCN[0] = PCN[1];
        PCN[n3] = PCN[n];
        if (bd.msub == 3) {
            S.time = S.time + S.dt;
            if (S.status == 0 && S.nf_pending) S.status = ST_NONFINITE;
        }
    }
    SERIAL_END
}

// The level set of a case restarted from a field file: LVSINIT with the
// band read (ITRACKING=1, two_phase.f:146-189) -- ALPHI_EXT from the band,
// ALPHI_BC, CAL_PSIF(1,N_MAX) (the band formula on every layer and the
// column fill elsewhere, two_phase.f:2835-3164) -- then GRID_COUPLING(1)
// (4863-5118) and the CFL input of the first DETERMINE_DT.  The host has
// put W, P, T, the band (S.bn, S.bk, S.alphi), the time, DT and
// VOL_TOT_ORI in place, as PREFLD reads them (1d.f:972-1097).
DEV void body_init_fld(const BatchDev &bd, int inst, TransportSh &T)
{
    InstState &S = bd.st[inst];
    BandSh &B = T.B;
    CaseP cp = case_ptrs(bd, inst);
    const InstParam &P = *cp.P;
    const int n = P.n, n3 = P.n3, n3f = P.n3f, n3fm = P.n3fm;
    band_load(B, P, S);
    if (B.status) { SERIAL { S.status = B.status; } SERIAL_END return; }
    PFOR(i, 0, WMAX) { B.phw[i] = 0.0; }
    PFOR_END
    PFOR(e, 0, B.off[LICA_MLAYER + 1]) { B.phw[B.ent_k[e] - B.wlo] = S.alphi[B.ent_n[e]][B.ent_q[e]]; }
    PFOR_END
    SERIAL { alphi_bc(B, P); }
    SERIAL_END
    double *PSIF = cp.R(RF_PSIF);
    psif_mtp_band(B, P, LICA_MLAYER, B.pm1);
    SERIAL {
        const Fill f = fill_params(B, P);
        B.iv[0] = f.knum; B.iv[1] = f.kc_low; B.iv[2] = f.khigh;
        B.v[0] = f.tmp1; B.v[1] = f.tmp_outer;
    }
    SERIAL_END
    Fill f;
    f.knum = B.iv[0]; f.kc_low = B.iv[1]; f.khigh = B.iv[2]; f.tmp1 = B.v[0]; f.tmp_outer = B.v[1];
    PFOR(k, 1, n3fm + 1) { PSIF[k] = fill_val(f, k); }
    PFOR_END
    PFOR(e, 0, B.off[LICA_MLAYER + 1]) { PSIF[B.ent_k[e]] = B.pm1[e]; }
    PFOR_END
    SERIAL {
        if (B.status) S.status = B.status;
        for (int k = 0; k <= P.r; k++) { PSIF[-k] = PSIF[1]; PSIF[n3f + k] = PSIF[n3fm]; }
    }
    SERIAL_END
    if (S.status) return;
    // GRID_COUPLING(1): every NS cell
    double *PZN = cp.A(NS_PSI_ZN), *PCN = cp.A(NS_PSI_CN);
    const double *W = cp.A(NS_W), *SSDZ = cp.A(NS_SSDZ);
    const double *ZP = cp.A(NS_ZP), *Z = cp.A(NS_Z), *ZF = cp.R(RF_ZF);
    const int *KCOU1 = cp.AI(NSI_KCOU1), *KCOU2 = cp.AI(NSI_KCOU2);
    const int *KCOUMP1 = cp.AI(NSI_KCOUMP1), *KCOUMP2 = cp.AI(NSI_KCOUMP2);
    auto fpsif = [&](int kk) { return PSIF[kk]; };
    PFOR(k, 1, n + 1) {
        if (k >= 2) PZN[k] = snap3(box_avg(ZF, P.ssdzf, fpsif, ZP[k - 1], ZP[k], KCOUMP2[k - 1], KCOUMP1[k]));
        else PZN[k] = snap3(PZN[k]);
        PCN[k] = snap3(box_avg(ZF, P.ssdzf, fpsif, Z[k], Z[k + 1], KCOU2[k], KCOU1[k + 1]));
    }
    PFOR_END
    SERIAL {
        PCN[0] = PCN[1];
        PCN[n3] = PCN[n];
    }
    SERIAL_END
    double c3max = red_max(1, n + 1, 0.0, [&](int k) { return fabs(W[k] + W[k + 1]) * SSDZ[k]; }, B.red);
    SERIAL { S.c3max = c3max; }
    SERIAL_END
}
