// TEMP_GRAD (chphase.f:55-368) on two waves per case: phase marking of the
// NS cells under the band, the pseudo-time temperature extrapolation
// TEMP_EXTP (467-803) on the NS window under the band (one wave per phase),
// TF_GRAD (959-1159) and VOL_TOT_MFDOT (1328-1493).
#pragma once
#include "lica_kern_band.h"

// the N-tube entries (layers 1..NN2) precede the outermost layer's
static_assert(LICA_NN2 + 1 <= LICA_MLAYER, "layer order of the band entries");

// LDS of a workgroup (the band, its windows, the TEMP_EXTP window); the
// window views of psif_mtp that VOL_TOT_MFDOT and CAL_PSI_TMP need reuse
// B.tf0/B.mask_s and B.tf1/B.slot, free by then (TF_GRAD and the
// reinitialisation are done), which keeps six workgroups on a CU
// TEMP_EXTP's window of NS cells [nlo, nhi] under the band
// The NS window of TEMP_EXTP, cells nlo..nhi (psitmp: CAL_PSI_TMP of cells
// nlo-1 .. nhi+1).  Only t0 has LDS of its own; the other arrays use band
// arrays that are dead while they are live, which keeps the workgroup
// within 20 KB of LDS (eight workgroups per CU): nphi and psih live from the
// window's set-up to the end of the extrapolation, in B.phs and B.mfd
// (both written afresh by VOL_TOT_MFDOT before they are read); t1 in B.wf
// (TEMP_GRAD has no use for it); psitmp, after TEMPASSIGN has read t0, in t0.
struct ExtpWin {
    int nlo, nhi;
    double *t0, *t1, *nphi, *psih, *psitmp;
};

struct TgradSh {
    BandSh B;
    int nlo, nhi;
    double t0[NSWMAX];
    double xit[2], xhv[2];   // TEMP_EXTP iterations / halvings of phases 0, 1
    int xst[2];
};
static_assert(NSWMAX <= WMAX && NSWMAX <= LICA_MAXENT && NSWMAX <= WMAX + 8,
              "the window arrays must fit the band arrays they use");

// One point of TEMP_EXTP_RK3 (chphase.f:806-956): the upwind pseudo-time
// update of cell value a with neighbours al/ar; the normal is fixed to +z
// (chphase.f:897-899) and x/y differences vanish on the flat column.
DEV double extp_point(double a, double al, double ar, double np, double npl, double npr, double vv, double vvp,
                      double psih, double dtn, int im, int nd, double redtl, double sm_grid, double sdx1)
{
    double tn = a;
    if (np != 0.0) {
        double dpx = 0.5 * (np - np) * (1.0 / sdx1);
        double dpz;
        if (npl == 0.0) dpz = (npr - np) * vvp;
        else if (npr == 0.0) dpz = (np - npl) * vv;
        else dpz = 0.5 * (npr - npl) * vv;
        double dalphi = dpx * dpx + dpx * dpx + dpz * dpz;
        if (dalphi != 0.0) {
            double nx = 0.0, ny = 0.0, nz = 1.0;
            double apxm = (a - a) * (1.0 / sdx1);
            double apxp = apxm;
            double apzm = (a - al) * vv;
            double apzp = (ar - a) * vv;
            if (im == 1) { nx = (-1.0) * nx; ny = (-1.0) * ny; nz = (-1.0) * nz; }
            double ax = fmin(nx, 0.0) * apxp + fmax(nx, 0.0) * apxm;
            double ay = fmin(ny, 0.0) * apxp + fmax(ny, 0.0) * apxm;
            double az = fmin(nz, 0.0) * apzp + fmax(nz, 0.0) * apzm;
            double drhs;
            if (im == 1) {
                drhs = (ax + ay + az) + (double)nd * dtn;
                tn = a - redtl * drhs * (1.0 - psih);
            } else {
                drhs = (ax + ay + az) - (double)nd * dtn;
                tn = a - redtl * drhs * psih;
            }
            if (nd == 0 && fabs(np) <= sm_grid) {
                drhs = (ax + ay + az);
                tn = a - redtl * drhs;
            }
        }
    }
    return tn;
}

// The pseudo-time loops of TEMP_EXTP (chphase.f:619-679 for dT/dn, 704-769
// for T) held in registers: window cell i = CPL*lane + s of slot s of the
// lane (one or two cells per lane of the phase's wave; the CPU logic-check
// build keeps the whole window in one lane), neighbours by lane moves, the
// RAT test by a wave vote.  Every cell update equals extp_point in the reference's
// Jacobi order (all cells from the previous stage); the parts of it that
// depend only on NPHI/PSIH are evaluated once per loop.
template <int CPL>
struct XCells {
    double vv[CPL], coef[CPL], dtn[CPL];
    int upd[CPL], act[CPL], small[CPL];   // act: updated and on a live cell
};

#ifdef LICA_EMUL
template <int CPL>
DEV int xcell(int s) { return s; }
template <int CPL>
DEV void xneighbors(const double *a, double *l, double *r)
{
    for (int s = 0; s < CPL; s++) {
        l[s] = s > 0 ? a[s - 1] : a[s];
        r[s] = s + 1 < CPL ? a[s + 1] : a[s];
    }
}
DEV bool xwave_any(bool f) { return f; }
DEV void xwave_urgent(bool on) { (void)on; }
#else
template <int CPL>
DEV int xcell(int s) { return CPL * (int)(threadIdx.x & 63) + s; }
template <int CPL>
DEV void xneighbors(const double *a, double *l, double *r)
{
    if constexpr (CPL == 1) {
        l[0] = lane_up1(a[0]);
        r[0] = lane_dn1(a[0]);
    } else {
        double up = lane_up1(a[1]);
        double dn = lane_dn1(a[0]);
        l[0] = up; r[0] = a[1];
        l[1] = a[0]; r[1] = dn;
    }
}
DEV bool xwave_any(bool f) { return __any((int)f) != 0; }
// Most loops converge in their first 100 iterations or after one halving; a
// loop that needs a second halving is several times longer than the batch's
// typical one (ten times at the tail) and ends its case's stage last: its
// wave takes the SIMD's issue priority over the other waves (which have the
// parallel slack) until the extrapolation ends.
DEV void xwave_urgent(bool on)
{
    if (on) __builtin_amdgcn_s_setprio(2);
    else __builtin_amdgcn_s_setprio(0);
}
#endif

// one forward-Euler stage on a (cells outside [ia,ib] stay as they are);
// the dT/dn pass (ND=0) uses the array being advanced as its DTEMPN.  The
// x/y terms of extp_point, min/max(nx,0) * (v-v)/sdx1 summed twice, are a
// zero of one sign z0 for finite v and NaN otherwise, i.e. exactly
// (v-v)*z0, and AX+AY+AZ = (z+z)+az = z+az; the z factors min/max(nz,0)
// are -1/0 (IM=1) or 0/1 (IM=0).
template <int CPL, int IM, int ND>
DEV void extp_stage(double *a, const XCells<CPL> &c, double redtl, double z0)
{
    double l[CPL], r[CPL], tn[CPL];
    xneighbors<CPL>(a, l, r);
    // every slot computes the update (no branch between the slots' chains);
    // cells that do not move keep their value
    for (int s = 0; s < CPL; s++) {
        const double v = a[s];
        const double z = (v - v) * z0;
        const double apzm = (v - l[s]) * c.vv[s];
        const double apzp = (r[s] - v) * c.vv[s];
        const double az = IM ? (-apzp) + 0.0 * apzm : 0.0 * apzp + apzm;
        const double sum = z + az;
        double drhs;
        if (ND == 0) drhs = IM ? sum + 0.0 * v : sum - 0.0 * v;
        else drhs = IM ? sum + c.dtn[s] : sum - c.dtn[s];
        double t = v - redtl * drhs * c.coef[s];
        if (ND == 0) t = c.small[s] ? v - redtl * sum : t;
        tn[s] = c.act[s] ? t : v;
    }
    for (int s = 0; s < CPL; s++) a[s] = tn[s];
}

// fabs(x / rk - 1.0) >= eps for rk != 0, decided without the division where
// the margin is clear: for |x| <= 1e290 and 1e-290 <= |rk| <= 1e290,
// |x - rk| below (above) eps |rk| (1 -+ 2^-30) puts the exact quotient's
// distance to 1 below (above) eps (1 -+ 2^-31), which the rounding of the
// quotient (2^-53 relative) cannot cross; everything else takes the division
DEV bool rel_change_fails(double x, double rk, double eps)
{
    const double ar = fabs(rk);
    if (ar >= 1.0e-290 && ar <= 1.0e290 && fabs(x) <= 1.0e290) {
        const double d = fabs(x - rk), th = eps * ar;
        if (d < th * (1.0 - 0x1p-30)) return false;
        if (d > th * (1.0 + 0x1p-30)) return true;
    }
    return fabs(x / rk - 1.0) >= eps;
}

// DTEMPN of phase IM at the start of TEMP_EXTP (chphase.f:510-596) for window
// cell i (NS cell k in [ia,ib]).  NPHI vanishes on the three cells at either
// end of the window, so the neighbours i-1, i+1 are read only inside it.
template <int IM>
DEV double extp_dtn_start(const ExtpWin &X, const InstParam &P, const double *PCN, const double *VVDZ,
                          const double *tw, int i, int k, double sm_grid)
{
    const double np = X.nphi[i], pc = PCN[k];
    double v = 0.0;
    if (((fabs(np) > 1.5 * sm_grid) && (np != 0.0)) || ((fabs(pc - 0.5) == 0.5) && (np != 0.0))) {
        const double npl = X.nphi[i - 1], npr = X.nphi[i + 1];
        double dpx = 0.5 * (np - np) * (1.0 / P.sdx1);
        double dtdx = 0.5 * (tw[i] - tw[i]) * (1.0 / P.sdx1);
        double dpz, dtdz;
        if (npl == 0.0) { dpz = (npr - np) * VVDZ[k + 1]; dtdz = (tw[i + 1] - tw[i]) * VVDZ[k + 1]; }
        else if (npr == 0.0) { dpz = (np - npl) * VVDZ[k]; dtdz = (tw[i] - tw[i - 1]) * VVDZ[k]; }
        else { dpz = 0.5 * (npr - npl) * VVDZ[k]; dtdz = 0.5 * (tw[i + 1] - tw[i - 1]) * VVDZ[k]; }
        double dalphi = dpx * dpx + dpx * dpx + dpz * dpz;
        if (dalphi != 0.0) {
            double nx = 0.0, ny = 0.0, nz = 1.0;
            if (IM == 1) v = nx * dtdx + ny * dtdx + nz * dtdz * X.psih[i];
            else v = nx * dtdx + ny * dtdx + nz * dtdz * (1.0 - X.psih[i]);
        }
    } else if (((fabs(np) <= 1.5 * sm_grid) || (fabs(pc - 0.5) != 0.5)) && ((np != 1.0e-7) && (np != 0.0))) {
        if (IM == 1) v = (tw[i] - P.tsat) / np * X.psih[i];
        else v = (tw[i] - P.tsat) / np * (1.0 - X.psih[i]);
    }
    return v;
}

// TEMP_EXTP's two pseudo-time loops for phase IM on the LDS window tw, in
// registers of one wave.  Returns ST_EXTRAP after 20 halvings.
// RAT < EPS over the updated cells (a NaN ratio never counts, as with MAX)
// is the wave's vote that no cell's ratio reaches EPS.
template <int CPL, int IM>
DEV int temp_extp_loops(const ExtpWin &X, const InstParam &P, const double *PCN, const double *VVDZ, double *tw, int ia,
                        int ib, double sm_grid, double eps, double *iters, double *halv)
{
    const int im = IM;
    const int nlo = X.nlo, nw = X.nhi - X.nlo + 1;
    const double inv_sdx1 = 1.0 / P.sdx1;
    double nx = 0.0;
    if (im == 1) nx = (-1.0) * nx;
    const double fmn = fmin(nx, 0.0), fmx = fmax(nx, 0.0);
    const double z0 = fmn * (0.0 * inv_sdx1) + fmx * (0.0 * inv_sdx1);
    XCells<CPL> c;
    double a[CPL], rk[CPL];
    for (int s = 0; s < CPL; s++) {
        const int i = xcell<CPL>(s), k = nlo + i;
        const bool ok = i < nw;
        const double np = ok ? X.nphi[i] : 0.0;
        const double npl = (ok && i > 0) ? X.nphi[i - 1] : 0.0;
        const double npr = (ok && i + 1 < nw) ? X.nphi[i + 1] : 0.0;
        const double vv = ok ? VVDZ[k] : 0.0, vvp = ok ? VVDZ[k + 1] : 0.0;
        const double ps = ok ? X.psih[i] : 0.0;
        // the static part of extp_point
        int live = 0;
        if (np != 0.0) {
            double dpx = 0.5 * (np - np) * (1.0 / P.sdx1);
            double dpz;
            if (npl == 0.0) dpz = (npr - np) * vvp;
            else if (npr == 0.0) dpz = (np - npl) * vv;
            else dpz = 0.5 * (npr - npl) * vv;
            double dalphi = dpx * dpx + dpx * dpx + dpz * dpz;
            live = dalphi != 0.0;
        }
        c.vv[s] = vv;
        c.coef[s] = (im == 1) ? (1.0 - ps) : ps;
        c.small[s] = fabs(np) <= sm_grid;
        c.act[s] = live && ok && k >= ia && k <= ib;
        c.upd[s] = ok && k >= ia && k <= ib;
        c.dtn[s] = c.upd[s] ? extp_dtn_start<IM>(X, P, PCN, VVDZ, tw, i, k, sm_grid) : 0.0;
        a[s] = c.dtn[s];
    }
    // (a) dT/dn (ND=0), continuing from the current field after a halving
    double redtl = 0.2 * sm_grid;
    int halvings = 0;
    for (;;) {
        int conv = 0;
        for (int it = 1; it <= 100; it++) {
            for (int s = 0; s < CPL; s++) rk[s] = a[s];
            extp_stage<CPL, IM, 0>(a, c, redtl, z0);
            extp_stage<CPL, IM, 0>(a, c, redtl, z0);
            for (int s = 0; s < CPL; s++) a[s] = c.upd[s] ? 0.75 * rk[s] + 0.25 * a[s] : a[s];
            extp_stage<CPL, IM, 0>(a, c, redtl, z0);
            bool fail = false;
            for (int s = 0; s < CPL; s++) {
                a[s] = c.upd[s] ? 1.0 / 3.0 * rk[s] + 2.0 / 3.0 * a[s] : a[s];
                if (!c.upd[s]) continue;
                // rk = 0: |rk/a - 1| is 1 for any a but 0 and NaN
                if (rk[s] != 0.0) fail = fail || rel_change_fails(a[s], rk[s], eps);
                else fail = fail || (a[s] != 0.0 && a[s] == a[s]);
            }
            *iters += 1.0;
            if (!xwave_any(fail)) { conv = 1; break; }
        }
        if (conv) break;
        if (++halvings > 20) { xwave_urgent(false); return ST_EXTRAP; }
        *halv += 1.0;
        redtl = 0.5 * redtl;
        xwave_urgent(halvings >= 2);
    }
    for (int s = 0; s < CPL; s++) { c.dtn[s] = a[s]; c.small[s] = 0; }
    // (b) T (ND=1) with the extrapolated dT/dn as source
    double t[CPL];
    for (int s = 0; s < CPL; s++) { int i = xcell<CPL>(s); t[s] = (i < nw) ? tw[i] : 0.0; }
    redtl = 0.2 * sm_grid;
    halvings = 0;
    for (;;) {
        int conv = 0;
        for (int it = 1; it <= 100; it++) {
            for (int s = 0; s < CPL; s++) { rk[s] = t[s]; a[s] = t[s]; }
            extp_stage<CPL, IM, 1>(a, c, redtl, z0);
            extp_stage<CPL, IM, 1>(a, c, redtl, z0);
            for (int s = 0; s < CPL; s++) a[s] = c.upd[s] ? 0.75 * rk[s] + 0.25 * a[s] : a[s];
            extp_stage<CPL, IM, 1>(a, c, redtl, z0);
            bool fail = false;
            for (int s = 0; s < CPL; s++) {
                a[s] = c.upd[s] ? 1.0 / 3.0 * rk[s] + 2.0 / 3.0 * a[s] : a[s];
                t[s] = c.upd[s] ? a[s] : t[s];
                if (!c.upd[s]) continue;
                if (rk[s] != 0.0) fail = fail || rel_change_fails(t[s], rk[s], eps);
            }
            *iters += 1.0;
            if (!xwave_any(fail)) { conv = 1; break; }
        }
        if (conv) break;
        if (++halvings > 20) { xwave_urgent(false); return ST_EXTRAP; }
        *halv += 1.0;
        redtl = 0.5 * redtl;
        xwave_urgent(halvings >= 2);
    }
    xwave_urgent(false);
    for (int s = 0; s < CPL; s++) {
        int i = xcell<CPL>(s);
        if (i < nw) tw[i] = t[s];
    }
    return ST_OK;
}

// PHINSINTP (chphase.f:1258-1325) on the flat column; call inside SERIAL
DEV double phinsintp(const BandSh &B, const CaseP &cp, int ko, int *ic)
{
    const InstParam &P = *cp.P;
    const int *KCOU2 = cp.AI(NSI_KCOU2), *KCOUMP2 = cp.AI(NSI_KCOUMP2);
    const double *ZP = cp.A(NS_ZP), *ZPF = cp.R(RF_ZPF);
    int k1 = KCOU2[ko], k2 = KCOUMP2[ko];
    double xd = P.phins_xd, yd = P.phins_yd, zd;
    zd = (k1 == k2) ? 0.0 : (ZP[ko] - ZPF[k1]) / (ZPF[k2] - ZPF[k1]);
    double a = phi_get(B, k1, P.n3f), b = phi_get(B, k2, P.n3f);
    double c00 = a * (1.0 - xd) + a * xd, c10 = a * (1.0 - xd) + a * xd;
    double c01 = b * (1.0 - xd) + b * xd, c11 = b * (1.0 - xd) + b * xd;
    double c0 = c00 * (1.0 - yd) + c10 * yd, c1 = c01 * (1.0 - yd) + c11 * yd;
    double v = c0 * (1.0 - zd) + c1 * zd;
    if (v > 0.0) *ic = 1;
    if (v < 0.0) *ic = 0;
    if (a == 0.0 || b == 0.0) v = 0.0;
    return v;
}

// TEMPASSIGN (chphase.f:372-463): x/y weights are zero on the flat column
DEV double tempassign(int nlo, const CaseP &cp, int k, const double *tw)
{
    const double *ZP = cp.A(NS_ZP), *ZPF = cp.R(RF_ZPF), *VDZ = cp.A(NS_VDZ);
    const int *KCMV = cp.RI(RFI_KCOUMP_VEL);
    int kk = KCMV[k], kk1, kk2;
    double zd;
    if (ZP[kk] > ZPF[k]) {
        kk1 = kk - 1; kk2 = kk;
        zd = (kk1 == kk2) ? 0.0 : (ZPF[k] - (ZP[kk] - VDZ[kk])) / (ZP[kk2] - ZP[kk1]);
    } else {
        kk1 = kk; kk2 = kk + 1;
        zd = (kk1 == kk2) ? 0.0 : (ZPF[k] - ZP[kk1]) / (ZP[kk2] - ZP[kk1]);
    }
    double xd = 0.0, yd = 0.0;
    double t_kk = tw[kk - nlo], t_k2 = tw[kk2 - nlo];
    double c00 = t_kk * (1.0 - xd) + t_kk * xd;
    double c10 = t_kk * (1.0 - xd) + t_kk * xd;
    double c01 = t_k2 * (1.0 - xd) + t_k2 * xd;
    double c11 = t_k2 * (1.0 - xd) + t_k2 * xd;
    double c0 = c00 * (1.0 - yd) + c10 * yd;
    double c1 = c01 * (1.0 - yd) + c11 * yd;
    return c0 * (1.0 - zd) + c1 * zd;
}

DEV double tf_get(const BandSh &B, const double *tfw, int k)
{
    if (k < B.wlo || k > B.whi) return 0.0;
    return tfw[k - B.wlo];
}

DEV void body_temp_grad(const BatchDev &bd, int inst, TgradSh &TS)
{
    InstState &S = bd.st[inst];
    if (S.status || S.m == 1) return;   // TEMP_GRAD runs from the second step (1d.f:3654-3667)
    BandSh &B = TS.B;
    CaseP cp = case_ptrs(bd, inst);
    const InstParam &P = *cp.P;
    const int n = P.n, n3 = P.n3, n3f = P.n3f;
    const double *T = cp.A(NS_T), *PCN = cp.A(NS_PSI_CN), *VVDZ = cp.A(NS_VVDZ);
    const double *ZP = cp.A(NS_ZP), *Z = cp.A(NS_Z), *ZPF = cp.R(RF_ZPF), *ZF = cp.R(RF_ZF);
    const int *KCMV = cp.RI(RFI_KCOUMP_VEL);
    const int *KCOU1 = cp.AI(NSI_KCOU1), *KCOU2 = cp.AI(NSI_KCOU2);
    int *IMASK = cp.AI(NSI_IMASK);
    double *SFGAS = cp.R(RF_SFGAS), *SFGAS2 = cp.R(RF_SFGAS2);
    const double *VPREF = cp.R(RF_VPREFIX);

    band_load(B, P, S);
    PFOR(j, 0, (LICA_MLAYER + 1) * LICA_MAXP) { S.mflvs[j / LICA_MAXP][j % LICA_MAXP] = 0.0; }
    PFOR_END
    {
        // NS window under the band
        const int ne = B.off[LICA_MLAYER + 1];
        const double kmn = red_min(0, ne, 1073741824.0, [&](int e) { return (double)KCMV[B.ent_k[e]]; }, B.red);
        const double kmx = red_max(0, ne, -1.0, [&](int e) { return (double)KCMV[B.ent_k[e]]; }, B.red);
        SERIAL {
            int kmin = (int)kmn, kmax = (int)kmx;
            if (kmax < 0) { kmin = 1; kmax = 1; }
            TS.nlo = kmin - 3 < 0 ? 0 : kmin - 3;
            TS.nhi = kmax + 3 > n3 ? n3 : kmax + 3;
            // CAL_PSI_TMP's rows reach one cell beyond either end
            if (TS.nhi - TS.nlo + 3 > NSWMAX) B.status = ST_WINDOW;
        }
        SERIAL_END
    }
    if (B.status) { SERIAL { S.status = B.status; } SERIAL_END return; }
    const int nlo = TS.nlo, nhi = TS.nhi;
    const ExtpWin X{nlo, nhi, TS.t0, B.wf, B.phs, B.mfd, TS.t0};
    PFOR(i, 0, WMAX) { B.phw[i] = 0.0; B.tf0[i] = 0.0; B.tf1[i] = 0.0; }
    PFOR_END
    PFOR(e, 0, B.off[LICA_MLAYER + 1]) { B.phw[B.ent_k[e] - B.wlo] = S.alphi[B.ent_n[e]][B.ent_q[e]]; }
    PFOR_END
    PFOR(k, nlo, nhi + 1) {
        int i = k - nlo;
        X.t0[i] = T[k]; X.t1[i] = T[k]; X.nphi[i] = 0.0; X.psih[i] = 0.0;
    }
    PFOR_END
    // (1) phase of the NS cells under the band (chphase.f:174-211).  The
    // lookups and PHINSINTP of every entry run in parallel (PHINSINTP is a
    // function of the NS cell; xe = the phase flag it sets, -1 if it leaves
    // IC as it is); the reference's order of updates -- which entry sets a
    // cell, the running flag IC -- is replayed serially from LDS.  B.pm1,
    // B.pm2 and B.mcell are free until step (3).
    const int e0 = B.off[1], e1 = B.off[LICA_NN2 + 1], e2 = B.off[LICA_MLAYER], e3 = B.off[LICA_MLAYER + 1];
    int *xk = B.mcell, *xe = B.mcell + LICA_MAXENT;
    PFOR(e, e0, e1) {
        const int kk = KCMV[B.ent_k[e]];
        int ic = -1;
        B.pm1[e] = phinsintp(B, cp, kk, &ic);
        B.pm2[e] = PCN[kk];
        xk[e] = kk;
        xe[e] = ic;
    }
    PFOR_END
    PFOR(e, e2, e3) {
        const int k = B.ent_k[e];
        const int kk = KCMV[k];
        const double dz = ZPF[k] - ZP[kk];
        const double d2 = P.dxy2_min + dz * dz;
        B.pm1[e] = phi_get(B, k, n3f) + sqrt(d2);
        B.pm2[e] = PCN[kk];
        xk[e] = kk;
    }
    PFOR_END
    SERIAL {
        int ic = 0;
        for (int e = e0; e < e1; e++) {
            const int i = xk[e] - nlo;
            if (X.nphi[i] == 0.0) {
                if (xe[e] >= 0) ic = xe[e];
                X.nphi[i] = B.pm1[e];
                const double pc = B.pm2[e];
                if (fabs(pc - 0.5) != 0.5) X.psih[i] = (double)ic;
                else X.psih[i] = pc;
            }
        }
        for (int e = e2; e < e3; e++) {
            const int i = xk[e] - nlo;
            X.nphi[i] = fmin(X.nphi[i], B.pm1[e]);
            X.psih[i] = B.pm2[e];
        }
    }
    SERIAL_END
    // (2) TEMP_EXTP
    const double sm_grid = fmin(fmin(2.0 * P.sdxf, 2.0 * P.sdyf), 2.0 * P.sdzf);
    const double eps = 1.0e-4;
    const int ia = nlo < 1 ? 1 : nlo, ib = nhi > n ? n : nhi;
    // the two phases are independent (phase im advances its own copy of the
    // window, X.t0 or X.t1, from its own DTEMPN): on the GPU wave im runs
    // phase im.  Counters and status combine in the reference's order: a
    // failed phase 0 ends TEMP_GRAD before phase 1 starts, so phase 1's
    // counts are dropped then.
#ifdef LICA_EMUL
    for (int im = 0; im <= 1; im++) {
        double iters = 0.0, halv = 0.0;
        TS.xst[im] = im ? temp_extp_loops<NSWMAX, 1>(X, P, PCN, VVDZ, X.t1, ia, ib, sm_grid, eps, &iters, &halv)
                        : temp_extp_loops<NSWMAX, 0>(X, P, PCN, VVDZ, X.t0, ia, ib, sm_grid, eps, &iters, &halv);
        TS.xit[im] = iters;
        TS.xhv[im] = halv;
    }
#else
    {
        const int im = (int)(threadIdx.x >> 6);
        double iters = 0.0, halv = 0.0;
        int xst;
        if (nhi - nlo + 1 <= 64)
            xst = im ? temp_extp_loops<1, 1>(X, P, PCN, VVDZ, X.t1, ia, ib, sm_grid, eps, &iters, &halv)
                     : temp_extp_loops<1, 0>(X, P, PCN, VVDZ, X.t0, ia, ib, sm_grid, eps, &iters, &halv);
        else
            xst = im ? temp_extp_loops<2, 1>(X, P, PCN, VVDZ, X.t1, ia, ib, sm_grid, eps, &iters, &halv)
                     : temp_extp_loops<2, 0>(X, P, PCN, VVDZ, X.t0, ia, ib, sm_grid, eps, &iters, &halv);
        if ((threadIdx.x & 63) == 0) {
            TS.xst[im] = xst;
            TS.xit[im] = iters;
            TS.xhv[im] = halv;
        }
    }
#endif
    BARRIER();
    SERIAL {
        S.n_extrap_iters += TS.xit[0];
        S.n_halvings += TS.xhv[0];
        if (TS.xst[0]) {
            S.status = TS.xst[0];
        } else {
            S.n_extrap_iters += TS.xit[1];
            S.n_halvings += TS.xhv[1];
            if (TS.xst[1]) S.status = TS.xst[1];
        }
    }
    SERIAL_END
    if (TS.xst[0] || TS.xst[1]) return;
    // (3) TF on the T-band, (4)-(6) TF_GRAD
    PFOR(e, B.off[1], B.off[LICA_NN1 + 1]) {
        int k = B.ent_k[e];
        B.tf0[k - B.wlo] = tempassign(X.nlo, cp, k, X.t0);
        B.tf1[k - B.wlo] = tempassign(X.nlo, cp, k, X.t1);
    }
    PFOR_END
    psif_mtp_band(B, P, LICA_NN1, B.pm1);
    const Fill fill1 = fill_params_all(B, P);
    PFOR(e, B.off[1], B.off[LICA_NN1 - 2 + 1]) {
        int k = B.ent_k[e];
        double gx = 0.5 * (0.0 - 0.0) * (1.0 / P.sdxf);
        double gz = 0.5 * (phi_get(B, k + 1, n3f) - phi_get(B, k - 1, n3f)) * P.ssdzf;
        double agi = 1.0 / sqrt(gx * gx + gx * gx + gz * gz);
        double dtx = 0.5 * (0.0 - 0.0) * (1.0 / P.sdxf);
        double dtz0 = 0.5 * (tf_get(B, B.tf0, k + 1) - tf_get(B, B.tf0, k - 1)) * P.ssdzf;
        double dtdn0 = (dtx * gx + dtx * gx + dtz0 * gz) * agi;
        double dtz1 = 0.5 * (tf_get(B, B.tf1, k + 1) - tf_get(B, B.tf1, k - 1)) * P.ssdzf;
        double dtdn1 = (dtx * gx + dtx * gx + dtz1 * gz) * agi;
        S.mflvs[B.ent_n[e]][B.ent_q[e]] = P.ja * P.prai * P.rei * (-P.tcp * dtdn1 + P.tcm * dtdn0);
    }
    PFOR_END
    // VOL_TOT_MFDOT: provisional advance, reinitialisation, new volume
    PFOR(i, 0, WMAX) { B.phs[i] = B.phw[i]; }
    PFOR_END
    PFOR(e, B.off[1], B.off[LICA_NN1 - 2 + 1]) {
        int k = B.ent_k[e];
        double mf = S.mflvs[B.ent_n[e]][B.ent_q[e]];
        B.mfd[e] = (phi_get(B, k, n3f) > 0.0) ? mf / P.denr : mf;
    }
    PFOR_END
    const double dtconst = S.dtconst;
    PFOR(e, B.off[1], B.off[LICA_NN1 - 2 + 1]) {
        int k = B.ent_k[e];
        double a = phi_get(B, k, n3f);
        double v = a - 2.0 * dtconst * B.mfd[e];
        v = 0.75 * a + 0.25 * v;
        v = v - dtconst * B.mfd[e];
        v = 1.0 / 3.0 * a + 2.0 / 3.0 * v;
        phi_set(B, k, n3f, v);
    }
    PFOR_END
    reinit_band(B, P, S, LICA_NN1, LICA_NN2, LICA_NN2);
    psif_mtp_band(B, P, LICA_NN1, B.pm2);
    double *pmw2 = B.tf0;
    int *tmk2 = B.mask_s;
    psif_view(B, LICA_NN1, B.pm2, pmw2, tmk2);
    // the interface plane: the ascending scan keeps the last sign change,
    // the highest one of the fill scan
    const Fill f2 = fill_params_all(B, P);
    SERIAL {
        S.place = (f2.knum > 0) ? ZPF[f2.khigh] + fabs(phi_get(B, f2.khigh, n3f)) : 0.0;
        S.vol_tot_iphs = vol_implicit(B, P, VPREF, f2, pmw2, tmk2);
    }
    SERIAL_END
    // SFGAS/SFGAS2: clear the previous window, then write the new one
    const int sfl = S.sf_lo, sfh = S.sf_hi;
    PFOR(k, sfl, sfh + 1) { SFGAS[k] = 0.0; SFGAS2[k] = 0.0; }
    PFOR_END
    PFOR(e, B.off[1], B.off[LICA_NN1 - 2 + 1]) {
        int nl = B.ent_n[e], q = B.ent_q[e], k = B.ent_k[e];
        if (S.mflvs[nl][q] != 0.0) {
            if (fabs(B.pm1[e] - 0.5) != 0.5) {
                double d = B.pm2[e] - B.pm1[e];
                if (S.alphi[nl][q] > 0.0) { SFGAS[k] = -d * P.denr; SFGAS2[k] = -d; }
                else { SFGAS[k] = -d; SFGAS2[k] = -d; }
            }
        }
    }
    PFOR_END
    SERIAL {
        int lo = 1 << 30, hi = -1;
        for (int e = B.off[1]; e < B.off[LICA_NN1 - 2 + 1]; e++) {
            int k = B.ent_k[e];
            if (k < lo) lo = k;
            if (k > hi) hi = k;
        }
        if (hi < 0) { lo = 1; hi = 0; }
        S.sf_lo = lo; S.sf_hi = hi;
    }
    SERIAL_END
    PFOR(i, 0, WMAX) { B.phw[i] = B.phs[i]; }
    PFOR_END
    // (7) CAL_PSI_TMP(1) and IMASK_PH (chphase.f:300-358)
    double *pmw1 = B.tf1;
    int *tmk1 = B.slot;
    psif_view(B, LICA_NN1, B.pm1, pmw1, tmk1);
    const int ka = (nlo - 1) < 1 ? 1 : nlo - 1, kb = (nhi + 1) > n ? n : nhi + 1;
    PFOR(k, ka, kb + 1) {
        auto fpm = [&](int kk) { return psif_implicit(B, fill1, pmw1, tmk1, kk); };
        double ps = box_avg(ZF, P.ssdzf, fpm, Z[k], Z[k + 1], KCOU2[k], KCOU1[k + 1]);
        if (ps <= 1.0e-3) ps = 0.0;
        else if (ps >= 0.999) ps = 1.0;
        X.psitmp[k - ka] = ps;
    }
    PFOR_END
    const int iml = S.im_lo, imh = S.im_hi;
    PFOR(k, iml, imh + 1) { IMASK[k] = 0; }
    PFOR_END
    const int ja = (ka + 1) < 2 ? 2 : ka + 1;
    PFOR(k, ja, kb + 1) {
        if (fabs(X.psitmp[k - ka] - X.psitmp[k - 1 - ka]) != 0.0) IMASK[k] = 1;
    }
    PFOR_END
    SERIAL {
        S.im_lo = ja; S.im_hi = kb;
        if (B.status) S.status = B.status;
    }
    SERIAL_END
}

// ---------------------------------------------------------------------
// Lead workgroups of TEMP_GRAD.  A case's extrapolation takes from about a
// hundred to over a thousand iterations per substep (one dependent chain per
// phase), and a case that is slow in one substep is slow in the next: if it
// starts late it alone sets the end of the launch.  A launch with lead slots
// runs LICA_TG_LEAD lead workgroups before its cases.  Lead j takes the case
// whose key is in slot j of the previous substep's row -- the costliest case
// among the positions p = j (mod LICA_TG_LEAD) -- and the workgroup of that
// case's position returns at once.  Both read the same key, which no
// workgroup of the launch writes (they record into the current substep's
// row), and decide from it and the case's step counter alone, so every case
// runs exactly once.  Key: substep index 3m + msub (28 bits), cost (the
// slower phase's iterations, 12 bits), position (24 bits); atomicMax keeps
// the latest substep's costliest case.  A key of another substep (the first
// launch, a restored state) leads nothing.
#define TG_SUB_MASK ((1ull << 28) - 1)
DEV unsigned long long tg_sub(const InstState &S, int msub)
{
    return (unsigned long long)(3 * S.m + msub) & TG_SUB_MASK;
}

// position of the case that workgroup wg runs (-1: none)
DEV int tg_lead_pos(const BatchDev &bd, int wg)
{
    const bool lead = wg < LICA_TG_LEAD;
    const int pos = lead ? -1 : wg - LICA_TG_LEAD;
    const int j = lead ? wg : pos % LICA_TG_LEAD;
    const int prev = (bd.msub == 1) ? 3 : bd.msub - 1;
    const unsigned long long key = bd.lead[prev * LICA_TG_LEAD + j];
    const int kp = (int)(key & 0xFFFFFFull);
    bool take = key != 0ull && kp < bd.ninst && kp % LICA_TG_LEAD == j;
    if (take) take = (key >> 36) == ((tg_sub(bd.st[bd.order[kp]], bd.msub) + TG_SUB_MASK) & TG_SUB_MASK);
    if (lead) return take ? kp : -1;
    return (take && kp == pos) ? -1 : pos;
}

#ifndef LICA_EMUL
// the case's cost into the current substep's row (TS.xit negative: no
// extrapolation ran)
DEV void tg_lead_record(const BatchDev &bd, int inst, int pos, const TgradSh &TS)
{
    if (!(TS.xit[0] >= 0.0 && TS.xit[1] >= 0.0)) return;
    const double c = fmax(TS.xit[0], TS.xit[1]);
    const unsigned long long cost = (c >= 4095.0) ? 4095ull : (unsigned long long)c;
    const unsigned long long key = (tg_sub(bd.st[inst], bd.msub) << 36) | (cost << 24) | (unsigned long long)pos;
    atomicMax(&bd.lead[bd.msub * LICA_TG_LEAD + pos % LICA_TG_LEAD], key);
}
#endif
