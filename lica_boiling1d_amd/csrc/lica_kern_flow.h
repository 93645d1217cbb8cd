// NS-column kernels: momentum + projection (1d.f:1603-1627, 1749-1923,
// 3006-3169, 3180-3627), CONVRGE1 (3306-3362) and the initial state
// (MAKEFLD 1d.f:821-946, LVSINIT two_phase.f:116-458, GRID_COUPLING(1)).
#pragma once
#include "lica_kern_tgrad.h"
#include "lica_kern_mg.h"

// 1/DENF_Z from PSI_Z (CAL_CONT_BC 1d.f:2986-2994 incl. the end copies)
DEV double denf_zi(const InstParam &P, const double *PZ, int k)
{
    int kk = k < 1 ? 1 : (k > P.n ? P.n : k);
    double d = P.denm + P.den_diff * PZ[kk];
    return 1.0 / d;
}

// NREF: the most refinement steps its partitioned solve may take (TRI_NREF
// for the direct Poisson mode's kernel, 0 for the others, whose kernel then
// carries no refinement code and needs fewer registers)
template <int NREF>
DEV void body_momentum(const BatchDev &bd, int inst, BlkSh &sh)
{
    InstState &S = bd.st[inst];
    if (S.status) return;
    CaseP cp = case_ptrs(bd, inst);
    const InstParam &P = *cp.P;
    const int n = P.n, n3 = P.n3, msub = bd.msub;
    double *__restrict__ W = cp.A(NS_W), *__restrict__ PP = cp.A(NS_P), *__restrict__ PZN = cp.A(NS_PSI_ZN), *__restrict__ PCN = cp.A(NS_PSI_CN);
    double *__restrict__ DFZ = cp.A(NS_DF_Z), *__restrict__ VFZ = cp.A(NS_VF_Z);
    double *__restrict__ PZ = cp.A(NS_PSI_Z), *__restrict__ PC = cp.A(NS_PSI_C), *__restrict__ RK3ZO = cp.A(NS_RK3ZO);
    double *__restrict__ TA = cp.A(NS_TA), *__restrict__ TB = cp.A(NS_TB), *__restrict__ TC = cp.A(NS_TC), *__restrict__ TG = cp.A(NS_TG);
    double *__restrict__ PA = cp.A(NS_PA), *__restrict__ PB = cp.A(NS_PB), *__restrict__ PCc = cp.A(NS_PC), *__restrict__ PG = cp.A(NS_PG);
    const double *__restrict__ ZP = cp.A(NS_ZP), *__restrict__ Z = cp.A(NS_Z), *__restrict__ SSDZ = cp.A(NS_SSDZ), *__restrict__ VVDZ = cp.A(NS_VVDZ);
    const double *__restrict__ AKW = cp.A(NS_AKW), *__restrict__ CKW = cp.A(NS_CKW), *__restrict__ ZF = cp.R(RF_ZF);
    const double *__restrict__ SFGAS = cp.R(RF_SFGAS), *__restrict__ SFGAS2 = cp.R(RF_SFGAS2);
    const int *__restrict__ KCOU1 = cp.AI(NSI_KCOU1), *__restrict__ KCOU2 = cp.AI(NSI_KCOU2), *__restrict__ IMASK = cp.AI(NSI_IMASK);
    const long long m = S.m;
    const double dt = S.dt, alpha = S.alpha, acoef = S.acoef, dtconst = S.dtconst, dtconsti = S.dtconsti;
    const int sfl = S.sf_lo, sfh = S.sf_hi;
    // a refined range [k1,k2] that misses the nonzero SFGAS window averages to 0;
    // the boxes that touch it read ZF/SFGAS/SFGAS2 staged in LDS
    auto touches = [&](int k1, int k2) { return !(k2 < sfl || k1 > sfh); };
    stage_refwin(*sh.win, sfl - 2 * P.r - 4, sfh + 2 * P.r + 4, P.m3l, P.n3f + P.r, ZF, SFGAS, SFGAS2);
    const RefWin &win = *sh.win;
    auto zfa = [&](int kk) { return (kk >= win.lo && kk <= win.hi + 1) ? win.zf[kk - win.lo] : ZF[kk]; };
    auto fsf = [&](int kk) { return (kk >= win.lo && kk <= win.hi) ? win.f0[kk - win.lo] : SFGAS[kk]; };
    auto fsf2 = [&](int kk) { return (kk >= win.lo && kk <= win.hi) ? win.f1[kk - win.lo] : SFGAS2[kk]; };

    if (m != 1) {
        // CAL_CONTINUITY2 (1d.f:3006-3169).  Outside region 2 -- cells more
        // than two from every jump of PSI_CN -- it only clamps PSI_Z and
        // PSI_C to [0, 1], which CONTI_Z/C of this substep already did (a
        // NaN stays NaN either way): the loop runs over the cells within
        // reach of a jump.
        int jlo, jhi;
        jump_range(PCN, n, sh.red, &jlo, &jhi);
        const int sa = (jhi >= 0) ? (jlo - 2 < 1 ? 1 : jlo - 2) : 1;
        const int sb = (jhi >= 0) ? (jhi + 3 > n ? n : jhi + 3) : 0;
        PFOR(k, sa, sb + 1) {
            int rg = region_of(PCN, n, k);
            if (k >= 2) {
                double pz = PZ[k];
                if (rg == 2 && touches(KCOU2[k], KCOU1[k + 1]))
                    pz = pz - box_avg_acc(zfa, P.ssdzf, fsf2, ZP[k - 1], ZP[k], KCOU2[k], KCOU1[k + 1]);
                else if (rg == 2)
                    pz = pz - 0.0;
                if (pz > 1.0) pz = 1.0;
                else if (pz < 0.0) pz = 0.0;
                PZ[k] = pz;
            }
            double pc = PC[k];
            if (rg == 2 && touches(KCOU2[k], KCOU1[k + 1]))
                pc = pc - box_avg_acc(zfa, P.ssdzf, fsf2, Z[k], Z[k + 1], KCOU2[k], KCOU1[k + 1]);
            if (pc > 1.0) pc = 1.0;
            else if (pc < 0.0) pc = 0.0;
            PC[k] = pc;
        }
        PFOR_END
        SERIAL {
            PZ[1] = PZ[2]; PZ[n3] = PZ[n];
            PC[1] = PC[2]; PC[n3] = PC[n];
        }
        SERIAL_END
    }
    // RHSZ (1d.f:1603-1745), RHS1/DENF_Z - W (1557-1560), LHSW (1749-1809)
    const double g = P.gamma[msub - 1], ro = P.ro[msub - 1], a2 = 2.0 * alpha;
    auto lhsw_row = [&](int k, double &ta, double &tb, double &tc, double &tg) {
        int kp1 = k + 1, km1 = k - 1;
        double wt = 0.5 * (W[kp1] + W[k]);
        double wb = 0.5 * (W[k] + W[km1]);
        double anzz = (DFZ[k] * (wt * wt) - DFZ[km1] * (wb * wb)) * VVDZ[k];
        double anz = -0.0 - 0.0 - anzz;
        double alz5 = (W[kp1] - W[k]) * SSDZ[k];
        double alz6 = (W[k] - W[km1]) * SSDZ[km1];
        double alzz = (VFZ[k] * alz5 - VFZ[km1] * alz6) * VVDZ[k];
        double alz_sum1 = 0.0 + 0.0 + alzz + alzz;
        double alz_sum2 = 0.0 + 0.0;
        double pgr_z = (PP[k] - PP[km1]) * VVDZ[k];
        double den_zn = P.denm + P.den_diff * PZN[k];
        double rk3z = anz + alz_sum2;
        double cn2z = alz_sum1;
        double rhs1 = den_zn * W[k] +
                      (g * rk3z + ro * RK3ZO[k] + a2 * cn2z + a2 * (0.0 - pgr_z) - a2 * 0.0 - a2 * (0.0 + 0.0)) * dt;
        if (IMASK[k] == 1 && fabs(PZN[k] - 0.5) != 0.5)
            rhs1 = rhs1 + g * dt * den_zn * W[k] * (0.0 * 1.0 + 0.0 * 1.0 + (wt - wb) * VVDZ[k]);
        RK3ZO[k] = rk3z;
        double dzi = denf_zi(P, PZ, k);
        rhs1 = rhs1 * dzi - W[k];
        double dk1 = 2.0 * acoef * dzi * VFZ[km1] * AKW[k];
        double dk3 = 2.0 * acoef * dzi * VFZ[k] * CKW[k];
        double dk2 = -dk1 - dk3;
        ta = dk1; tc = dk3; tb = 1.0 + dk2; tg = rhs1;
    };
    // in the multigrid modes the rows reach the solve's registers through
    // LDS (no row arrays); otherwise through TA, TB, TC, TG
    tri_solve_rows<NREF>(TA, TB, TC, TG, 2, n, bd.tri_mode, sh.tri, lhsw_row, [&](int k, double x) { TG[k] = x; });
    // the solve's workspace overlays the window: DIVGS reads it again
    stage_refwin(*sh.win, sfl - 2 * P.r - 4, sfh + 2 * P.r + 4, P.m3l, P.n3f + P.r, ZF, SFGAS, SFGAS2);
    // The solve's correction, BC + BC_FLUX (1d.f:1812-1923), FSM_CHOI
    // (1d.f:3365-3400), DIVGS (3180-3303) and the POISSON matrix (3529-3543)
    // in one pass over chunks of the workgroup's size.  The thread of cell k
    // also evaluates FSM_CHOI's W(k+1) (the same expression of the same
    // values) for DIVGS's difference, and a chunk writes W after a barrier,
    // when every thread of the chunk has read the old W(k), W(k+1) (a later
    // chunk reads only its own cells and the one after it).  W(1) = 0 and
    // W(n3) = W(n) + UBC from the corrected W(n) before FSM_CHOI, with
    // UBC = -(W(n3) - W(1)) SDX1 SDY1 / AA.
    auto bc = [&](double wn) {
        const double w1 = 0.0, wn3 = wn;
        double qq3 = 0.0 + (wn3 - w1) * P.sdx1 * P.sdy1;
        double aa = 0.0 + 2.0 * P.xl * P.yl;
        double ubc = -qq3 / aa;
        W[1] = 0.0;
        W[n3] = wn3 + ubc;
    };
    auto fsm = [&](int k, double zi) {
        const double wk = TG[k] + W[k];
        return wk + dtconst * ((PP[k] - PP[k - 1]) * VVDZ[k]) * zi;
    };
    if (n < 2) {
        SERIAL { bc(W[n]); }
        SERIAL_END
    }
    for (int k0 = 1; k0 <= n; k0 += PSTRIDE) {
        const int k = k0 + PTID;
        double wk = 0.0, wn = 0.0;
        if (k <= n) {
            const double zk = denf_zi(P, PZ, k), zk1 = denf_zi(P, PZ, k + 1);
            if (k >= 2) wk = fsm(k, zk);
            if (k == n) wn = TG[n] + W[n];
            double d = (k <= n - 1) ? (fsm(k + 1, zk1) - wk) * SSDZ[k] : 0.0;
            double ss = touches(KCOU2[k], KCOU1[k + 1]) ? box_avg_acc(zfa, P.ssdzf, fsf, Z[k], Z[k + 1], KCOU2[k], KCOU1[k + 1]) : 0.0;
            d = d - ss * (P.den_diff / (P.denp * P.denm));
            if (k >= 2) d = d * dtconsti;
            double fixkl = (k == 1) ? 1.0 : 0.0;
            double ak = VVDZ[k] * SSDZ[k] * zk * (1.0 - fixkl);
            double ck = VVDZ[k + 1] * SSDZ[k] * zk1;
            PA[k] = ak; PCc[k] = ck; PB[k] = -(ak + ck); PG[k] = d;
        }
        BARRIER();
        if (k >= 2 && k <= n) W[k] = wk;
        if (k == n && n >= 2) bc(wn);
    }
    BARRIER();
}

// POISSON solve (1d.f:3544-3576) of the system written by body_momentum
DEV void body_poisson_direct(const BatchDev &bd, int inst, BlkSh &sh)
{
    InstState &S = bd.st[inst];
    if (S.status) return;
    CaseP cp = case_ptrs(bd, inst);
    const InstParam &P = *cp.P;
    const int n = P.n, n3 = P.n3;
    double *PP = cp.A(NS_P), *PA = cp.A(NS_PA), *PB = cp.A(NS_PB), *PCc = cp.A(NS_PC), *PG = cp.A(NS_PG);
    tri_solve(PA, PB, PCc, PG, 1, n, bd.tri_mode, sh.tri);
    PFOR(k, 1, n + 1) { PP[k] = PG[k]; }
    PFOR_END
    SERIAL {
        PP[0] = PP[1];
        PP[n3] = PP[n];
        S.poisson_solves += 1.0;
    }
    SERIAL_END
}

// POISSON by multigrid (the same system, solved to max|g-Ap| <= rtol max|g|)
DEV void body_poisson_mg(const BatchDev &bd, int inst, MgSh &sh)
{
    InstState &S = bd.st[inst];
    if (S.status) return;
    CaseP cp = case_ptrs(bd, inst);
    const InstParam &P = *cp.P;
    const int n = P.n, n3 = P.n3;
    double *PP = cp.A(NS_P);
    const double *PA = cp.A(NS_PA), *PCc = cp.A(NS_PC), *PG = cp.A(NS_PG);
    const bool warm = LICA_MG_WARM(bd.pmode) && bd.guess && bd.hist;
    const double *x0 = warm ? bd.guess + P.gs_off : nullptr;
    double res = 0.0;
    // PB = -(PA + PC) row by row (the momentum stage's POISSON matrix): the
    // solve evaluates it instead of reading it
    const int cyc = mg_solve(PA, nullptr, PCc, PG, PP, n, bd.mg + P.mg_off, bd.pmode, bd.mg_rtol, bd.mg_maxc, x0, sh,
                             &res);
    if (warm) {
        // the solution joins this substep's history ring
        const int s = bd.msub - 1, head = S.hist_head[s];
        double *Hs = bd.hist + P.hist_off + ((size_t)s * MG_HIST + (size_t)head) * (size_t)n;
        PFOR(i, 0, n) { Hs[i] = PP[i + 1]; }
        PFOR_END
        SERIAL {
            S.hist_head[s] = (head + 1) % MG_HIST;
            S.hist_cnt[s] = (S.hist_cnt[s] < MG_HIST) ? S.hist_cnt[s] + 1 : MG_HIST;
        }
        SERIAL_END
    }
    SERIAL {
        PP[0] = PP[1];
        PP[n3] = PP[n];
        S.poisson_solves += 1.0;
        S.poisson_cycles += (cyc < 0) ? (double)bd.mg_maxc : (double)cyc;
        S.poisson_resid = fmax(S.poisson_resid, res);
        if (cyc < 0) S.status = ST_NOT_CONVERGED;
    }
    SERIAL_END
}

// capture rows of every case (3 x (n+2) at 3*gs_off): before the POISSON
// solve (which 0) the right-hand side and the starting point -- the
// warm-start base in the warm-started modes, else the last solution
// P + pshift -- after it (which 1) the new solution
DEV void body_capture(const BatchDev &bd, int inst, int which)
{
    InstState &S = bd.st[inst];
    if (S.status || !bd.cap) return;
    CaseP cp = case_ptrs(bd, inst);
    const InstParam &P = *cp.P;
    const double *PP = cp.A(NS_P), *PG = cp.A(NS_PG);
    double *base = bd.cap + 3 * P.gs_off;
    double *dst = base + (long long)which * (P.n + 2);
    double *rhs = base + 2LL * (P.n + 2);
    const double sh0 = (which == 0) ? S.pshift : 0.0;
    const double *start = (which == 0 && LICA_MG_WARM(bd.pmode) && bd.guess) ? bd.guess + P.gs_off : nullptr;
    PFOR(k, 1, P.n + 1) {
        dst[k] = start ? start[k] : PP[k] + sh0;
        if (which == 0) rhs[k] = PG[k];
    }
    PFOR_END
}

// Warm start of the multigrid solve (LICA_POISSON_PROJECTED, _LEARNED): the
// POISSON solution of a substep differs from the one of the same substep a
// step earlier by a slowly drifting factor and a small local change, so the
// base is the combination B = sum_j c_j v_j of the case's solutions at this
// substep in the last MG_HIST steps and the last solution Q = P + pshift
// that minimises |g - A B|_2 (FP64 normal equations, Cholesky with
// vectors that add nothing new dropped).  B goes to the guess rows.
struct WarmSh {
    double red[16 * ((MG_HIST + 1) * (MG_HIST + 2) / 2 + MG_HIST + 1)];
    double coef[MG_HIST + 1];
};

// Least squares over the vectors 0..m-1 from their normal equations G c = h:
// Cholesky in vector order, dropping a vector whose pivot falls below
// 1e-12 of its diagonal (nearly dependent on the ones kept before it); the
// dropped ones get c = 0.  Every index is a compile-time constant (the
// kept set is a mask), so G, L and the solves stay in registers.
template <int MV>
DEV void normal_solve(const double (&G)[MV][MV], const double (&h)[MV], int m, double (&c)[MV])
{
    double L[MV][MV];
    bool keep[MV];
#pragma unroll
    for (int j = 0; j < MV; j++) {
#pragma unroll
        for (int i = 0; i < j; i++) {
            double v = G[j][i];
#pragma unroll
            for (int u = 0; u < i; u++) v = (keep[u] && keep[i]) ? v - L[j][u] * L[i][u] : v;
            L[j][i] = keep[i] ? v / L[i][i] : 0.0;
        }
        double d = G[j][j];
#pragma unroll
        for (int i = 0; i < j; i++) d = keep[i] ? d - L[j][i] * L[j][i] : d;
        keep[j] = j < m && G[j][j] > 0.0 && d > 1.0e-12 * G[j][j];
        L[j][j] = keep[j] ? sqrt(d) : 1.0;
    }
    double y[MV];
#pragma unroll
    for (int i = 0; i < MV; i++) {
        double v = h[i];
#pragma unroll
        for (int u = 0; u < i; u++) v = (keep[u] && keep[i]) ? v - L[i][u] * y[u] : v;
        y[i] = keep[i] ? v / L[i][i] : 0.0;
    }
#pragma unroll
    for (int i = MV - 1; i >= 0; i--) {
        double v = y[i];
#pragma unroll
        for (int u = i + 1; u < MV; u++) v = (keep[u] && keep[i]) ? v - L[u][i] * c[u] : v;
        c[i] = keep[i] ? v / L[i][i] : 0.0;
    }
}

DEV void body_warm_base(const BatchDev &bd, int inst, WarmSh &sh)
{
    InstState &S = bd.st[inst];
    if (S.status || !bd.hist || !bd.guess) {
        // a stopped case gets no network correction
        SERIAL {
            if (bd.wscale) {
                bd.wscale[LICA_WARM_INFO * inst + 1] = 0.0;
                bd.wscale[LICA_WARM_INFO * inst + 6] = 0.0;
            }
        }
        SERIAL_END
        return;
    }
    CaseP cp = case_ptrs(bd, inst);
    const InstParam &P = *cp.P;
    const int n = P.n, s = bd.msub - 1;
    const double *__restrict__ PA = cp.A(NS_PA), *__restrict__ PCc = cp.A(NS_PC);
    // PB = -(PA + PC) row by row (the momentum stage's POISSON matrix),
    // evaluated instead of read
    auto pbv = [&](int row) { return -(PA[row] + PCc[row]); };
    const double *__restrict__ PG = cp.A(NS_PG), *__restrict__ PP = cp.A(NS_P);
    double *B = bd.guess + P.gs_off;
    const double *Hs = bd.hist + P.hist_off + (size_t)s * MG_HIST * (size_t)n;
    const int cnt = S.hist_cnt[s], head = S.hist_head[s];
    const double pq = S.pshift;
    const int m = cnt + 1;   // history vectors (most recent first), then Q
    auto val = [&](int j, int row) -> double {
        if (row < 1 || row > n) return 0.0;
        if (j < cnt) return Hs[(size_t)((head - 1 - j + 2 * MG_HIST) % MG_HIST) * n + (row - 1)];
        return PP[row] + pq;
    };
    auto aval = [&](int j, int row) -> double {   // (A v_j)(row), rows 1..n of the POISSON operator
        double v = pbv(row) * val(j, row);
        if (row > 1) v = v + PA[row] * val(j, row - 1);
        if (row < n) v = v + PCc[row] * val(j, row + 1);
        return v;
    };
    // G_ij = <A v_i, A v_j> (i <= j, row by row), h_i = <A v_i, g>
    constexpr int MV = MG_HIST + 1, NG = MV * (MV + 1) / 2;
    double sums[NG + MV];
    red_sumN<NG + MV>(1, n + 1, [&](int row, double *v) {
        double w[MV];
        for (int j = 0; j < MV; j++) w[j] = (j < m) ? aval(j, row) : 0.0;
        const double g = PG[row];
        int q = 0;
        for (int i = 0; i < MV; i++)
            for (int j = i; j < MV; j++) v[q++] = w[i] * w[j];
        for (int j = 0; j < MV; j++) v[NG + j] = w[j] * g;
    }, sh.red, sums);
    SERIAL {
        double G[MV][MV], h[MV];
        {
            int q = 0;
            for (int i = 0; i < MV; i++)
                for (int j = i; j < MV; j++) { G[i][j] = sums[q]; G[j][i] = sums[q]; q++; }
            for (int j = 0; j < MV; j++) h[j] = sums[NG + j];
        }
        double c[MV];
        normal_solve<MV>(G, h, m, c);
        for (int j = 0; j < MV; j++) sh.coef[j] = c[j];
    }
    SERIAL_END
    auto base = [&](int row) -> double {
        double v = 0.0;
        for (int j = 0; j < m; j++) v = v + sh.coef[j] * val(j, row);
        return v;
    };
    if (!bd.wscale) {
        PFOR(row, 1, n + 1) { B[row] = base(row); }
        PFOR_END
        return;
    }
    // B, then (reading B back, its neighbours written by other threads
    // before the barrier) the scales of the network's features -- max|r| of
    // B's residual g - (bB + aB- + cB+), max|b| -- and the maxima of the
    // multigrid solve's start test -- max|g|, max(|a|+|b|+|c|), max|B| and
    // max|r| with r accumulated as that test does, ((g - bB) - aB-) - cB+ --
    // so that the solve need not read the rows again when B already meets
    // the tolerance.
    PFOR(row, 1, n + 1) { B[row] = base(row); }
    PFOR_END
    double mx[6];
    red_maxN<6>(1, n + 1, [&](int row, double *v) {
        const double x = B[row];
        const double xl = (row > 1) ? B[row - 1] : 0.0, xr = (row < n) ? B[row + 1] : 0.0;
        double ax = pbv(row) * x;
        if (row > 1) ax = ax + PA[row] * xl;
        if (row < n) ax = ax + PCc[row] * xr;
        double rm = PG[row] - pbv(row) * x;
        if (row > 1) rm = rm - PA[row] * xl;
        if (row < n) rm = rm - PCc[row] * xr;
        const double ak = (row > 1) ? fabs(PA[row]) : 0.0, ck = (row < n) ? fabs(PCc[row]) : 0.0;
        v[0] = fabs(PG[row] - ax);
        v[1] = fabs(pbv(row));
        v[2] = ak + fabs(pbv(row)) + ck;
        v[3] = fabs(x);
        v[4] = fabs(PG[row]);
        v[5] = fabs(rm);
    }, sh.red, mx);
    SERIAL {
        double *inf = bd.wscale + (size_t)LICA_WARM_INFO * inst;
        const double den = mx[2] * mx[3] + mx[4];
        inf[1] = (den > 0.0) ? mx[0] / den : mx[0];
        inf[2] = (mx[0] == 0.0) ? 1.0 : mx[0];
        inf[3] = (mx[1] == 0.0) ? 1.0 : mx[1];
        const double res0 = (den > 0.0) ? mx[5] / den : mx[5];
        inf[5] = res0;
        inf[6] = (!(mx[5] > mx[4]) && !(res0 > bd.mg_rtol)) ? 1.0 : 0.0;
    }
    SERIAL_END
}

// CONVRGE1 (1d.f:3306-3362): max |dw/dz - S (rho_l - rho_v)/rho_l| over the
// cells (S the cell's box average of SFGAS), and the CFL input
// max |W(k)+W(k+1)| SSDZ(k) of the next DETERMINE_DT (1d.f:1325-1373)
DEV void conv_maxima(const CaseP &cp, const InstState &S, BlkSh &sh, double *dvmax, double *c3max)
{
    const InstParam &P = *cp.P;
    const int n = P.n;
    const double *__restrict__ W = cp.A(NS_W);
    const double *__restrict__ Z = cp.A(NS_Z), *__restrict__ SSDZ = cp.A(NS_SSDZ), *__restrict__ ZF = cp.R(RF_ZF);
    const double *__restrict__ SFGAS = cp.R(RF_SFGAS);
    const int *__restrict__ KCOU1 = cp.AI(NSI_KCOU1), *__restrict__ KCOU2 = cp.AI(NSI_KCOU2);
    const int sfl = S.sf_lo, sfh = S.sf_hi;
    stage_refwin(*sh.win, sfl - 2 * P.r - 4, sfh + 2 * P.r + 4, P.m3l, P.n3f + P.r, ZF, SFGAS, nullptr);
    const RefWin &win = *sh.win;
    auto zfa = [&](int kk) { return (kk >= win.lo && kk <= win.hi + 1) ? win.zf[kk - win.lo] : ZF[kk]; };
    auto fsf = [&](int kk) { return (kk >= win.lo && kk <= win.hi) ? win.f0[kk - win.lo] : SFGAS[kk]; };
    *dvmax = red_max(1, n + 1, 0.0, [&](int k) {
        double d = (W[k + 1] - W[k]) * SSDZ[k];
        bool t = !(KCOU1[k + 1] < sfl || KCOU2[k] > sfh);
        double ss = t ? box_avg_acc(zfa, P.ssdzf, fsf, Z[k], Z[k + 1], KCOU2[k], KCOU1[k + 1]) : 0.0;
        d = d - ss * (P.den_diff / P.denp);
        return fabs(d);
    }, sh.red);
    *c3max = red_max(1, n + 1, 0.0, [&](int k) { return fabs(W[k] + W[k + 1]) * SSDZ[k]; }, sh.red);
}

// CONVRGE1 on its own (the seam with the reference's argument list, after
// the caller's W was copied in)
DEV void body_convrge1(const BatchDev &bd, int inst, BlkSh &sh)
{
    InstState &S = bd.st[inst];
    if (S.status) return;
    CaseP cp = case_ptrs(bd, inst);
    double dvmax, c3max;
    conv_maxima(cp, S, sh, &dvmax, &c3max);
    SERIAL {
        S.dvmax = dvmax;
        S.c3max = c3max;
    }
    SERIAL_END
}

// UCALC (1d.f:3582-3627), p -= p(1) (3433-3470), CONVRGE1 (3306-3362), the
// CFL input of the next DETERMINE_DT, and VEL_INTP's UT/UO (two_phase.f:5570-5616)
DEV void body_ucalc(const BatchDev &bd, int inst, BlkSh &sh)
{
    InstState &S = bd.st[inst];
    if (S.status) return;
    CaseP cp = case_ptrs(bd, inst);
    const InstParam &P = *cp.P;
    const int n = P.n, n3 = P.n3, msub = bd.msub;
    double *__restrict__ W = cp.A(NS_W), *__restrict__ PP = cp.A(NS_P), *__restrict__ PZ = cp.A(NS_PSI_Z), *__restrict__ WO = cp.A(NS_WO), *__restrict__ WT = cp.A(NS_WT);
    const double *__restrict__ T = cp.A(NS_T);
    const double *__restrict__ VVDZ = cp.A(NS_VVDZ);
    const double *__restrict__ Z = cp.A(NS_Z), *__restrict__ SSDZ = cp.A(NS_SSDZ), *__restrict__ ZF = cp.R(RF_ZF);
    const double *__restrict__ SFGAS = cp.R(RF_SFGAS);
    const int *__restrict__ KCOU1 = cp.AI(NSI_KCOU1), *__restrict__ KCOU2 = cp.AI(NSI_KCOU2);
    const double dtconst = S.dtconst;
    const long long m = S.m;
    const int sfl = S.sf_lo, sfh = S.sf_hi;
    // p(1) before the shift (nothing below writes it before the shift pass)
    const double phiref = PP[1];
    stage_refwin(*sh.win, sfl - 2 * P.r - 4, sfh + 2 * P.r + 4, P.m3l, P.n3f + P.r, ZF, SFGAS, nullptr);
    const RefWin &win = *sh.win;
    auto zfa = [&](int kk) { return (kk >= win.lo && kk <= win.hi + 1) ? win.zf[kk - win.lo] : ZF[kk]; };
    auto fsf = [&](int kk) { return (kk >= win.lo && kk <= win.hi) ? win.f0[kk - win.lo] : SFGAS[kk]; };
    // UCALC with the end values W(1) = 0, W(n3) = W(n) (from the thread of
    // cell n, which has the corrected W(n))
    PFOR(k, 2, n + 1) {
        const double wk = W[k] - dtconst * (PP[k] - PP[k - 1]) * VVDZ[k] * denf_zi(P, PZ, k);
        W[k] = wk;
        if (k == n) { W[1] = 0.0; W[n3] = wk; }
    }
    PFOR_END
    if (n < 2) {
        SERIAL { W[1] = 0.0; W[n3] = W[n]; }
        SERIAL_END
    }
    // one pass over k = 0..n3 (W no longer changes): p -= p(1) (1d.f:3433-3470,
    // p(0) = p(1) after it), CONVRGE1's DVMAX and the CFL input (as
    // conv_maxima), MAIN's end-of-step finiteness check (the last substep;
    // W, T and P no longer change after it, and it takes effect after that
    // substep's TRANSPORT, whose own STOPs come first) and VEL_INTP's
    // UT = (U + UO)/2, UO = U (two_phase.f:5570-5616)
    const int first = (m == 1 && msub == 1);
    double mx[3];
    red_maxN<3>(0, n3 + 1, [&](int k, double *v) {
        double pk = PP[k];
        if (k >= 1 && k <= n) { pk = pk - phiref; PP[k] = pk; }
        else if (k == 0) { pk = phiref - phiref; PP[0] = pk; }
        v[0] = 0.0;
        v[1] = 0.0;
        const double wk = W[k];
        if (k >= 1 && k <= n) {
            double d = (W[k + 1] - wk) * SSDZ[k];
            bool t = !(KCOU1[k + 1] < sfl || KCOU2[k] > sfh);
            double ss = t ? box_avg_acc(zfa, P.ssdzf, fsf, Z[k], Z[k + 1], KCOU2[k], KCOU1[k + 1]) : 0.0;
            d = d - ss * (P.den_diff / P.denp);
            v[0] = fabs(d);
            v[1] = fabs(wk + W[k + 1]) * SSDZ[k];
        }
        v[2] = (msub == 3 && !(isfinite(wk) && isfinite(T[k]) && isfinite(pk))) ? 1.0 : 0.0;
        const double wo = first ? wk : WO[k];
        WT[k] = 0.5 * (wk + wo);
        WO[k] = wk;
    }, sh.red, mx);
    SERIAL {
        S.pshift = phiref;
        S.dvmax = mx[0];
        S.c3max = mx[1];
        S.nf_pending = (mx[2] > 0.0) ? 1 : 0;
    }
    SERIAL_END
}

// ---------------------------------------------------------------------
// initial state: MAKEFLD, LVSINIT (ITRACKING=2), GRID_COUPLING(1)
struct InitSh {
    BlkSh blk;
    int mask_s[WMAX];
    int bn[LICA_MLAYER + 1];
    int bk[LICA_MLAYER + 1][LICA_MAXP];
    int iv[16];
    double v[16];
    int inband[WMAX];
    double pm[LICA_MAXENT];
    int ek[LICA_MAXENT];
};

DEV double phi0(const InstParam &P, const double *ZPF, int k) { return (ZPF[k] - P.z0) + 1.0e-12; }

DEV void body_init(const BatchDev &bd, int inst, InitSh &is)
{
    InstState &S = bd.st[inst];
    CaseP cp = case_ptrs(bd, inst);
    const InstParam &P = *cp.P;
    const int n = P.n, n3 = P.n3, n3f = P.n3f, n3fm = P.n3fm;
    double *W = cp.A(NS_W), *T = cp.A(NS_T), *PZN = cp.A(NS_PSI_ZN), *PCN = cp.A(NS_PSI_CN);
    double *PSIF = cp.R(RF_PSIF);
    const double *ZP = cp.A(NS_ZP), *Z = cp.A(NS_Z), *SSDZ = cp.A(NS_SSDZ), *ZF = cp.R(RF_ZF), *ZPF = cp.R(RF_ZPF);
    const int *KCOU1 = cp.AI(NSI_KCOU1), *KCOU2 = cp.AI(NSI_KCOU2);
    const int *KCOUMP1 = cp.AI(NSI_KCOUMP1), *KCOUMP2 = cp.AI(NSI_KCOUMP2);
    // MAKEFLD (1d.f:879-917)
    PFOR(k, 1, n + 1) {
        double pbody = (ZP[k] - P.z0) + 1.0e-12;
        if (pbody > 0.0) { T[k] = 1.0; W[k] = 1.0; }
        else { T[k] = P.tslope * ZP[k]; W[k] = 0.0; }
    }
    PFOR_END
    int any_liq = red_or(1, n + 1, [&](int k) { return ((ZP[k] - P.z0) + 1.0e-12) > 0.0; }, is.blk.red);
    int any_vap = red_or(1, n + 1, [&](int k) { return !(((ZP[k] - P.z0) + 1.0e-12) > 0.0); }, is.blk.red);
    SERIAL {
        if (any_liq) W[n3] = 1.0;
        if (any_vap) { T[0] = 0.0; W[1] = 0.0; }
        S.dt = P.dt0; S.time = 0.0; S.m = 0; S.status = 0;
        S.n_bandgen = 1.0;   // the seed BAND_GENERATION of LVSINIT
        S.sf_lo = 1; S.sf_hi = 0; S.im_lo = 1; S.im_hi = 0;
        // LVSINIT seed (two_phase.f:209-252): a contiguous range of planes
        double cri = 1.2 * fmax(fmax(P.sdxf, P.sdyf), P.sdzf);
        int slo = -1, shi = -2;
        for (int k = 1; k <= n3fm; k++) {
            if (fabs(phi0(P, ZPF, k)) <= cri) { if (slo < 0) slo = k; shi = k; }
        }
        long long nseed = (slo < 0) ? 0 : (shi - slo + 1);
        long long rr = (long long)P.r * P.r, total = nseed * rr, crin = total / (LICA_MLAYER - 1);
        long long keep = (crin * (LICA_MLAYER - 1) + rr - 1) / rr;
        if (keep > nseed) keep = nseed;
        int scan_hi = slo + (int)keep - 1;
        auto seedphi = [&](int k) { return (slo >= 0 && k >= slo && k <= shi) ? phi0(P, ZPF, k) : 0.0; };
        // S-tube over the scanned seed planes (two_phase.f:1667-1694)
        for (int nl = 0; nl <= LICA_MLAYER; nl++) is.bn[nl] = 0;
        int st = 0;
        for (int k = slo; slo >= 0 && k <= scan_hi; k++) {
            double a = seedphi(k);
            if (a * seedphi(k + 1) < 0.0 || a * seedphi(k - 1) < 0.0) {
                if (is.bn[1] >= LICA_MAXP) { st = ST_BAND_CAPACITY; break; }
                is.bk[1][is.bn[1]++] = k;
            }
        }
        int klo = 1 << 30, khi = -1;
        for (int q = 0; q < is.bn[1]; q++) { if (is.bk[1][q] < klo) klo = is.bk[1][q]; if (is.bk[1][q] > khi) khi = is.bk[1][q]; }
        if (khi < 0) { klo = 1; khi = 1; }
        int wlo = klo - 14 < -4 ? -4 : klo - 14;
        is.iv[0] = wlo;
        for (int i = 0; i < WMAX; i++) { is.mask_s[i] = 0; is.inband[i] = 0; }
        auto msk = [&](int k) -> int & { return is.mask_s[k - wlo]; };
        for (int k = (slo > wlo ? slo : wlo); slo >= 0 && k <= scan_hi && k < wlo + WMAX; k++) msk(k) = 2;
        for (int q = 0; q < is.bn[1]; q++) msk(is.bk[1][q]) = 1;
        // layers 2..N_MAX (two_phase.f:1700-1960), planes only
        for (int nl = 2; nl <= LICA_MLAYER && !st; nl++)
            for (int q = 0; q < is.bn[nl - 1] && !st; q++) {
                int k = is.bk[nl - 1][q];
                for (int k2 = k - 1; k2 <= k + 1; k2++) {
                    if (k2 == 1 || k2 == n3f) continue;
                    if (k2 < wlo || k2 >= wlo + WMAX) { st = ST_WINDOW; break; }
                    if (msk(k2) != 1) {
                        if (is.bn[nl] >= LICA_MAXP) { st = ST_BAND_CAPACITY; break; }
                        is.bk[nl][is.bn[nl]++] = k2;
                        msk(k2) = 1;
                    }
                }
            }
        for (int nl = 1; nl <= LICA_MLAYER; nl++)
            if ((long long)is.bn[nl] * P.r * P.r > P.mf_band) st = ST_BAND_OVERFLOW;
        int e = 0;
        for (int nl = 1; nl <= LICA_MLAYER; nl++) {
            S.bn[nl] = is.bn[nl];
            for (int q = 0; q < is.bn[nl]; q++) {
                int k = is.bk[nl][q];
                S.bk[nl][q] = k;
                S.alphi[nl][q] = phi0(P, ZPF, k);
                is.inband[k - wlo] = 1;
                is.ek[e++] = k;
            }
        }
        is.iv[1] = e;
        is.iv[2] = slo; is.iv[3] = shi;
        S.status = st;
    }
    SERIAL_END
    if (S.status) return;
    const int wlo = is.iv[0], ne = is.iv[1], slo = is.iv[2], shi = is.iv[3];
    auto phinit = [&](int k) {
        bool seed = (slo >= 0 && k >= slo && k <= shi);
        bool band = (k >= wlo && k < wlo + WMAX) ? (is.inband[k - wlo] != 0) : false;
        return (seed || band) ? phi0(P, ZPF, k) : 0.0;
    };
    // CAL_PSIF(1,N_MAX): band formula on all layers, column fill elsewhere
    PFOR(e, 0, ne) {
        int k = is.ek[e];
        double d3 = 0.5 * fabs(phinit(k + 1) - phinit(k - 1));
        double v = d_psif_cell(phinit(k), d3);
        is.pm[e] = snap5(v);
        if (v == -1.0) S.status = ST_PSIF_BUG;
    }
    PFOR_END
    SERIAL {
        int knum = 0, khigh = 0, klow = 0;
        int kt = shi < n3fm ? shi : n3fm;
        int kl = slo < 1 ? 1 : slo;
        for (int k = kt; slo >= 0 && k >= kl; k--)
            if (phinit(k) * phinit(k + 1) < 0.0) { if (knum == 0) khigh = k; knum++; klow = k; }
        // band planes outside the seed range
        for (int e = 0; e < ne; e++) {
            int k = is.ek[e];
            if (k > kt || k < kl)
                if (phinit(k) * phinit(k + 1) < 0.0) { knum++; if (k > khigh) khigh = k; if (k < klow || klow == 0) klow = k; }
        }
        is.iv[4] = knum; is.iv[5] = klow;
        is.v[0] = (knum > 0) ? ((phinit(khigh) < 0.0) ? 0.0 : 1.0) : 1.0;
        is.v[1] = (knum > 0) ? ((phinit(klow + 1) < 0.0) ? 0.0 : 1.0) : 1.0;
    }
    SERIAL_END
    const int knum = is.iv[4], klow = is.iv[5];
    const double tmp1 = is.v[0], tmpo = is.v[1];
    PFOR(k, 1, n3fm + 1) {
        double v = (knum == 0) ? 1.0 : ((k <= klow) ? tmp1 : tmpo);
        PSIF[k] = v;
    }
    PFOR_END
    PFOR(e, 0, ne) { PSIF[is.ek[e]] = is.pm[e]; }
    PFOR_END
    SERIAL {
        for (int k = 1; k <= n3fm; k++) PSIF[k] = snap5(PSIF[k]);
        for (int k = 0; k <= P.r; k++) { PSIF[-k] = PSIF[1]; PSIF[n3f + k] = PSIF[n3fm]; }
        // VOL_TOT_ORI (two_phase.f:423-433)
        double vol = 0.0, dvol = P.sdzf;
        for (int k = 1; k <= n3fm; k++) vol = vol + (1.0 - PSIF[k]) * dvol;
        S.vol_tot_ori = vol;
        S.vol_tot_iphs = vol;
    }
    SERIAL_END
    // GRID_COUPLING(1)
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
    double c3max = red_max(1, n + 1, 0.0, [&](int k) { return fabs(W[k] + W[k + 1]) * SSDZ[k]; }, is.blk.red);
    SERIAL { S.c3max = c3max; }
    SERIAL_END
}
