// Device/host shared layout of the batched film-boiling state in HBM.
//
// Every case (instance) owns three contiguous blocks:
//   NS block      NS_NF fields x ln doubles, field f, NS index k in [-2, n+5]
//                 at ns[f*ln + k + 2]
//   refined block RF_NF fields x lf doubles, refined index k in [m3l, ...]
//                 at rf[f*lf + k - m3l]          (m3l = -2r)
//   int blocks    NSI_NF x lni ints (k at [f*lni + k + 2]),
//                 RFI_NF x lfi ints (k at [f*lfi + k])
// Fields of one case are structure-of-arrays so that a workgroup walking k
// reads consecutive addresses.  The narrow band (<= 11 layers of z-planes)
// and all scalars live in a fixed-size InstState record per case.
#pragma once
#include <stdint.h>

#define LICA_MLAYER 11
#define LICA_NN1 8
#define LICA_NN2 10
#define LICA_MAXP 16
#define LICA_MAXENT (LICA_MLAYER * LICA_MAXP)

enum NsField {
    NS_Z = 0, NS_ZP, NS_SDZ, NS_VDZ, NS_SSDZ, NS_VVDZ, NS_AKW, NS_CKW, NS_AKUV, NS_CKUV,
    NS_W, NS_P, NS_T, NS_PSI_ZN, NS_PSI_CN, NS_WO, NS_RK3ZO, NS_WT,
    NS_DF_Z, NS_VF_Z, NS_DF_C, NS_PSI_Z, NS_PSI_C, NS_RHSPS,
    NS_TA, NS_TB, NS_TC, NS_TG,           // tridiagonal work (energy, momentum)
    NS_PA, NS_PB, NS_PC, NS_PG,           // Poisson system of the substep
    NS_NF
};
enum NsiField { NSI_KCOU1 = 0, NSI_KCOU2, NSI_KCOUMP1, NSI_KCOUMP2, NSI_IMASK, NSI_NF };
enum RfField { RF_ZF = 0, RF_ZPF, RF_PSIF, RF_SFGAS, RF_SFGAS2, RF_VPREFIX, RF_NF };
enum RfiField { RFI_KCOU_VEL = 0, RFI_KCOUMP_VEL, RFI_NF };

struct InstParam {
    int n, r, n3, n3f, n3fm, m3l;
    int ln, lf, lni, lfi;
    int mf_band, idtopt;
    long long ns_off, rf_off, nsi_off, rfi_off;
    long long mg_off;            // multigrid scratch (doubles), multigrid modes only
    long long gs_off;            // guess / capture rows (n+2 per case)
    long long hist_off;          // solution history of the warm start (3 substeps x MG_HIST x n)
    // properties (INPUT_SIMULATION 1d.f:434-480)
    double re, prm, denr, ja, tsat, cflmax, dt0, z0, tslope;
    double rei, denm, denp, den_diff, vism, visp, denscm, densc_diff;
    double tcm, tcp, tc_diff, prai, time_st, time_gv;
    double gamma[3], ro[3];
    // geometry scalars
    double sdzf, ssdzf, sdxf, sdyf, sdx1, sdy1, xl, yl, xp1, yp1;
    double phins_xd, phins_yd;   // PHINSINTP x/y weights
    double dxy2_min;             // min over the r*r columns of DX^2+DY^2 (chphase.f:198-211)
};

struct InstState {
    double time, dt, alpha, acoef, dtconst, dtconsti;
    long long m;
    int msub, status;
    int nf_pending;       // W, T or P non-finite after the step's last UCALC: ST_NONFINITE once TRANSPORT ends
    int pad0;
    double vol_tot_ori, vol_tot_iphs, place, dvmax, cflm, c3max;
    double n_subcycle_warn, n_mass_warn, n_halvings, n_extrap_iters;
    double n_reinit, n_bandgen, n_masscor;
    double poisson_cycles, poisson_solves, poisson_resid;
    double pshift;        // p(1) removed by the last UCALC: P + pshift is the last solution
    int hist_head[3], hist_cnt[3];   // warm-start history ring per substep: next slot, entries held
    int sf_lo, sf_hi;     // refined window of nonzero SFGAS/SFGAS2
    int im_lo, im_hi;     // NS window of IMASK_PH==1
    int bn[LICA_MLAYER + 1];
    int bk[LICA_MLAYER + 1][LICA_MAXP];
    double alphi[LICA_MLAYER + 1][LICA_MAXP];
    double mflvs[LICA_MLAYER + 1][LICA_MAXP];
};

// kernel argument block: the whole batch, one workgroup per case
struct BatchDev {
    const InstParam *par;
    InstState *st;
    double *ns, *rf;
    int *nsi, *rfi;
    const int *order;    // workgroup -> case (largest cases first)
    int ninst;
    int tri_mode;        // 0: serial Thomas, 1: partitioned + refinement, 2: partitioned
    int msub;            // RK3 substep 1..3
    int pmode;           // Poisson mode (0/1 direct, 2 MG64, 3 mixed, 4 learned)
    int mg_maxc;         // V-cycle cap per solve
    double mg_rtol;      // relative residual target
    double *mg;          // multigrid scratch of the batch
    long long mg_lds;    // dynamic LDS bytes of the multigrid kernel (x, f of the coarse levels)
    double *guess;       // initial guesses of the warm-started modes (rows at gs_off)
    double *hist;        // solution history of the warm-started modes
    double *cap;         // P before / after the last substep's solve, its RHS (3 x (n+2) at 3*gs_off)
    double *wscale;      // per case LICA_WARM_INFO doubles (learned mode): network window start,
                         // backward error of the warm-start base, its residual scale s_r, the row scale
                         // s_b, the network step alpha, the base's backward error as the multigrid start
                         // test computes it, 1 if the solve may take the guess as it is
    unsigned long long *lead;   // TEMP_GRAD's lead slots of this launch's case set (4 x LICA_TG_LEAD
                                // keys, by substep), or null: no lead workgroups
};
#define LICA_WARM_INFO 7
// TEMP_GRAD's lead workgroups: a launch with lead slots runs LICA_TG_LEAD
// workgroups ahead of its cases, each taking the costliest case (by the
// previous substep's extrapolation iterations) of one residue class of the
// positions
#define LICA_TG_LEAD 32

#define LICA_MG_FP64_MODE 2
#define LICA_MG_MIXED_MODE 3
#define LICA_MG_LEARNED_MODE 4
#define LICA_MG_PROJ_MODE 5
#ifndef MG_HIST
#define MG_HIST 4            // solutions per substep kept for the warm start
#endif
// modes whose V-cycles start from the warm-start guess
#define LICA_MG_WARM(mode) ((mode) == LICA_MG_LEARNED_MODE || (mode) == LICA_MG_PROJ_MODE)

// status codes shared with include/lica_mi355x.h
#define ST_OK 0
#define ST_BAND_OVERFLOW 1
#define ST_PSIF_BUG 2
#define ST_GRID 3
#define ST_EXTRAP 4
#define ST_BAND_CAPACITY 5
#define ST_NONFINITE 6
#define ST_WINDOW 7
#define ST_NOT_CONVERGED 32
