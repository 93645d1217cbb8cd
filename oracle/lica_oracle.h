/*
 * lica_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * Serial FP64 CPU restatement of the per-step path of the 1-D film-boiling
 * solver (reference files 1d.f, two_phase.f, chphase.f, grid.f).  It is the
 * parity checker for the HIP path and the CPU baseline of bench.py; nothing
 * in lica_boiling1d_amd/ may link, load or call it.
 *
 * The reference stores the refined level-set grid as (r x r x rN) with x/y
 * columns that stay identical; this restatement keeps one z column and keeps
 * the narrow-band lists as ordered lists of z-planes (each plane standing for
 * the r*r list entries the reference holds for it).
 */
#ifndef LICA_ORACLE_H
#define LICA_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int64_t n;          /* NS cells in z (N3M)                       */
    int64_t r;          /* level-set refinement factor (N3F_BD)      */
    double zl;          /* domain height (grid.in ZL)                */
    double re, visr, denr;            /* lica.in line 21           */
    double surf_j, fr;                /* lica.in line 23           */
    double prm, scr, tcr;             /* lica.in line 25           */
    double tsat, ja;                  /* lica.in line 27           */
    double dt0, cflmax;               /* lica.in line 6            */
    int64_t idtopt;
    double z0;          /* initial interface height (ASSIGN_LVS)     */
    double tslope;      /* initial vapour temperature slope (MAKEFLD) */
} orc_case;

typedef struct orc_inst orc_inst;

#define ORC_M_MAX 11    /* band layers (N_MAX)            */
#define ORC_MAXP 16     /* z-planes stored per band layer */

/* status codes */
enum {
    ORC_OK = 0,
    ORC_BAND_OVERFLOW = 1,       /* two_phase.f:361,1989 */
    ORC_PSIF_BUG = 2,            /* two_phase.f:2905 */
    ORC_GRID = 3,                /* 1d.f:4761 / coupling search failure */
    ORC_EXTRAP_NOT_CONVERGED = 4,/* chphase.f:676-678 (bounded here) */
    ORC_BAND_CAPACITY = 5,       /* more z-planes per layer than stored */
    ORC_NONFINITE = 6
};

/* field ids for orc_get */
enum {
    ORC_F_W = 0, ORC_F_P = 1, ORC_F_T = 2, ORC_F_PSI_CN = 3, ORC_F_PSI_ZN = 4,
    ORC_F_PSIF = 5, ORC_F_PHI = 6, ORC_F_SFGAS = 7, ORC_F_WO = 8,
    ORC_F_BAND = 9,   /* per layer: count then planes (int as double) */
    ORC_F_ALPHI = 10, ORC_F_MFLVS = 11
};

/* scalar ids for orc_scalars (array of ORC_NSCAL doubles) */
enum {
    ORC_S_TIME = 0, ORC_S_DT = 1, ORC_S_STEP = 2, ORC_S_PLACE = 3,
    ORC_S_DVMAX = 4, ORC_S_CFLM = 5, ORC_S_VOL_ORI = 6, ORC_S_VOL_IPHS = 7,
    ORC_S_QWALL = 8, ORC_S_STATUS = 9, ORC_S_SUBCYCLE_WARN = 10,
    ORC_S_MASS_WARN = 11, ORC_S_EXTRAP_HALVINGS = 12, ORC_S_EXTRAP_ITERS = 13,
    ORC_S_REINIT_CALLS = 14, ORC_S_BANDGEN_CALLS = 15, ORC_S_MASSCOR_ITERS = 16,
    ORC_NSCAL = 24
};

orc_inst *orc_new(const orc_case *c);
void orc_free(orc_inst *s);
int orc_init(orc_inst *s);
/* restart from a field file's contents (1d.f:117-143 with IREAD=1,
 * ITRACKING=1): w, p, t indexed 0..N3; bn[1..M_MAX]; bk and alphi rows
 * nl = 0..M_MAX of ORC_MAXP slots; call on a fresh instance              */
int orc_init_fld(orc_inst *s, const double *w, const double *p, const double *t, double time, double dt,
                 double vol_tot_ori, const int64_t *bn, const int64_t *bk, const double *alphi);
int orc_step(orc_inst *s, int64_t nsteps);
int64_t orc_get(orc_inst *s, int field, double *out, int64_t cap);
void orc_scalars(orc_inst *s, double *out);
int64_t orc_n(orc_inst *s);
/* the Poisson inputs (a, b, c, g of POISSON) of the last substep */
int64_t orc_last_poisson(orc_inst *s, double *a, double *b, double *c, double *g, double *p);

/* CPU baseline: nc cases set up and advanced nsteps steps each (or until
 * they stop), one case per OpenMP thread; set-up and steps timed apart
 * (seconds, wall clock); returns the threads used */
/* share (on = 1) or not the per-cell loops of one instance among the OpenMP threads */
void orc_set_loop_threads(orc_inst *s, int on);
/* one case with its per-cell loops shared by `threads` OpenMP threads
 * (0: the default count); returns the threads used                        */
int orc_run_single(const orc_case *c, int64_t nsteps, int64_t threads, int64_t *status, int64_t *steps_out,
                   double *setup_s, double *step_s);
int orc_run_cases(const orc_case *cases, int64_t nc, int64_t nsteps, int64_t *status, int64_t *steps_out,
                  double *setup_s, double *step_s);

/* unit-level hooks of the pinning tests (each runs one restated routine on
 * an initialised instance with inputs the test prescribes; see the .c file) */
void orc_rk3_coefficients(const orc_inst *s, double *gamma, double *ro);
double orc_test_box_avg(orc_inst *s, const double *f, double z1, double z2, int64_t k1, int64_t k2);
void orc_test_determine_dt(orc_inst *s, const double *w, const double *psi_cn, double dt, double *dt_out,
                           double *cflm_out);
int orc_test_mass_correction(orc_inst *s, double vol_target, int64_t *passes);
void orc_test_vel_intp(orc_inst *s, const double *w, double *wf);
int orc_test_temp_grad(orc_inst *s, const double *t, double dtconst, double *temp0, double *temp1, double *mflvs,
                       double *sfgas2, double *place, double *vol_iphs);

int64_t orc_last_poisson_inputs(orc_inst *s, double *denf_zi, double *divgsum);

/* unit-level restatements used by the tests */
void orc_trdiag3(int64_t n, const double *a, const double *b, const double *c,
                 double *g);
double orc_hjweno(double a, double b, double c, double d);
double orc_godunov(double asign, double p, double m);
double orc_psif_cell(double phi, double d3);
void orc_grid_nodes(int64_t n, double zl, double *z);

/* The line solvers of the 3-D code (no live caller on the 1-D column).
 * Arrays are column-major m1 x m2 (element (i, j), 1-based, at
 * (i-1) + (j-1)*m1), as the reference's A(M2,M1) / A(M1,M2) dummies.
 * TRDIAG1/2/3 (1d.f:3721-3812): systems along j = l1..l2 for lines
 * i = ll1..ll2; overwrite_r (TRDIAG3) also stores UU in R(i, l1..l2-1). */
void orc_trdiag_lines(const double *a, const double *b, const double *c, double *r, double *uu, int64_t m1,
                      int64_t m2, int64_t l1, int64_t l2, int64_t ll1, int64_t ll2, int overwrite_r);
/* TRDIAG1P/2P/3P (1d.f:3815-3985): periodic systems along j = j1..j2 for
 * lines k = l1..l2, F in and out */
void orc_trdiag_periodic_lines(const double *a, const double *b, const double *c, double *f, int64_t m1, int64_t m2,
                               int64_t j1, int64_t j2, int64_t l1, int64_t l2);
/* PTDIAG1A (1d.f:3988-4034): pentadiagonal systems along j = j1..jn
 * (E u(j-2) + D u(j-1) + A u(j) + B u(j+1) + C u(j+2) = F) for lines
 * i = i1..in; O, Q, R receive the elimination's coefficients */
void orc_ptdiag_lines(const double *a, const double *b, const double *c, const double *d, const double *e,
                      const double *f, double *uu, int64_t m1, int64_t m2, int64_t i1, int64_t in, int64_t j1,
                      int64_t jn, double *o, double *q, double *r);

#ifdef __cplusplus
}
#endif
#endif
