/*
 * lica_mi355x.h -- C ABI of the MI355X (gfx950) implementation of the
 * per-step hot path of the 1-D film-boiling solver lica_boiling1d.
 *
 * Plain C types only: int64_t/double scalars and pointers.  Arrays passed
 * across the boundary are the reference's column arrays for the single
 * (I=1,J=1) column, index 0..N3 (N3 = n+1), i.e. Fortran X(1,1,0:N3).
 *
 * Reference interfaces replaced (file:line in jaehee016/lica_boiling1d):
 *   lica_parse_deck      INPUT_SIMULATION 1d.f:323-492 + grid.in/grid.f:10-322
 *   lica_init            MAIN init 1d.f:70-190 (ALLOCINIT, GEOM, MAKEFLD,
 *                        LVSINIT two_phase.f:116, GRID_COUPLING(1) 1d.f:143)
 *   lica_step            MAIN time loop 1d.f:196-295
 *   lica_determine_dt    DETERMINE_DT(CFLM,U,V,W,PSI_CN) 1d.f:683 (caller 1d.f:201)
 *   lica_rhsnlhs         RHSNLHS(ITRACKING,U,V,W,P,PSI_XN..PSI_CN,RK3XO..ANPSO,
 *                        QVOL_ORI,VOL1) 1d.f:1444 (caller 1d.f:227)
 *   lica_convrge1        CONVRGE1(DVMAX,U,V,W) 1d.f:3306 (caller 1d.f:234)
 *   lica_transport       TRANSPORT(U,V,W,PSI_CN,VOL_TOT_ORI) two_phase.f:701
 *                        (caller 1d.f:242)
 *   lica_grid_coupling   GRID_COUPLING(INI_PSI,PSI_XN,PSI_YN,PSI_ZN,PSI_CN)
 *                        two_phase.f:4863 (callers 1d.f:143,244)
 *   lica_poisson         POISSON(P,DIVGSUM,DENF_XI,DENF_YI,DENF_ZI,DENF_Y,DENF_C)
 *                        1d.f:3492 (caller 1d.f:3428), batched
 *   lica_batch_*         batched driver of the same loop for many cases
 *   lica_batch_write_fld WRITEFIELD 1d.f:1121-1232 (the fld###### restart/post file)
 *   lica_batch_save/load no counterpart: the reference's fld restart keeps 12
 *                        digits and omits PSIF/TEMP/WO/MFLVS; these keep every
 *                        byte of the device state a later step reads
 * Errors are returned as status codes (the reference STOPs).
 */
#ifndef LICA_MI355X_H
#define LICA_MI355X_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* one boiling case: the deck values of lica.in/grid.in it depends on */
typedef struct {
    int64_t n;        /* NS cells in z (N3M)                         */
    int64_t r;        /* level-set refinement factor (N3FM/N3M)      */
    double zl;        /* domain height                              */
    double re, visr, denr;
    double surf_j, fr;
    double prm, scr, tcr;
    double tsat, ja;
    double dt0, cflmax;
    int64_t idtopt;
    double z0;        /* initial interface height (ASSIGN_LVS)      */
    double tslope;    /* initial vapour temperature slope (MAKEFLD) */
} lica_case_desc;

typedef struct lica_batch lica_batch;

/* status codes */
enum {
    LICA_OK = 0,
    LICA_BAND_OVERFLOW = 1,        /* two_phase.f:361,1989  */
    LICA_PSIF_BUG = 2,             /* two_phase.f:2905      */
    LICA_GRID_TOO_LARGE = 3,       /* 1d.f:4761,4775        */
    LICA_EXTRAP_NOT_CONVERGED = 4, /* chphase.f:676-678     */
    LICA_BAND_CAPACITY = 5,
    LICA_NONFINITE = 6,
    LICA_WINDOW = 7,               /* band outside the device window */
    LICA_BAD_ARG = 20,
    LICA_UNSUPPORTED_DECK = 21,
    LICA_IO = 22,
    LICA_TOO_LARGE = 23,           /* beyond the fixed workspace of lica_poisson_noalloc */
    LICA_HIP_ERROR = 30,
    LICA_NO_DEVICE = 31,
    LICA_NOT_CONVERGED = 32        /* Poisson iteration hit its cycle cap */
};

/* Poisson solver modes (lica_batch_set_poisson) */
enum {
    LICA_POISSON_THOMAS = 0,    /* serial Thomas per case, bit-identical order */
    LICA_POISSON_DIRECT = 1,    /* parallel partitioned tridiagonal (FP64)      */
    LICA_POISSON_MG64 = 2,      /* FP64 V-cycles from zero to tolerance          */
    LICA_POISSON_MIXED = 3,     /* FP32 V-cycles inside FP64 defect correction   */
    LICA_POISSON_LEARNED = 4,   /* MIXED from the projection + conv-net guess    */
    LICA_POISSON_PROJECTED = 5  /* MIXED from the history projection alone       */
};

/* field ids for lica_batch_get_field: NS fields have n+2 entries (0..N3),
 * refined fields r*n entries (1..N3FM)                                  */
enum {
    LICA_F_W = 0, LICA_F_P = 1, LICA_F_T = 2, LICA_F_PSI_CN = 3,
    LICA_F_PSI_ZN = 4, LICA_F_PSIF = 5, LICA_F_PHI = 6, LICA_F_SFGAS = 7,
    LICA_F_WO = 8, LICA_F_BAND = 9, LICA_F_ALPHI = 10, LICA_F_MFLVS = 11,
    /* rows 1..n of the last POISSON system (n entries each); PG holds the
     * solution after a direct solve, the captured PRHS its right-hand side */
    LICA_F_PA = 12, LICA_F_PB = 13, LICA_F_PC = 14, LICA_F_PG = 15,
    LICA_F_PPREV = 16, LICA_F_PSOL = 17, LICA_F_PRHS = 18
};

/* scalar ids for lica_batch_get_scalars (LICA_NSCAL doubles) */
enum {
    LICA_S_TIME = 0, LICA_S_DT = 1, LICA_S_STEP = 2, LICA_S_PLACE = 3,
    LICA_S_DVMAX = 4, LICA_S_CFLM = 5, LICA_S_VOL_ORI = 6, LICA_S_VOL_IPHS = 7,
    LICA_S_QWALL = 8, LICA_S_STATUS = 9, LICA_S_SUBCYCLE_WARN = 10,
    LICA_S_MASS_WARN = 11, LICA_S_EXTRAP_HALVINGS = 12, LICA_S_EXTRAP_ITERS = 13,
    LICA_S_REINIT_CALLS = 14, LICA_S_BANDGEN_CALLS = 15, LICA_S_MASSCOR_ITERS = 16,
    LICA_S_POISSON_CYCLES = 17, LICA_S_POISSON_SOLVES = 18, LICA_S_POISSON_RESID = 19,
    LICA_S_DTCONST = 20, LICA_S_PSHIFT = 21,
    LICA_NSCAL = 24
};

/* deck: reads lica.in (list-directed, 1d.f:347-386) and grid.in
 * (grid.f:29-53); only the 1-D film-boiling configuration is accepted  */
int lica_parse_deck(const char *lica_in_path, const char *grid_in_path, lica_case_desc *out);

/* ---- batched API ---------------------------------------------------- */
int lica_batch_create(int64_t ninst, const lica_case_desc *cases, lica_batch **out);
int lica_batch_init(lica_batch *b);                       /* 1d.f:70-190 */
int lica_batch_run(lica_batch *b, int64_t nsteps);         /* 1d.f:196-295 x nsteps */
int lica_batch_set_poisson(lica_batch *b, int64_t mode, double rtol, int64_t max_cycles);
int lica_batch_set_block(lica_batch *b, int64_t threads);  /* 64..1024, power of 2 */
int lica_batch_use_graph(lica_batch *b, int64_t on);
/* per-kernel GPU time (events between launches) for the next runs; no graph */
int lica_batch_set_timing(lica_batch *b, int64_t on);
int lica_batch_sync(lica_batch *b);
int lica_batch_get_field(lica_batch *b, int64_t field, int64_t inst, double *out, int64_t cap, int64_t *count);
int lica_batch_get_scalars(lica_batch *b, int64_t inst, double *out);
int lica_batch_status(lica_batch *b, int64_t *first_bad_inst);
/* per case: status word and completed time steps (arrays of lica_batch_size) */
int lica_batch_get_progress(lica_batch *b, int64_t *status_out, int64_t *steps_out);
int64_t lica_batch_size(lica_batch *b);
/* GPU time of the last lica_batch_run, per kernel class (ms), summed over
 * the launches (L per kernel and substep with L lanes); returns count */
int lica_batch_last_timings(lica_batch *b, double *ms_out, int64_t cap);
void lica_batch_destroy(lica_batch *b);

/* ---- checkpoint / restart -------------------------------------------- */
/* binary checkpoint of the device state a later step reads (fields, band,
 * scalars, counters, the Poisson mode, tolerance and cycle cap, and in the
 * warm-started modes the solution history): a restart continues bit for
 * bit; the learned mode's network is not kept (set it again after loading) */
int lica_batch_save(lica_batch *b, const char *path);
int lica_batch_load(const char *path, lica_batch **out);
/* WRITEFIELD text file (1d.f:1121-1232, read by post.f READFLD 238-401) of one case */
int lica_batch_write_fld(lica_batch *b, int64_t inst, const char *path, int64_t ihist);
/* Restart of case inst of an initialized batch from a field file, as MAIN
 * does with IREAD=1 and ITRACKING=1 (1d.f:117-143): PREFLD's reads
 * (1d.f:972-1097: W, P, T, the band and VOL_TOT_ORI of the 1-D column; U
 * and V must be zero, x/y planes uniform), then LVSINIT from the band read
 * (two_phase.f:146-189) and GRID_COUPLING(1) on the device.  flags: the
 * deck's LICA_FLD_IRESET (TIME = 0), LICA_FLD_IPZERO (P = 0),
 * LICA_FLD_IDTOLD (DT = the file's DT, else the deck's).  The step counter
 * starts again at 1, as MAIN's loop does; the case's Poisson history is
 * cleared.  LICA_IO for a file that is not such a field file of this
 * case's grid; otherwise the case's status (0, or a reference STOP such as
 * LICA_PSIF_BUG found while rebuilding the level set).                   */
#define LICA_FLD_IRESET 1
#define LICA_FLD_IPZERO 2
#define LICA_FLD_IDTOLD 4
int lica_batch_read_fld(lica_batch *b, int64_t inst, const char *path, int64_t flags);

/* ---- warm start of the Poisson solve ----------------------------------- */
/* No counterpart in the reference.  LICA_POISSON_PROJECTED starts the
 * mixed-precision V-cycles from B, the combination of the case's solutions at
 * the same substep in the last two steps and its last solution that
 * minimises |g - A B|_2.  LICA_POISSON_LEARNED adds a 1-D convolutional
 * network's correction (4 layers, 21 taps, 32 channels, BF16 weights, FP32
 * accumulation on MFMA): from per-cell features (r/max|r|, a/max|b|,
 * b/max|b|, c/max|b| with r = g - A B) over 256 rows around the interface
 * it gives a direction d, and the guess is B + alpha d with
 * alpha = <A d, r>/<A d, A d>; cases whose B meets the tolerance skip it.
 * The V-cycles always run to the tolerance.                                */
typedef struct lica_net lica_net;
int64_t lica_net_param_count(void);
int lica_net_create(uint64_t seed, lica_net **out);
void lica_net_destroy(lica_net *net);
/* parameters layer by layer: W[tap][cout][32], then bias[cout]; cout 32,32,32,16 */
int lica_net_get_params(lica_net *net, float *out);
int lica_net_set_params(lica_net *net, const float *in);
/* features [cells][4] of nseq sequences of lens[s] cells -> y[cells] */
int lica_net_forward(lica_net *net, const float *feat, const int64_t *lens, int64_t nseq, float *y);
/* loss = mean (y - target)^2 and its gradient in the parameter layout */
int lica_net_grad(lica_net *net, const float *feat, const float *target, const int64_t *lens, int64_t nseq,
                  float *grad, double *loss);
/* training windows kept on the GPU ([window][cell][4] features, [window][cell]
 * targets); set_data replaces them, add_pairs appends windows_per_case
 * windows of len cells from every healthy case of a batch run with capture
 * on in a multigrid mode (features of the captured system and starting
 * point B around the interface, target p - B scaled to unit maximum over
 * the window), cases shorter than len are skipped                         */
int lica_net_set_data(lica_net *net, const float *feat, const float *target, int64_t n, int64_t len);
int lica_net_add_pairs(lica_net *net, lica_batch *b, int64_t windows_per_case, int64_t len, uint64_t seed,
                       int64_t *added);
int lica_net_data_size(lica_net *net, int64_t *n, int64_t *len);
int lica_net_get_data(lica_net *net, float *feat, float *target);
int lica_net_train(lica_net *net, int64_t iters, int64_t batch, double lr, uint64_t seed, double *loss);
/* No counterpart in the reference (tuning): cap on the LDS bytes a multigrid
 * workgroup keeps its coarse levels in (-1: 60 KB for batches of >= 512
 * cases, else the device limit); levels beyond the cap use global memory */
int lica_batch_set_mg_lds(lica_batch *b, int64_t bytes);
/* No counterpart in the reference (scheduling): L = 2..8 runs the batch's
 * cases as L sets (every L-th case of the largest-first order) whose step
 * launch sequences go to L streams at once, so that one set's latency-bound
 * kernels share the CUs with another's bandwidth-bound ones; 1 runs one
 * sequence.  L is capped at the number of cases.  Default: 2 for batches of
 * 256 cases or more.  Results do not change (cases are independent).  Each
 * lane needs a hardware queue of its own (GPU_MAX_HW_QUEUES, 4 by default in
 * HIP: set it to at least L + 4 before HIP starts for more than 2 lanes)  */
int lica_batch_set_lanes(lica_batch *b, int64_t lanes);
/* No counterpart in the reference (scheduling): split 1 makes the lanes
 * contiguous size bins of the largest-first order holding equal work (n
 * plus 512 rows per case), each launched with threads and multigrid LDS
 * fitted to its own largest case; 0 (default) every L-th case.  Results do
 * not change.                                                              */
int lica_batch_set_lane_split(lica_batch *b, int64_t split);
/* No counterpart in the reference (tuning): coordinate descent on the
 * batch's step time.  For the kernel classes whose shape is free
 * (the multigrid or direct POISSON solve, the momentum and energy stages,
 * UCALC, DETERMINE_DT, in that order) and for each lane, tries 64..1024
 * threads per workgroup, then multigrid LDS sizes per lane, each trial
 * `steps` (1..16) steps of every lane at once as in a run; a candidate
 * within 0.5 % of the best or faster is timed again over 4 x `steps` steps
 * in turn with the kept shape, twice, and taken when more than 0.5 %
 * faster there.
 * With the binned split it then moves
 * the split's cuts so that the lanes' measured times (each from its first
 * launch to its last, trials of 4 x `steps` steps) come out equal, keeping
 * a new split unless the batch gets more than 0.5 % slower.  The batch's
 * device state is copied before and put back after every trial, so the run
 * continues from where it was; no result depends on the shapes or the split
 * chosen (maxima only, chunks of the partitioned solves set by the row
 * count, cases independent).  Call after lica_batch_set_lanes /
 * set_lane_split / set_poisson, which drop it.                             */
int lica_batch_autotune(lica_batch *b, int64_t steps);
/* per lane (3 + LICA_KERNEL_CLASSES values): cases, largest n, multigrid LDS
 * bytes, then threads per workgroup of every kernel class in the library's
 * order (init, dt, cont_energy, temp_grad, momentum, poisson_direct, ucalc,
 * transport, poisson_mg, capture_pre, capture_post, warm_base, convrge1);
 * *count receives the number of values (out may be null to ask for it)    */
#define LICA_KERNEL_CLASSES 13
int lica_batch_get_tuning(lica_batch *b, int64_t *out, int64_t cap, int64_t *count);
/* the network a batch uses in LICA_POISSON_LEARNED (not owned by the batch) */
int lica_batch_set_net(lica_batch *b, lica_net *net);
/* keep, for every case, the starting point of substep msub's POISSON solve
 * (1..3; 0 switches capture off) -- the warm-start base B in the PROJECTED
 * and LEARNED modes, else the last solution -- the new solution and the
 * right-hand side (fields LICA_F_PPREV, LICA_F_PSOL, LICA_F_PRHS)          */
int lica_batch_set_capture(lica_batch *b, int64_t on);
/* No counterpart in the reference (diagnostic of the warm start): per case
 * 7 doubles of the last learned-mode substep -- first row of the network's
 * window, backward error of the warm-start base, its residual scale s_r and
 * row scale s_b, the network step alpha (0 when the network did not run),
 * the base's backward error as the multigrid start test computes it, and 1
 * if the solve took the guess as it was (zeros before the first);
 * len >= 7 * cases                                                         */
int lica_batch_warm_info(lica_batch *b, double *out, int64_t len);

/* ---- single-case API mirroring the reference call sequence ----------- */
int lica_init(lica_batch **ctx, const char *lica_in_path, const char *grid_in_path);
int lica_step(lica_batch *ctx, int64_t nsteps);
int lica_determine_dt(lica_batch *ctx, double *cflm, double *dt);
int lica_rhsnlhs(lica_batch *ctx, int64_t msub, double *w, double *p, double *psi_zn, double *psi_cn);
int lica_convrge1(lica_batch *ctx, double *dvmax);
int lica_transport(lica_batch *ctx);
int lica_grid_coupling(lica_batch *ctx, int64_t ini_psi, double *psi_zn, double *psi_cn);
int lica_end_step(lica_batch *ctx);   /* TIME=TIME+DT (1d.f:260) */

/* ---- seams with the reference's own argument lists -------------------- */
/* Arrays are the reference's REAL*8 (0:M1,0:M2,0:M3) of the 1-D
 * configuration, M1 = M2 = 2 (1dmodulesX.f:5), in Fortran order
 * ((I,J,K) at I + 3 (J + 3 K)), scalars by reference, so a Fortran driver
 * passes its own arrays unchanged.  Each call copies the columns I = J = 1
 * it reads into the context at entry and writes the ones the subroutine
 * changes back at exit; the context owns the module state (T, the band,
 * PSIF, SFGAS, WO, ...).  x/y arrays, QVOL_ORI, VOL1 and ANPSO are accepted
 * and not used (identically zero or unused in this configuration).
 *   lica_f_determine_dt  DETERMINE_DT(CFLM,U,V,W,PSI_CN) 1d.f:683 (caller
 *                        1d.f:201); DT, a module variable there, in and out
 *                        as the last argument
 *   lica_f_rhsnlhs       RHSNLHS(ITRACKING,U,V,W,P,PSI_XN,PSI_YN,PSI_ZN,PSI_CN,
 *                        RK3XO,RK3YO,RK3ZO,ANPSO,QVOL_ORI,VOL1) 1d.f:1444
 *                        (caller 1d.f:227); MSUB, MAIN's loop variable, first
 *   lica_f_convrge1      CONVRGE1(DVMAX,U,V,W) 1d.f:3306 (caller 1d.f:234)
 *   lica_f_transport     TRANSPORT(U,V,W,PSI_CN,VOL_TOT_ORI) two_phase.f:701
 *                        (caller 1d.f:242); W must be the W the preceding
 *                        RHSNLHS returned (LICA_BAD_ARG otherwise)
 *   lica_f_grid_coupling GRID_COUPLING(INI_PSI,PSI_XN,PSI_YN,PSI_ZN,PSI_CN)
 *                        two_phase.f:4863 (caller 1d.f:244), INI_PSI = 0
 *   lica_f_poisson       POISSON(P,DIVGSUM,DENF_XI,DENF_YI,DENF_ZI,DENF_Y,DENF_C)
 *                        1d.f:3492 (caller 1d.f:3428): DIVGSUM(1,1,1:N3M),
 *                        the matrix from DENF_ZI and the context's geometry,
 *                        solved in the context's Poisson mode (not the
 *                        warm-started ones)
 *   lica_f_temp_grad     TEMP_GRAD(PSI_CN) chphase.f:55 (caller 1d.f:3667)  */
int lica_f_determine_dt(lica_batch *ctx, double *cflm, const double *u, const double *v, const double *w,
                        const double *psi_cn, double *dt);
int lica_f_rhsnlhs(lica_batch *ctx, const int64_t *msub, const int64_t *itracking, double *u, double *v, double *w,
                   double *p, double *psi_xn, double *psi_yn, double *psi_zn, double *psi_cn, double *rk3xo,
                   double *rk3yo, double *rk3zo, double *anpso, double *qvol_ori, double *vol1);
int lica_f_convrge1(lica_batch *ctx, double *dvmax, const double *u, const double *v, const double *w);
int lica_f_transport(lica_batch *ctx, const double *u, const double *v, const double *w, double *psi_cn,
                     double *vol_tot_ori);
int lica_f_grid_coupling(lica_batch *ctx, const int64_t *ini_psi, double *psi_xn, double *psi_yn, double *psi_zn,
                         double *psi_cn);
int lica_f_poisson(lica_batch *ctx, double *p, const double *divgsum, const double *denf_xi, const double *denf_yi,
                   const double *denf_zi, const double *denf_y, const double *denf_c);
int lica_f_temp_grad(lica_batch *ctx, const double *psi_cn);
/* the context's module state into the caller's arrays (after lica_init: the
 * state MAKEFLD, LVSINIT and GRID_COUPLING(1) leave, 1d.f:70-190); NULL
 * pointers are skipped */
int lica_f_state(lica_batch *ctx, double *w, double *p, double *t, double *psi_zn, double *psi_cn,
                 double *vol_tot_ori, double *time, double *dt);

/* Line solvers of the 3-D code (no live caller on the 1-D column).  Arrays
 * are column-major m1 x m2 doubles (element (i, j), 1-based, at
 * (i-1) + (j-1)*m1): the reference's A(M2,M1) dummies of TRDIAG1 have
 * m1 = M2, m2 = M1, those of TRDIAG2 m1 = M1, m2 = M2, and so on.  The
 * systems run along the second index, one per line of the first; results
 * are the reference's serial loops' bit for bit.  Host arrays: the call
 * copies them to the GPU and back and returns when done; elements outside
 * the ranges are left as they were.
 *
 * TRDIAG1 (1d.f:3721-3749), TRDIAG2 (3752-3780), TRDIAG3 (3783-3812):
 * A u(j-1) + B u(j) + C u(j+1) = R for j = l1..l2 on lines i = ll1..ll2
 * into UU; TRDIAG3 also leaves UU in R(i, l1..l2-1), as the reference does. */
int lica_trdiag1(const double *a, const double *b, const double *c, const double *r, double *uu, int64_t m1,
                 int64_t m2, int64_t l1, int64_t l2, int64_t ll1, int64_t ll2);
int lica_trdiag2(const double *a, const double *b, const double *c, const double *r, double *uu, int64_t m1,
                 int64_t m2, int64_t l1, int64_t l2, int64_t ll1, int64_t ll2);
int lica_trdiag3(const double *a, const double *b, const double *c, double *r, double *uu, int64_t m1,
                 int64_t m2, int64_t l1, int64_t l2, int64_t ll1, int64_t ll2);
/* TRDIAG1P (1d.f:3815-3869), TRDIAG2P (3872-3926), TRDIAG3P (3929-3985):
 * periodic systems along j = j1..j2 (j2 > j1; A(j1) couples to u(j2), C(j2)
 * to u(j1)) on lines k = l1..l2; F holds the right-hand side in and the
 * solution out. */
int lica_trdiag1p(const double *a, const double *b, const double *c, double *f, int64_t m1, int64_t m2, int64_t j1,
                  int64_t j2, int64_t l1, int64_t l2);
int lica_trdiag2p(const double *a, const double *b, const double *c, double *f, int64_t m1, int64_t m2, int64_t j1,
                  int64_t j2, int64_t l1, int64_t l2);
int lica_trdiag3p(const double *a, const double *b, const double *c, double *f, int64_t m1, int64_t m2, int64_t j1,
                  int64_t j2, int64_t l1, int64_t l2);
/* PTDIAG1A (1d.f:3988-4034): E u(j-2) + D u(j-1) + A u(j) + B u(j+1) +
 * C u(j+2) = F for j = j1..jn (jn > j1) on lines i = i1..in into UU; O, Q,
 * R receive the elimination's coefficients, as in the reference. */
int lica_ptdiag1a(const double *a, const double *b, const double *c, const double *d, const double *e,
                  const double *f, double *uu, int64_t m1, int64_t m2, int64_t i1, int64_t in, int64_t j1,
                  int64_t jn, double *o, double *q, double *r);
/* The same three solvers on DEVICE arrays, asynchronous on `stream` (a
 * hipStream_t; NULL for the null stream), for data that stays on the GPU.
 * work: m1 x m2 doubles (TRDIAG's multipliers) or 3 x m1 x m2 (the periodic
 * solve's), owned by the caller; nothing is allocated.  overwrite_r = 1 is
 * TRDIAG3.  No counterpart in the reference beyond the routines above. */
int lica_trdiag_lines_dev(const double *a, const double *b, const double *c, double *r, double *uu, double *work,
                          int64_t m1, int64_t m2, int64_t l1, int64_t l2, int64_t ll1, int64_t ll2,
                          int64_t overwrite_r, void *stream);
int lica_trdiag_periodic_lines_dev(const double *a, const double *b, const double *c, double *f, double *work,
                                   int64_t m1, int64_t m2, int64_t j1, int64_t j2, int64_t l1, int64_t l2,
                                   void *stream);
int lica_ptdiag_lines_dev(const double *a, const double *b, const double *c, const double *d, const double *e,
                          const double *f, double *uu, int64_t m1, int64_t m2, int64_t i1, int64_t in, int64_t j1,
                          int64_t jn, double *o, double *q, double *r, void *stream);

/* batched POISSON system of 1d.f:3535-3564: rows k=1..n of each of nsys
 * systems (a(1) is ignored as in the reference), contiguous per system.  */
int lica_poisson(int64_t nsys, int64_t n, const double *a, const double *b, const double *c,
                 const double *g, double *p, int64_t mode, double rtol, int64_t max_cycles,
                 int64_t *cycles_out);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc): lica_poisson on DEVICE arrays a, b, c, g -> p (nsys x n,
 * contiguous per system, the same meaning and results as lica_poisson) with
 * no allocation of any kind.  The kernels' workspace is one fixed 64 MiB
 * device array per GPU, held until the process ends: the multigrid modes fit
 * nsys * (11 * n + a few) doubles, the direct modes 4 * nsys * n doubles;
 * larger requests return LICA_TOO_LARGE.  Work goes to `stream` (a
 * hipStream_t; NULL for the null stream) and the call returns when it is
 * done.  Every call on one GPU uses that GPU's workspace: calls that overlap
 * on one GPU race.                                                          */
int lica_poisson_noalloc(int64_t nsys, int64_t n, const double *a, const double *b, const double *c,
                         const double *g, double *p, int64_t mode, double rtol, int64_t max_cycles,
                         int64_t *cycles_out, void *stream);

const char *lica_status_string(int status);

#ifdef __cplusplus
}
#endif
#endif
