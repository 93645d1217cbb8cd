// Host-side view of the kernel launchers (shared by the GPU build and the
// CPU logic-check build of the same kernel source).
#pragma once
#include "lica_types.h"

enum KernelId {
    KL_INIT = 0, KL_DT, KL_CONT_ENERGY, KL_TEMP_GRAD, KL_MOMENTUM,
    KL_POISSON_DIRECT, KL_UCALC, KL_TRANSPORT, KL_POISSON_MG, KL_CAPTURE_PRE, KL_CAPTURE_POST, KL_WARM_BASE,
    KL_CONVRGE1, KL_COUNT
};

// chunks of the partitioned tridiagonal solve of nrow rows: a kernel that
// solves systems of up to nmax rows needs lica_tri_chunks(nmax) threads to
// take the partitioned path
int lica_tri_chunks(int nrow);

int lica_launch(int which, const BatchDev &bd, int threads, void *stream);
// the level set of cases restarted from a field file (one wave per case)
int lica_launch_init_fld(const BatchDev &bd, void *stream);
int lica_launch_tri(int nsys, int n, double *a, double *b, double *c, double *g, int mode, int threads, void *stream);
int lica_launch_mg(int nsys, int n, const double *a, const double *b, const double *c, const double *g, double *p,
                   double *scratch, int mode, double rtol, int maxc, int *cycles, double *resid, int threads,
                   void *stream);
long long lica_mg_scratch_doubles(int n);
// the line solvers of the 3-D code (lica_kern_lines.h) on device arrays,
// column-major m1 x m2, 1-based ranges
int lica_launch_trdiag_lines(const double *a, const double *b, const double *c, double *r, double *uu, double *gam,
                             long long m1, int l1, int l2, int i1, int i2, int overwrite_r, void *stream);
int lica_launch_trdiag_periodic_lines(const double *a, const double *b, const double *c, double *f, double *q,
                                      double *s, double *qe, long long m1, int j1, int j2, int k1, int k2,
                                      void *stream);
int lica_launch_ptdiag_lines(const double *a, const double *b, const double *c, const double *d, const double *e,
                             const double *f, double *uu, double *o, double *q, double *r, long long m1, int i1,
                             int i2, int j1, int jn, void *stream);
// the fixed workspace of lica_poisson_noalloc: this device's address of the
// array (0, or -1 when HIP cannot give it)
#define LICA_NOALLOC_SCRATCH_BYTES (64ull << 20)
#define LICA_NOALLOC_MAX_DEVICES 64
int lica_noalloc_scratch(double **out);
long long lica_mg_lds_need(int n, int tsize);   // LDS for x, f of every level
long long lica_mg_lds_max();                   // largest dynamic LDS of a multigrid workgroup

// warm-start network (lica_net.hip): one POISSON system per case
struct NetCase {
    const double *a, *b, *c, *g;       // rows 1..n of the case's POISSON system (index k)
    double *guess;                     // rows 1..n: the warm-start base B, corrected by s * network
    const double *place;               // interface height of the case (its state)
    double dzi;                        // 1 / NS cell height (uniform grid)
    int n, win;                        // rows; cells of the corrected window around the interface
};
// one case's captured POISSON system (rows 1..n): a, b, c, right-hand side,
// the starting point (p0) and the solution (p1)
struct NetPairSrc {
    const double *a, *b, *c, *g, *p0, *p1;
    int n;
    int kif;     // NS row of the interface (windows are centred there, with jitter)
};
struct lica_net;
// append up to wpc windows of len cells of every listed case to the network's
// training set (features as k_net_features, target (p1 - p0) over its
// maximum magnitude in the window)
int lica_net_append(lica_net *t, const NetPairSrc *hsrc, int ncase, int64_t wpc, int64_t len, uint64_t seed,
                    int64_t *added);
int64_t lica_net_window();   // cells of the corrected window around the interface
// network correction of the warm-start guesses of a batch (one launch);
// dinfo holds LICA_WARM_INFO doubles per case, [1..3] = backward error, s_r,
// s_b of the base (written by the warm-start base kernel), [0] = window
// start and [4] = alpha (written here), [6] = the multigrid start verdict
// on the base (cleared here for the cases whose guess moves)
// (order: the cases of the launch, ncase of them; null for cases 0..ncase-1)
int lica_net_guess(lica_net *t, const NetCase *dcases, const int *order, int ncase, double *dinfo, double rtol,
                   void *stream);
int lica_net_cast(lica_net *t, void *stream);   // BF16 weights from the FP32 masters
