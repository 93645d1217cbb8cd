// Host-side view of the kernel launchers (shared by the GPU build and the
// CPU logic-check build of the same kernel source).
#pragma once
#include "lica_types.h"

enum KernelId {
    KL_INIT = 0, KL_DT, KL_CONT_ENERGY, KL_TEMP_GRAD, KL_MOMENTUM,
    KL_POISSON_DIRECT, KL_UCALC, KL_TRANSPORT, KL_POISSON_MG, KL_CAPTURE_PRE, KL_CAPTURE_POST, KL_COUNT
};

int lica_launch(int which, const BatchDev &bd, int threads, void *stream);
int lica_launch_tri(int nsys, int n, double *a, double *b, double *c, double *g, int mode, int threads, void *stream);
int lica_launch_mg(int nsys, int n, const double *a, const double *b, const double *c, const double *g, double *p,
                   double *scratch, int mode, double rtol, int maxc, int *cycles, double *resid, int threads,
                   void *stream);
long long lica_mg_scratch_doubles(int n);
long long lica_mg_lds_need(int n, int tsize);   // LDS for x, f of every level
long long lica_mg_lds_max();                   // largest dynamic LDS of a multigrid workgroup

// warm-start network (lica_net.hip): one POISSON system per case
struct NetCase {
    const double *a, *b, *c, *g, *p;   // rows 1..n of the case (index k)
    const double *pshift;              // P + *pshift is the case's last solution
    double *guess;                     // rows 1..n: receives P + s_p * network
    long long off;                     // first cell of the case in the activation arena
    int n;
};
// one case's captured POISSON system (rows 1..n): a, b, c, right-hand side,
// P before (p0) and after (p1) the solve
struct NetPairSrc {
    const double *a, *b, *c, *g, *p0, *p1;
    int n;
};
struct lica_net;
// append up to wpc windows of len cells of every listed case to the network's
// training set (features as k_net_features, target (p1 - p0) / max|p0|)
int lica_net_append(lica_net *t, const NetPairSrc *hsrc, int ncase, int64_t wpc, int64_t len, uint64_t seed,
                    int64_t *added);
int lica_net_guess(lica_net *t, const NetCase *dcases, int ncase, const int64_t *lens, double *dscale, void *stream);
int lica_net_prepare(lica_net *t, const int64_t *lens, int ncase, const NetCase *dcases);
