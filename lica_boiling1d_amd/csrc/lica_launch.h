// Host-side view of the kernel launchers (shared by the GPU build and the
// CPU logic-check build of the same kernel source).
#pragma once
#include "lica_types.h"

enum KernelId {
    KL_INIT = 0, KL_DT, KL_CONT_ENERGY, KL_TEMP_GRAD, KL_MOMENTUM,
    KL_POISSON_DIRECT, KL_UCALC, KL_TRANSPORT, KL_POISSON_MG, KL_COUNT
};

int lica_launch(int which, const BatchDev &bd, int threads, void *stream);
int lica_launch_tri(int nsys, int n, double *a, double *b, double *c, double *g, int mode, int threads, void *stream);
int lica_launch_mg(int nsys, int n, const double *a, const double *b, const double *c, const double *g, double *p,
                   double *scratch, int mode, double rtol, int maxc, int *cycles, double *resid, int threads,
                   void *stream);
long long lica_mg_scratch_doubles(int n);
