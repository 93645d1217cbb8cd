// Kernel entry points and launchers of the per-step path (gfx950).
#include "lica_kern_flow.h"
#include "lica_kern_lines.h"
#include "lica_launch.h"
#include <mutex>

long long lica_mg_scratch_doubles(int n) { return mg_scratch_doubles(n); }
int lica_tri_chunks(int nrow) { return nrow >= 4 ? tri_chunks(nrow) : 1; }
long long lica_mg_lds_need(int n, int tsize) { return mg_lds_bytes(n, tsize); }

// Fixed device workspace of lica_poisson_noalloc (callers that cannot
// allocate): one array per GPU, held until the process ends.  Host code
// never uses the array's name as an address: the address of each GPU's copy
// comes from hipGetSymbolAddress with that device current, once per device.
#ifdef LICA_EMUL
static double scratch[LICA_NOALLOC_SCRATCH_BYTES / sizeof(double)];
int lica_noalloc_scratch(double **out)
{
    *out = scratch;
    return 0;
}
#else
static __device__ __align__(256) double scratch[LICA_NOALLOC_SCRATCH_BYTES / sizeof(double)];
int lica_noalloc_scratch(double **out)
{
    static std::mutex mu;
    static double *addr[LICA_NOALLOC_MAX_DEVICES] = {};
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= LICA_NOALLOC_MAX_DEVICES) return -1;
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[dev]) {
        void *ptr = nullptr;
        if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return -1;
        addr[dev] = (double *)ptr;
    }
    *out = addr[dev];
    return 0;
}
#endif

#ifndef LICA_EMUL
// dynamic LDS, sized by the launch: the partitioned solves' workspace over
// the refined window (BlkSh kernels, blk_dyn_bytes), or
// the multigrid cycle's vectors
extern __shared__ double lica_tri_lds[];

DEV void tri_attach(BlkSh &sh)
{
    if (threadIdx.x == 0) {
        sh.tri.ws = lica_tri_lds;
        sh.tri.ws_rows = tri_ws_rows((int)blockDim.x);
        sh.win = (RefWin *)lica_tri_lds;
    }
    __syncthreads();
}

__global__ void __launch_bounds__(1024) k_init(BatchDev bd)
{
    __shared__ InitSh sh;
    body_init(bd, bd.order[blockIdx.x], sh);
}
__global__ void __launch_bounds__(1024) k_dt(BatchDev bd)
{
    __shared__ BlkSh sh;
    tri_attach(sh);
    body_dt(bd, bd.order[blockIdx.x], sh);
}
__global__ void __launch_bounds__(1024) k_cont_energy(BatchDev bd)
{
    __shared__ BlkSh sh;
    tri_attach(sh);
    body_cont_energy(bd, bd.order[blockIdx.x], sh);
}
// 20.1 KB of LDS and at most 128 VGPRs: eight two-wave workgroups per CU,
// four waves per SIMD to hide the latency of the extrapolation chains.  With
// lead slots the grid is LICA_TG_LEAD lead workgroups, then one per case.
__global__ void __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(4))) k_temp_grad(BatchDev bd)
{
    __shared__ TgradSh sh;
    const int pos = bd.lead ? tg_lead_pos(bd, (int)blockIdx.x) : (int)blockIdx.x;
    if (pos < 0) return;
    const int inst = bd.order[pos];
    if (threadIdx.x == 0) { sh.xit[0] = -1.0; sh.xit[1] = -1.0; }
    body_temp_grad(bd, inst, sh);
    if (bd.lead) {
        __syncthreads();
        if (threadIdx.x == 0) tg_lead_record(bd, inst, pos, sh);
    }
}
// the momentum stage of the Thomas and multigrid modes (no refinement in
// its solve)
__global__ void __launch_bounds__(1024) k_momentum(BatchDev bd)
{
    __shared__ BlkSh sh;
    tri_attach(sh);
    body_momentum<0>(bd, bd.order[blockIdx.x], sh);
}
// the direct Poisson mode's momentum stage (refined partitioned solve)
__global__ void __launch_bounds__(1024) k_momentum_refine(BatchDev bd)
{
    __shared__ BlkSh sh;
    tri_attach(sh);
    body_momentum<TRI_NREF>(bd, bd.order[blockIdx.x], sh);
}
__global__ void __launch_bounds__(1024) k_poisson_direct(BatchDev bd)
{
    __shared__ BlkSh sh;
    tri_attach(sh);
    body_poisson_direct(bd, bd.order[blockIdx.x], sh);
}
DEV void mg_attach(MgSh &sh, long long lds_bytes)
{
    if (threadIdx.x == 0) { sh.lds = (void *)lica_tri_lds; sh.lds_bytes = lds_bytes; }
    __syncthreads();
}
// six waves per SIMD (80 VGPRs, a few spills): three 512-thread workgroups
// per CU instead of two
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(6))) k_poisson_mg(BatchDev bd)
{
    __shared__ MgSh sh;
    mg_attach(sh, bd.mg_lds);
    body_poisson_mg(bd, bd.order[blockIdx.x], sh);
}

// nsys independent POISSON systems by multigrid (scratch: mg_scratch_doubles(n) each)
__global__ void __launch_bounds__(1024) k_mg_systems(const double *a, const double *b, const double *c, const double *g,
                                                     double *p, int n, double *scratch, int mode, double rtol,
                                                     int maxc, int *cycles, double *resid, long long lds_bytes)
{
    __shared__ MgSh sh;
    mg_attach(sh, lds_bytes);
    const size_t off = (size_t)blockIdx.x * (size_t)n - 1;
    double res = 0.0;
    int cyc = mg_solve(a + off, b + off, c + off, g + off, p + off, n,
                       scratch + (size_t)blockIdx.x * (size_t)mg_scratch_doubles(n), mode, rtol, maxc, nullptr, sh,
                       &res);
    if (threadIdx.x == 0) { cycles[blockIdx.x] = cyc; resid[blockIdx.x] = res; }
}

// largest dynamic LDS a multigrid workgroup may take (the static part is small)
long long lica_mg_lds_max()
{
    static long long cap = -1;
    if (cap < 0) {
        int dev = 0, optin = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, dev) != hipSuccess || optin <= 0)
            optin = 65536;
        // the static part (MgSh) of the larger of the two kernels comes off
        const void *ks[2] = {(const void *)k_poisson_mg, (const void *)k_mg_systems};
        long long stat = 4096;
        for (int i = 0; i < 2; i++) {
            hipFuncAttributes fa;
            if (hipFuncGetAttributes(&fa, ks[i]) == hipSuccess)
                stat = std::max(stat, (long long)fa.sharedSizeBytes);
        }
        cap = ((long long)optin - stat) / 256 * 256;
        for (int i = 0; i < 2; i++)
            if (hipFuncSetAttribute(ks[i], hipFuncAttributeMaxDynamicSharedMemorySize, (int)cap) != hipSuccess) {
                cap = (65536 - stat) / 256 * 256;
                break;
            }
    }
    return cap;
}

int lica_launch_mg(int nsys, int n, const double *a, const double *b, const double *c, const double *g, double *p,
                   double *scratch, int mode, double rtol, int maxc, int *cycles, double *resid, int threads,
                   void *stream)
{
    if (nsys <= 0) return 0;
    long long lds = mg_lds_bytes(n, mode == LICA_MG_FP64_MODE ? 8 : 4);
    lds = std::min(lds, std::min(lica_mg_lds_max(), (long long)(nsys >= 512 ? 65536 - 4096 : 1 << 30)));
    hipLaunchKernelGGL(k_mg_systems, dim3(nsys), dim3(threads), (size_t)lds, (hipStream_t)stream, a, b, c, g, p, n,
                       scratch, mode, rtol, maxc, cycles, resid, lds);
    return (int)hipGetLastError();
}

__global__ void __launch_bounds__(256) k_warm_base(BatchDev bd)
{
    __shared__ WarmSh sh;
    body_warm_base(bd, bd.order[blockIdx.x], sh);
}

__global__ void __launch_bounds__(1024) k_capture(BatchDev bd, int which)
{
    body_capture(bd, bd.order[blockIdx.x], which);
}

__global__ void __launch_bounds__(1024) k_ucalc(BatchDev bd)
{
    __shared__ BlkSh sh;
    tri_attach(sh);
    body_ucalc(bd, bd.order[blockIdx.x], sh);
}
__global__ void __launch_bounds__(1024) k_convrge1(BatchDev bd)
{
    __shared__ BlkSh sh;
    tri_attach(sh);
    body_convrge1(bd, bd.order[blockIdx.x], sh);
}
__global__ void __launch_bounds__(64) k_transport(BatchDev bd)
{
    __shared__ TransportSh sh;
    body_transport(bd, bd.order[blockIdx.x], sh);
}
__global__ void __launch_bounds__(64) k_init_fld(BatchDev bd)
{
    __shared__ TransportSh sh;
    body_init_fld(bd, bd.order[blockIdx.x], sh);
}

int lica_launch_init_fld(const BatchDev &bd, void *stream)
{
    if (bd.ninst <= 0) return 0;
    hipLaunchKernelGGL(k_init_fld, dim3(bd.ninst), dim3(64), 0, (hipStream_t)stream, bd);
    return (int)hipGetLastError();
}

static int tri_lds_setup();

// nsys independent tridiagonal systems of n rows, contiguous per system
__global__ void __launch_bounds__(1024) k_tri_systems(double *a, double *b, double *c, double *g, int n, int mode)
{
    __shared__ BlkSh sh;
    tri_attach(sh);
    const size_t off = (size_t)blockIdx.x * (size_t)n - 1;   // row k at [off + k], k = 1..n
    tri_solve(a + off, b + off, c + off, g + off, 1, n, mode, sh.tri);
}

int lica_launch_tri(int nsys, int n, double *a, double *b, double *c, double *g, int mode, int threads, void *stream)
{
    if (nsys <= 0) return 0;
    if (int rc = tri_lds_setup()) return rc;
    const size_t lds = (size_t)blk_dyn_bytes(threads, 1);
    hipLaunchKernelGGL(k_tri_systems, dim3(nsys), dim3(threads), lds, (hipStream_t)stream, a, b, c, g, n, mode);
    return (int)hipGetLastError();
}

// the line solvers (lica_kern_lines.h): one thread per line i = i1..i2
__global__ void __launch_bounds__(256) k_trdiag_lines(const double *a, const double *b, const double *c, double *r,
                                                      double *uu, double *gam, long long m1, int l1, int l2, int i1,
                                                      int i2, int overwrite_r)
{
    const int i = i1 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i <= i2) line_trdiag(a, b, c, r, uu, gam, m1, i, l1, l2, overwrite_r);
}
__global__ void __launch_bounds__(256) k_trdiag_periodic_lines(const double *a, const double *b, const double *c,
                                                               double *f, double *q, double *s, double *qe,
                                                               long long m1, int j1, int j2, int k1, int k2)
{
    const int k = k1 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (k <= k2) line_trdiag_periodic(a, b, c, f, q, s, qe, m1, k, j1, j2);
}
__global__ void __launch_bounds__(256) k_ptdiag_lines(const double *a, const double *b, const double *c,
                                                      const double *d, const double *e, const double *f, double *uu,
                                                      double *o, double *q, double *r, long long m1, int i1, int i2,
                                                      int j1, int jn)
{
    const int i = i1 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i <= i2) line_ptdiag(a, b, c, d, e, f, uu, o, q, r, m1, i, j1, jn);
}
static dim3 line_grid(int i1, int i2) { return dim3((unsigned)((i2 - i1 + 1 + 255) / 256)); }
int lica_launch_trdiag_lines(const double *a, const double *b, const double *c, double *r, double *uu, double *gam,
                             long long m1, int l1, int l2, int i1, int i2, int overwrite_r, void *stream)
{
    hipLaunchKernelGGL(k_trdiag_lines, line_grid(i1, i2), dim3(256), 0, (hipStream_t)stream, a, b, c, r, uu, gam, m1,
                       l1, l2, i1, i2, overwrite_r);
    return (int)hipGetLastError();
}
int lica_launch_trdiag_periodic_lines(const double *a, const double *b, const double *c, double *f, double *q,
                                      double *s, double *qe, long long m1, int j1, int j2, int k1, int k2,
                                      void *stream)
{
    hipLaunchKernelGGL(k_trdiag_periodic_lines, line_grid(k1, k2), dim3(256), 0, (hipStream_t)stream, a, b, c, f, q,
                       s, qe, m1, j1, j2, k1, k2);
    return (int)hipGetLastError();
}
int lica_launch_ptdiag_lines(const double *a, const double *b, const double *c, const double *d, const double *e,
                             const double *f, double *uu, double *o, double *q, double *r, long long m1, int i1,
                             int i2, int j1, int jn, void *stream)
{
    hipLaunchKernelGGL(k_ptdiag_lines, line_grid(i1, i2), dim3(256), 0, (hipStream_t)stream, a, b, c, d, e, f, uu, o,
                       q, r, m1, i1, i2, j1, jn);
    return (int)hipGetLastError();
}

#define LAUNCH(K, threads) \
    hipLaunchKernelGGL(K, dim3(bd.ninst), dim3(threads), 0, (hipStream_t)stream, bd)
#define LAUNCH_BLK(K, threads) \
    hipLaunchKernelGGL(K, dim3(bd.ninst), dim3(threads), (size_t)blk_dyn_bytes(threads, 0), (hipStream_t)stream, bd)
#define LAUNCH_TRI(K, threads) \
    hipLaunchKernelGGL(K, dim3(bd.ninst), dim3(threads), (size_t)blk_dyn_bytes(threads, 1), (hipStream_t)stream, bd)

// allow the dynamic LDS of the partitioned solves beyond 64 KB (1024 threads)
static int tri_lds_setup()
{
    static int done = 0, rc = 0;
    if (!done) {
        const int bytes = (int)blk_dyn_bytes(1024, 1);
        const void *ks[5] = {(const void *)k_cont_energy, (const void *)k_momentum, (const void *)k_momentum_refine,
                             (const void *)k_poisson_direct, (const void *)k_tri_systems};
        for (int i = 0; i < 5 && !rc; i++)
            rc = (int)hipFuncSetAttribute(ks[i], hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
        done = 1;
    }
    return rc;
}

int lica_launch(int which, const BatchDev &bd, int threads, void *stream)
{
    if (bd.ninst <= 0) return 0;
    if (int rc = tri_lds_setup()) return rc;
    switch (which) {
    case KL_INIT: LAUNCH(k_init, threads); break;
    case KL_DT: LAUNCH_BLK(k_dt, threads); break;
    case KL_CONT_ENERGY: LAUNCH_TRI(k_cont_energy, threads); break;
    case KL_TEMP_GRAD:
        hipLaunchKernelGGL(k_temp_grad, dim3(bd.ninst + (bd.lead ? LICA_TG_LEAD : 0)), dim3(128), 0, (hipStream_t)stream, bd);
        break;
    case KL_MOMENTUM:
        if (bd.tri_mode == 1) LAUNCH_TRI(k_momentum_refine, threads);
        else LAUNCH_TRI(k_momentum, threads);
        break;
    case KL_POISSON_DIRECT: LAUNCH_TRI(k_poisson_direct, threads); break;
    case KL_UCALC: LAUNCH_BLK(k_ucalc, threads); break;
    case KL_TRANSPORT: LAUNCH(k_transport, 64); break;
    case KL_CONVRGE1: LAUNCH_BLK(k_convrge1, threads); break;
    case KL_POISSON_MG:
        hipLaunchKernelGGL(k_poisson_mg, dim3(bd.ninst), dim3(threads), (size_t)bd.mg_lds, (hipStream_t)stream, bd);
        break;
    case KL_CAPTURE_PRE:
        hipLaunchKernelGGL(k_capture, dim3(bd.ninst), dim3(256), 0, (hipStream_t)stream, bd, 0);
        break;
    case KL_CAPTURE_POST:
        hipLaunchKernelGGL(k_capture, dim3(bd.ninst), dim3(256), 0, (hipStream_t)stream, bd, 1);
        break;
    case KL_WARM_BASE: LAUNCH(k_warm_base, 256); break;
    default: return -1;
    }
    return (int)hipGetLastError();
}
#else
// the CPU logic-check build models a gfx950 workgroup's 160 KB of LDS
long long lica_mg_lds_max() { return 160 * 1024 - 4096; }

static double *emul_tri_ws(int threads)
{
    static double *ws = nullptr;
    if (!ws) ws = new double[(size_t)blk_dyn_bytes(1024, 1) / 8 + 1];
    (void)threads;
    return ws;
}

// the same overlay as on the GPU: the window under the solve's workspace
static void attach(BlkSh &sh, int threads)
{
    sh.tri.ws = emul_tri_ws(threads);
    sh.tri.ws_rows = tri_ws_rows(threads);
    sh.win = (RefWin *)sh.tri.ws;
}

int lica_launch_tri(int nsys, int n, double *a, double *b, double *c, double *g, int mode, int threads, void *stream)
{
    (void)stream;
    static BlkSh *sh = nullptr;
    if (!sh) sh = new BlkSh();
    attach(*sh, threads);
    int saved = lica_emul_nthreads;
    lica_emul_nthreads = threads;
    for (int s = 0; s < nsys; s++) {
        const size_t off = (size_t)s * (size_t)n - 1;
        tri_solve(a + off, b + off, c + off, g + off, 1, n, mode, sh->tri);
    }
    lica_emul_nthreads = saved;
    return 0;
}

// the line solvers, one line after the other
int lica_launch_trdiag_lines(const double *a, const double *b, const double *c, double *r, double *uu, double *gam,
                             long long m1, int l1, int l2, int i1, int i2, int overwrite_r, void *stream)
{
    (void)stream;
    for (int i = i1; i <= i2; i++) line_trdiag(a, b, c, r, uu, gam, m1, i, l1, l2, overwrite_r);
    return 0;
}
int lica_launch_trdiag_periodic_lines(const double *a, const double *b, const double *c, double *f, double *q,
                                      double *s, double *qe, long long m1, int j1, int j2, int k1, int k2,
                                      void *stream)
{
    (void)stream;
    for (int k = k1; k <= k2; k++) line_trdiag_periodic(a, b, c, f, q, s, qe, m1, k, j1, j2);
    return 0;
}
int lica_launch_ptdiag_lines(const double *a, const double *b, const double *c, const double *d, const double *e,
                             const double *f, double *uu, double *o, double *q, double *r, long long m1, int i1,
                             int i2, int j1, int jn, void *stream)
{
    (void)stream;
    for (int i = i1; i <= i2; i++) line_ptdiag(a, b, c, d, e, f, uu, o, q, r, m1, i, j1, jn);
    return 0;
}

int lica_launch_mg(int nsys, int n, const double *a, const double *b, const double *c, const double *g, double *p,
                   double *scratch, int mode, double rtol, int maxc, int *cycles, double *resid, int threads,
                   void *stream)
{
    (void)stream;
    static MgSh *sh = nullptr;
    if (!sh) sh = new MgSh();
    static std::vector<double> lds;
    const long long lb = std::min(mg_lds_bytes(n, mode == LICA_MG_FP64_MODE ? 8 : 4), lica_mg_lds_max());
    lds.assign((size_t)(lb / 8 + 1), 0.0);
    sh->lds = lds.data();
    sh->lds_bytes = lb;
    int saved = lica_emul_nthreads;
    lica_emul_nthreads = threads;
    for (int s = 0; s < nsys; s++) {
        const size_t off = (size_t)s * (size_t)n - 1;
        double res = 0.0;
        cycles[s] = mg_solve(a + off, b + off, c + off, g + off, p + off, n,
                             scratch + (size_t)s * (size_t)mg_scratch_doubles(n), mode, rtol, maxc, nullptr, *sh,
                             &res);
        resid[s] = res;
    }
    lica_emul_nthreads = saved;
    return 0;
}

static double *emul_mg_lds = nullptr;
static long long emul_mg_lds_bytes = 0;
static void attach(MgSh &sh, int) { sh.lds = emul_mg_lds; sh.lds_bytes = emul_mg_lds_bytes; }
template <class Sh>
static void attach(Sh &, int) {}

template <class Sh, class F>
static void emul_run(const BatchDev &bd, int threads, F body)
{
    static Sh *sh = nullptr;
    if (!sh) sh = new Sh();
    int saved = lica_emul_nthreads;
    lica_emul_nthreads = threads;
    for (int b = 0; b < bd.ninst; b++) {
        memset((void *)sh, 0, sizeof(Sh));
        attach(*sh, threads);
        body(bd, bd.order[b], *sh);
    }
    lica_emul_nthreads = saved;
}

int lica_launch(int which, const BatchDev &bd, int threads, void *stream)
{
    (void)stream;
    switch (which) {
    case KL_INIT: emul_run<InitSh>(bd, threads, body_init); break;
    case KL_DT: emul_run<BlkSh>(bd, threads, body_dt); break;
    case KL_CONT_ENERGY: emul_run<BlkSh>(bd, threads, body_cont_energy); break;
    case KL_TEMP_GRAD: emul_run<TgradSh>(bd, 64, body_temp_grad); break;
    case KL_MOMENTUM:
        if (bd.tri_mode == 1) emul_run<BlkSh>(bd, threads, body_momentum<TRI_NREF>);
        else emul_run<BlkSh>(bd, threads, body_momentum<0>);
        break;
    case KL_POISSON_DIRECT: emul_run<BlkSh>(bd, threads, body_poisson_direct); break;
    case KL_UCALC: emul_run<BlkSh>(bd, threads, body_ucalc); break;
    case KL_TRANSPORT: emul_run<TransportSh>(bd, 64, body_transport); break;
    case KL_POISSON_MG: {
        static std::vector<double> lds;
        lds.assign((size_t)(bd.mg_lds / 8 + 1), 0.0);
        emul_mg_lds = lds.data();
        emul_mg_lds_bytes = bd.mg_lds;
        emul_run<MgSh>(bd, threads, body_poisson_mg);
        break;
    }
    case KL_CAPTURE_PRE:
        for (int b = 0; b < bd.ninst; b++) body_capture(bd, bd.order[b], 0);
        break;
    case KL_CAPTURE_POST:
        for (int b = 0; b < bd.ninst; b++) body_capture(bd, bd.order[b], 1);
        break;
    case KL_WARM_BASE: emul_run<WarmSh>(bd, 256, body_warm_base); break;
    case KL_CONVRGE1: emul_run<BlkSh>(bd, threads, body_convrge1); break;
    default: return -1;
    }
    return 0;
}

int lica_launch_init_fld(const BatchDev &bd, void *stream)
{
    (void)stream;
    emul_run<TransportSh>(bd, 64, body_init_fld);
    return 0;
}
#endif
