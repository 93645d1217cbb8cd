// Workgroup-level parallel primitives used by the kernels.
//
// The kernels are written as a uniform orchestration (executed by every
// thread of the workgroup) over three primitives:
//   PFOR(i, lo, hi) { body }  -- work items i in [lo,hi) spread over the
//                                workgroup, followed by a barrier
//   SERIAL { body }           -- executed by thread 0, followed by a barrier
//   red_max / red_min         -- exact order-free reductions
// Uniform decisions read shared (LDS) or global state after a barrier.
//
// With LICA_EMUL defined the same source compiles as plain C++ where a
// workgroup runs its work items in order; that build exists only for the
// CPU logic tests of the kernels and is never part of the product library.
#pragma once

#ifdef LICA_EMUL
#include <cmath>
#include <cstring>
#include <algorithm>
using std::isfinite;
#include <cstdio>
#include <cstdlib>
#define DEV inline
#define HOSTDEV inline
#define SHARED static
// Every barrier-bearing primitive must be reached by the whole workgroup:
// the logic-check build aborts when one is entered from inside a PFOR body
// or a SERIAL section (on the GPU that would be a barrier in divergent code).
extern int lica_emul_depth;
inline void lica_emul_check()
{
    if (lica_emul_depth) {
        fprintf(stderr, "lica: workgroup barrier inside divergent code\n");
        abort();
    }
}
struct EmulScope {
    EmulScope() { lica_emul_check(); ++lica_emul_depth; }
    ~EmulScope() { --lica_emul_depth; }
};
#define BARRIER() lica_emul_check()
#define PFOR(i, lo, hi) if (EmulScope lica_scope_{}; true) for (int i = (lo); i < (hi); ++i)
#define PFOR_END
#define SERIAL if (EmulScope lica_scope_{}; true)
#define SERIAL_END
extern int lica_emul_nthreads;
#define NTHREADS (lica_emul_nthreads)
#else
#include <hip/hip_runtime.h>
#define DEV __device__ __forceinline__
#define HOSTDEV __host__ __device__ __forceinline__
#define SHARED __shared__
#define BARRIER() __syncthreads()
#define PFOR(i, lo, hi) for (int i = (lo) + (int)threadIdx.x; i < (hi); i += (int)blockDim.x)
#define PFOR_END __syncthreads();
#define SERIAL if (threadIdx.x == 0)
#define SERIAL_END __syncthreads();
#define NTHREADS ((int)blockDim.x)
#endif

// exact max/min over f(i), i in [lo,hi), combined with init
template <class F>
DEV double red_max(int lo, int hi, double init, F f, double *scr)
{
#ifdef LICA_EMUL
    (void)scr;
    lica_emul_check();
    double m = init;
    for (int i = lo; i < hi; ++i) m = fmax(m, f(i));
    return m;
#else
    double m = init;
    for (int i = lo + (int)threadIdx.x; i < hi; i += (int)blockDim.x) m = fmax(m, f(i));
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
    const int w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    if ((threadIdx.x & 63) == 0) scr[w] = m;
    __syncthreads();
    double r = init;
    for (int i = 0; i < nw; ++i) r = fmax(r, scr[i]);
    __syncthreads();
    return r;
#endif
}

template <class F>
DEV double red_min(int lo, int hi, double init, F f, double *scr)
{
#ifdef LICA_EMUL
    (void)scr;
    lica_emul_check();
    double m = init;
    for (int i = lo; i < hi; ++i) m = fmin(m, f(i));
    return m;
#else
    double m = init;
    for (int i = lo + (int)threadIdx.x; i < hi; i += (int)blockDim.x) m = fmin(m, f(i));
    for (int off = 32; off > 0; off >>= 1) m = fmin(m, __shfl_xor(m, off, 64));
    const int w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    if ((threadIdx.x & 63) == 0) scr[w] = m;
    __syncthreads();
    double r = init;
    for (int i = 0; i < nw; ++i) r = fmin(r, scr[i]);
    __syncthreads();
    return r;
#endif
}

// two maxima at once over f(i, u, v) (u, v >= 0), i in [lo,hi); f may also
// write memory, which is visible to the whole workgroup on return
template <class F>
DEV void red_max2(int lo, int hi, F f, double *scr, double *m0, double *m1)
{
#ifdef LICA_EMUL
    (void)scr;
    lica_emul_check();
    double a = 0.0, b = 0.0;
    for (int i = lo; i < hi; ++i) {
        double u, v;
        f(i, u, v);
        a = fmax(a, u);
        b = fmax(b, v);
    }
    *m0 = a;
    *m1 = b;
#else
    double a = 0.0, b = 0.0;
    for (int i = lo + (int)threadIdx.x; i < hi; i += (int)blockDim.x) {
        double u, v;
        f(i, u, v);
        a = fmax(a, u);
        b = fmax(b, v);
    }
    for (int off = 32; off > 0; off >>= 1) {
        a = fmax(a, __shfl_xor(a, off, 64));
        b = fmax(b, __shfl_xor(b, off, 64));
    }
    const int w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    if ((threadIdx.x & 63) == 0) { scr[2 * w] = a; scr[2 * w + 1] = b; }
    __syncthreads();
    double ra = 0.0, rb = 0.0;
    for (int i = 0; i < nw; ++i) { ra = fmax(ra, scr[2 * i]); rb = fmax(rb, scr[2 * i + 1]); }
    __syncthreads();
    *m0 = ra;
    *m1 = rb;
#endif
}

// N sums at once over f(i, v[N]), i in [lo,hi); the order of the additions
// is fixed for a given workgroup size
template <int N, class F>
DEV void red_sumN(int lo, int hi, F f, double *scr, double *out)
{
#ifdef LICA_EMUL
    (void)scr;
    lica_emul_check();
    double acc[N];
    for (int q = 0; q < N; q++) acc[q] = 0.0;
    for (int i = lo; i < hi; ++i) {
        double v[N];
        f(i, v);
        for (int q = 0; q < N; q++) acc[q] += v[q];
    }
    for (int q = 0; q < N; q++) out[q] = acc[q];
#else
    double acc[N];
    for (int q = 0; q < N; q++) acc[q] = 0.0;
    for (int i = lo + (int)threadIdx.x; i < hi; i += (int)blockDim.x) {
        double v[N];
        f(i, v);
        for (int q = 0; q < N; q++) acc[q] += v[q];
    }
    for (int off = 32; off > 0; off >>= 1)
        for (int q = 0; q < N; q++) acc[q] += __shfl_xor(acc[q], off, 64);
    const int w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int q = 0; q < N; q++) scr[w * N + q] = acc[q];
    __syncthreads();
    for (int q = 0; q < N; q++) {
        double s = 0.0;
        for (int i = 0; i < nw; ++i) s += scr[i * N + q];
        out[q] = s;
    }
    __syncthreads();
#endif
}

// logical OR over work items (as an int reduction)
template <class F>
DEV int red_or(int lo, int hi, F f, double *scr)
{
    return red_max(lo, hi, 0.0, [&](int i) { return f(i) ? 1.0 : 0.0; }, scr) > 0.5;
}
