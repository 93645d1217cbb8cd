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
// a work item's first index and the stride between its indices
#define PTID 0
#define PSTRIDE 1
#define PLOOP if (EmulScope lica_scope_{}; true)
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
#define PTID ((int)threadIdx.x)
#define PSTRIDE ((int)blockDim.x)
#define PLOOP

// Lane moves of a 64-lane wave on DPP (data moved inside the VALU, no LDS
// crossbar): CTRL 0x138 wave_shr:1 (lane i gets lane i-1), 0x130 wave_shl:1
// (lane i gets lane i+1), 0x121..0x128 row_ror:1..8 (rotation within each
// row of 16 lanes).  A lane without a source keeps its own value, as
// __shfl_up/__shfl_down by one lane do at the wave's edge.
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v)
{
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    int lo = (int)(unsigned)(b & 0xffffffffull), hi = (int)(unsigned)(b >> 32);
    lo = __builtin_amdgcn_update_dpp(lo, lo, CTRL, 0xF, 0xF, false);
    hi = __builtin_amdgcn_update_dpp(hi, hi, CTRL, 0xF, 0xF, false);
    return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned long long)(unsigned)lo);
}
__device__ __forceinline__ double lane_up1(double v) { return dpp_mov<0x138>(v); }   // = __shfl_up(v, 1)
__device__ __forceinline__ double lane_dn1(double v) { return dpp_mov<0x130>(v); }   // = __shfl_down(v, 1)
__device__ __forceinline__ double lane_read(double v, int lane)
{
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b & 0xffffffffull), lane);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), lane);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
// maximum / minimum over the 64 lanes (exact in any order), the same value
// in every lane
__device__ __forceinline__ double wave_red_max(double m)
{
    m = fmax(m, dpp_mov<0x121>(m));
    m = fmax(m, dpp_mov<0x122>(m));
    m = fmax(m, dpp_mov<0x124>(m));
    m = fmax(m, dpp_mov<0x128>(m));
    return fmax(fmax(lane_read(m, 0), lane_read(m, 16)), fmax(lane_read(m, 32), lane_read(m, 48)));
}
__device__ __forceinline__ double wave_red_min(double m)
{
    m = fmin(m, dpp_mov<0x121>(m));
    m = fmin(m, dpp_mov<0x122>(m));
    m = fmin(m, dpp_mov<0x124>(m));
    m = fmin(m, dpp_mov<0x128>(m));
    return fmin(fmin(lane_read(m, 0), lane_read(m, 16)), fmin(lane_read(m, 32), lane_read(m, 48)));
}
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
    m = wave_red_max(m);
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
    m = wave_red_min(m);
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
    a = wave_red_max(a);
    b = wave_red_max(b);
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

// N maxima at once over f(i, v[N]) (v >= 0), i in [lo,hi); f may also write
// memory, which is visible to the whole workgroup on return
template <int N, class F>
DEV void red_maxN(int lo, int hi, F f, double *scr, double *out)
{
#ifdef LICA_EMUL
    (void)scr;
    lica_emul_check();
    double acc[N];
    for (int q = 0; q < N; q++) acc[q] = 0.0;
    for (int i = lo; i < hi; ++i) {
        double v[N];
        f(i, v);
        for (int q = 0; q < N; q++) acc[q] = fmax(acc[q], v[q]);
    }
    for (int q = 0; q < N; q++) out[q] = acc[q];
#else
    double acc[N];
    for (int q = 0; q < N; q++) acc[q] = 0.0;
    for (int i = lo + (int)threadIdx.x; i < hi; i += (int)blockDim.x) {
        double v[N];
        f(i, v);
        for (int q = 0; q < N; q++) acc[q] = fmax(acc[q], v[q]);
    }
    for (int q = 0; q < N; q++) acc[q] = wave_red_max(acc[q]);
    const int w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int q = 0; q < N; q++) scr[w * N + q] = acc[q];
    __syncthreads();
    for (int q = 0; q < N; q++) {
        double r = 0.0;
        for (int i = 0; i < nw; ++i) r = fmax(r, scr[i * N + q]);
        out[q] = r;
    }
    __syncthreads();
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
