// Convolutional warm start of the POISSON solve (gfx950, BF16 MFMA).
//
// The network maps per-cell features of a POISSON system (1d.f:3529-3543)
// and of the warm-start base B -- B's residual and the three diagonals, each
// scaled per system -- to a correction of B; the multigrid solve then starts
// from the corrected guess.
// Layers: NET_LAYERS 1-D convolutions of NET_TAPS taps over the cell index,
// NET_C channels (features padded to NET_C), ReLU between layers, one output
// channel.  Weights and activations are stored in BF16, every product is
// accumulated in FP32 by v_mfma_f32_16x16x32_bf16, master weights and the
// optimiser are FP32.
//
// Activation layout: channels-last, [cell][NET_C] BF16 per sequence, with
// sequences (one POISSON system, or one training window) zero-padded at both
// ends.  Weight layout for the MFMA B operand: [tap][co][ci] BF16.
//
// Operand layouts of v_mfma_f32_16x16x32_bf16 (wave of 64 lanes, l = lane):
//   A (16 x 32, rows x k): lane l holds row l%16, k = 8*(l/16) .. +7
//   B (32 x 16, k x cols): lane l holds col l%16, k = 8*(l/16) .. +7
//   D (16 x 16 FP32):      lane l holds col l%16, rows 4*(l/16) .. +3
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <stdint.h>
#include "lica_launch.h"

#define NET_C 32
#define NET_TAPS 21
#define NET_HALF (NET_TAPS / 2)
#define NET_LAYERS 4
#define NET_TILE 256         // cells per workgroup in the convolution kernels
#define NET_WIN 256          // cells around the interface the warm-start network corrects
#define NET_NF 4             // input features per cell
static_assert(NET_WIN <= NET_TILE, "a case's window is one tile of the per-case kernel");

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ __bf16 to_bf16(float v) { return (__bf16)v; }
__device__ __forceinline__ float from_bf16(__bf16 v) { return (float)v; }

struct NetSeq {
    long long off;   // first cell of the sequence in the activation arena
    int len;         // cells
    int tile0;       // first tile of the sequence in the tile list
};

// ---------------------------------------------------------------------
// The FP32 sums of one layer over a tile: xs holds the tile's NET_TILE input
// rows with NET_HALF halo rows on either side, ws the layer's weights;
// wave w gets rows 64w..64w+63 as four 16-row blocks (acc[rb][nt]: block
// rb, output channels 16nt..16nt+15).
// Taps t0 .. t1-1 only, with ws holding those taps (ws[t - t0]) and acc
// carried over from the previous taps when t0 > 0: the sums accumulate in
// the same order whatever the grouping.
template <int COUT>
__device__ __forceinline__ void net_tile_mfma(const __bf16 *xs, const __bf16 *ws, f32x4 (&acc)[4][COUT / 16],
                                              int t0 = 0, int t1 = NET_TAPS)
{
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int m = lane & 15, kq = lane >> 4;                  // operand row / k group
    constexpr int NT = COUT / 16;
    if (t0 == 0) {
#pragma unroll
        for (int rb = 0; rb < 4; rb++)
#pragma unroll
            for (int nt = 0; nt < NT; nt++) acc[rb][nt] = (f32x4){0.0f, 0.0f, 0.0f, 0.0f};
    }
    const int r0 = wave * 64;                                 // first tile row of this wave
    for (int t = t0; t < t1; t++) {
        bf16x8 b[NT];
#pragma unroll
        for (int nt = 0; nt < NT; nt++) b[nt] = *(const bf16x8 *)(ws + ((t - t0) * COUT + nt * 16 + m) * NET_C + kq * 8);
#pragma unroll
        for (int rb = 0; rb < 4; rb++) {
            const bf16x8 a = *(const bf16x8 *)(xs + (r0 + rb * 16 + m + t) * NET_C + kq * 8);
#pragma unroll
            for (int nt = 0; nt < NT; nt++) acc[rb][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[nt], acc[rb][nt], 0, 0, 0);
        }
    }
}

// ---------------------------------------------------------------------
// Y[cell][co] = act(bias[co] + sum_t sum_ci W[t][co][ci] X[cell + t - HALF][ci])
// grid: one workgroup per tile of NET_TILE cells (tile -> sequence via
// seq_of_tile), 256 threads: wave w computes rows 64w..64w+63 of the tile as
// four 16-row blocks, so one staging of the weights serves 256 cells.
// COUT is 32, or 16 for the last layer (one meaningful channel); relu 0/1;
// yf (optional) receives the FP32 pre-activation of channel 0.
template <int COUT>
__global__ void __launch_bounds__(256) k_net_conv(const __bf16 *__restrict__ X, const __bf16 *__restrict__ W,
                                                  const float *__restrict__ bias, __bf16 *__restrict__ Y,
                                                  float *__restrict__ yf, const NetSeq *__restrict__ seqs,
                                                  const int *__restrict__ seq_of_tile, int relu)
{
    __shared__ __attribute__((aligned(16))) __bf16 xs[(NET_TILE + NET_TAPS) * NET_C];
    __shared__ __attribute__((aligned(16))) __bf16 ws[NET_TAPS * COUT * NET_C];
    const int tile = blockIdx.x;
    const NetSeq sq = seqs[seq_of_tile[tile]];
    const int c0 = (tile - sq.tile0) * NET_TILE;      // first cell of the tile in the sequence
    if (c0 >= sq.len) return;                         // nothing of the sequence in this tile
    const int tid = threadIdx.x;
    // stage the input rows c0-HALF .. c0+TILE+HALF-1 (zero outside the sequence)
    for (int i = tid; i < (NET_TILE + NET_TAPS - 1) * NET_C / 8; i += blockDim.x) {
        const int row = i / (NET_C / 8), part = i % (NET_C / 8);
        const int cell = c0 - NET_HALF + row;
        bf16x8 v;
        if (cell >= 0 && cell < sq.len) v = *(const bf16x8 *)(X + (sq.off + cell) * NET_C + part * 8);
        else for (int j = 0; j < 8; j++) v[j] = to_bf16(0.0f);
        *(bf16x8 *)(xs + row * NET_C + part * 8) = v;
    }
    for (int i = tid; i < NET_TAPS * COUT * NET_C / 8; i += blockDim.x)
        *(bf16x8 *)(ws + i * 8) = *(const bf16x8 *)(W + i * 8);
    __syncthreads();
    const int wave = tid >> 6, lane = tid & 63;
    const int m = lane & 15, kq = lane >> 4;                  // operand row / k group
    constexpr int NT = COUT / 16;
    f32x4 acc[4][NT];
    net_tile_mfma<COUT>(xs, ws, acc);
    const int r0 = wave * 64;                                 // first tile row of this wave
    // epilogue: lane holds col (l%16), rows 4*(l/16)+i of each block
#pragma unroll
    for (int rb = 0; rb < 4; rb++)
#pragma unroll
        for (int nt = 0; nt < NT; nt++) {
            const int co = nt * 16 + m;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int cell = c0 + r0 + rb * 16 + kq * 4 + i;
                if (cell >= sq.len) continue;
                float v = acc[rb][nt][i] + bias[co];
                if (yf && co == 0) yf[sq.off + cell] = v;
                if (relu) v = fmaxf(v, 0.0f);
                Y[(sq.off + cell) * NET_C + co] = to_bf16(v);
            }
        }
    if (COUT == 16) {   // zero the padding channels 16..31 of the output rows
        for (int i = tid; i < NET_TILE * 16; i += blockDim.x) {
            const int row = i / 16, ch = 16 + i % 16;
            const int cell = c0 + row;
            if (cell < sq.len) Y[(sq.off + cell) * NET_C + ch] = to_bf16(0.0f);
        }
    }
}

// ---------------------------------------------------------------------
// dW[t][co][ci] += sum_cells dZ[cell][co] X[cell + t - HALF][ci] and
// db[co] += sum_cells dZ[cell][co] over the tile's cells (FP32 atomics).
// A = dZ^T (co x cells), B = X shifted (cells x ci): K runs over 32 cells.
__global__ void __launch_bounds__(256) k_net_wgrad(const __bf16 *__restrict__ X, const __bf16 *__restrict__ dZ,
                                                   float *__restrict__ dW, float *__restrict__ db,
                                                   const NetSeq *__restrict__ seqs,
                                                   const int *__restrict__ seq_of_tile, int cout)
{
    __shared__ __attribute__((aligned(16))) __bf16 xs[(NET_TILE + NET_TAPS) * NET_C];
    __shared__ __attribute__((aligned(16))) __bf16 zt[NET_C * NET_TILE];   // dZ transposed: [co][cell]
    const int tile = blockIdx.x;
    const NetSeq sq = seqs[seq_of_tile[tile]];
    const int c0 = (tile - sq.tile0) * NET_TILE;
    const int tid = threadIdx.x;
    for (int i = tid; i < (NET_TILE + NET_TAPS - 1) * NET_C; i += blockDim.x) {
        const int row = i / NET_C, ch = i % NET_C;
        const int cell = c0 - NET_HALF + row;
        xs[row * NET_C + ch] = (cell >= 0 && cell < sq.len) ? X[(sq.off + cell) * NET_C + ch] : to_bf16(0.0f);
    }
    for (int i = tid; i < NET_C * NET_TILE; i += blockDim.x) {
        const int co = i / NET_TILE, r = i % NET_TILE;
        const int cell = c0 + r;
        zt[co * NET_TILE + r] = (cell < sq.len && co < cout) ? dZ[(sq.off + cell) * NET_C + co] : to_bf16(0.0f);
    }
    __syncthreads();
    const int wave = tid >> 6, lane = tid & 63;
    const int m = lane & 15, kq = lane >> 4;
    const int mt_n = cout / 16;
    // taps split over the waves; 2 x 2 (or 1 x 2) output tiles per tap
    for (int t = wave; t < NET_TAPS; t += 4) {
        for (int mt = 0; mt < mt_n; mt++) {
            for (int nt = 0; nt < 2; nt++) {
                f32x4 acc = (f32x4){0.0f, 0.0f, 0.0f, 0.0f};
                for (int kb = 0; kb < NET_TILE; kb += 32) {
                    const bf16x8 a = *(const bf16x8 *)(zt + (mt * 16 + m) * NET_TILE + kb + kq * 8);
                    bf16x8 b;
                    for (int j = 0; j < 8; j++) b[j] = xs[(kb + kq * 8 + j + t) * NET_C + nt * 16 + m];
                    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
                }
                for (int i = 0; i < 4; i++) {
                    const int co = mt * 16 + kq * 4 + i, ci = nt * 16 + m;
                    atomicAdd(&dW[(t * cout + co) * NET_C + ci], acc[i]);
                }
            }
        }
    }
    if (tid < cout) {
        float s = 0.0f;
        for (int r = 0; r < NET_TILE; r++) s += from_bf16(zt[tid * NET_TILE + r]);
        atomicAdd(&db[tid], s);
    }
}

// dX = conv(dZ, flipped transposed W) is k_net_conv with these weights:
// Wt[t'][ci][co] = W[HALF*2 - t'][co][ci] (cin of the transposed layer = cout).
__global__ void k_net_flip(const float *__restrict__ W, __bf16 *__restrict__ Wt, int cout, int cin)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NET_TAPS * NET_C * NET_C) return;
    const int t = i / (NET_C * NET_C), ci = (i / NET_C) % NET_C, co = i % NET_C;   // output [t][ci][co]
    float v = 0.0f;
    if (co < cout && ci < cin) v = W[((NET_TAPS - 1 - t) * cout + co) * NET_C + ci];
    Wt[i] = to_bf16(v);
}

// BF16 copy of FP32 weights ([t][co][ci], n elements)
__global__ void k_net_cast(const float *__restrict__ w, __bf16 *__restrict__ wb, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) wb[i] = to_bf16(w[i]);
}

// the same for every layer in one launch: element i of the concatenation
// belongs to the first layer l with i < end[l]
struct NetCast {
    const float *src[NET_LAYERS];
    __bf16 *dst[NET_LAYERS];
    long long end[NET_LAYERS];
};
__global__ void k_net_cast_all(NetCast c)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long b = 0;
    for (int l = 0; l < NET_LAYERS; l++) {
        if (i < c.end[l]) { c.dst[l][i - b] = to_bf16(c.src[l][i - b]); return; }
        b = c.end[l];
    }
}

// dZ_prev = dX * (Yprev > 0)   (ReLU of the previous layer), channels-last
__global__ void k_net_relu_mask(__bf16 *__restrict__ dX, const __bf16 *__restrict__ Yprev, long long n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && !(from_bf16(Yprev[i]) > 0.0f)) dX[i] = to_bf16(0.0f);
}

// loss = mean (y - target)^2 over cells; dZ_last[cell][0] = 2 (y - target) / count
__global__ void k_net_loss(const float *__restrict__ y, const float *__restrict__ target, __bf16 *__restrict__ dZ,
                           long long ncell, float inv_count, float *__restrict__ loss)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.0f;
    if (i < ncell) {
        const float d = y[i] - target[i];
        l = d * d * inv_count;
        dZ[i * NET_C] = to_bf16(2.0f * d * inv_count);
        for (int c = 1; c < NET_C; c++) dZ[i * NET_C + c] = to_bf16(0.0f);
    }
    for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(loss, l);
}

// minibatch gather: window idx[s] of the data set ([cell][NET_NF]) into row
// block s of the activations ([cell][NET_C], channels beyond NET_NF zero)
__global__ void k_net_gather(const __bf16 *__restrict__ dx, const float *__restrict__ dt,
                             const long long *__restrict__ idx, __bf16 *__restrict__ x, float *__restrict__ t, int len,
                             int batch)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)batch * len) return;
    const int s = (int)(i / len), c = (int)(i % len);
    const long long src = idx[s] * len + c;
    for (int ch = 0; ch < NET_C; ch++) x[i * NET_C + ch] = (ch < NET_NF) ? dx[src * NET_NF + ch] : to_bf16(0.0f);
    t[i] = dt[src];
}

// Adam on FP32 master weights
__global__ void k_net_adam(float *__restrict__ w, float *__restrict__ m, float *__restrict__ v,
                           const float *__restrict__ g, int n, float lr, float b1, float b2, float eps, float c1,
                           float c2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// ---------------------------------------------------------------------
// Features of a POISSON system and a starting point B (the warm-start base):
// with r = g - A B the base's residual (rows 1..n of the POISSON operator:
// a(1) and c(n) do not couple), s_r = max|r| and s_b = max|b| (1 where 0),
// the features are r/s_r, a/s_b, b/s_b, c/s_b.  The network's output y over
// a window of rows around the interface -- where the error left by the
// projection is largest, as a and c jump with the density -- is a direction
// d; its length comes from the residual: guess = B + alpha d with
// alpha = <A d, r> / <A d, A d>, so a useless direction costs nothing.
// Training targets are (p - B) scaled to unit maximum over the window.

__device__ __forceinline__ double net_res(const double *a, const double *b, const double *c, const double *g,
                                          const double *x, int n, int k)
{
    double v = b[k] * x[k];
    if (k > 1) v = v + a[k] * x[k - 1];
    if (k < n) v = v + c[k] * x[k + 1];
    return g[k] - v;
}

// block reduction of NV maxima or sums (256-thread workgroup, all threads
// get the result)
template <int NV, bool SUM>
__device__ void net_reduce(double *v)
{
    __shared__ double red[NV][8];
    const int tid = threadIdx.x;
    for (int off = 32; off > 0; off >>= 1)
        for (int q = 0; q < NV; q++) {
            const double o = __shfl_xor(v[q], off, 64);
            v[q] = SUM ? v[q] + o : fmax(v[q], o);
        }
    if ((tid & 63) == 0)
        for (int q = 0; q < NV; q++) red[q][tid >> 6] = v[q];
    __syncthreads();
    for (int q = 0; q < NV; q++) {
        double r = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) r = SUM ? r + red[q][w] : fmax(r, red[q][w]);
        v[q] = r;
    }
    __syncthreads();
}

// s_r, s_b over rows 1..n; with want_berr also the backward error of x,
// max|r| / (max(|a|+|b|+|c|) max|x| + max|g|)
__device__ void net_scales(const double *a, const double *b, const double *c, const double *g, const double *x, int n,
                           double *sr_out, double *sb_out, double *berr_out)
{
    double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};   // max|r|, max|b|, max(|a|+|b|+|c|), max|x|, max|g|
    for (int k = 1 + (int)threadIdx.x; k <= n; k += blockDim.x) {
        v[0] = fmax(v[0], fabs(net_res(a, b, c, g, x, n, k)));
        v[1] = fmax(v[1], fabs(b[k]));
        const double ak = (k > 1) ? fabs(a[k]) : 0.0, ck = (k < n) ? fabs(c[k]) : 0.0;
        v[2] = fmax(v[2], ak + fabs(b[k]) + ck);
        v[3] = fmax(v[3], fabs(x[k]));
        v[4] = fmax(v[4], fabs(g[k]));
    }
    net_reduce<5, false>(v);
    const double den = v[2] * v[3] + v[4];
    if (berr_out) *berr_out = (den > 0.0) ? v[0] / den : v[0];
    *sr_out = (v[0] == 0.0) ? 1.0 : v[0];
    *sb_out = (v[1] == 0.0) ? 1.0 : v[1];
}

__device__ __forceinline__ void net_row(const double *a, const double *b, const double *c, const double *g,
                                        const double *x, int n, int k, double sr, double sb, __bf16 *row)
{
    row[0] = to_bf16((float)(net_res(a, b, c, g, x, n, k) / sr));
    row[1] = to_bf16((float)(a[k] / sb));
    row[2] = to_bf16((float)(b[k] / sb));
    row[3] = to_bf16((float)(c[k] / sb));
}

// first row of a case's window: cs.win rows centred on the interface
__device__ __forceinline__ int net_k0(const NetCase &cs)
{
    const int kif = (int)(*cs.place * cs.dzi) + 1;
    return min(max(kif - cs.win / 2, 1), cs.n - cs.win + 1);
}

// training windows from captured systems (one workgroup per case): window w
// of case i starts at row 1 + woff[i*wpc + w] (-1: none) and goes to data-set
// slot slot[i*wpc + w]; features as above ([cell][NET_NF] BF16), target
// (p1 - p0) / max over the window of |p1 - p0| (FP32; 0 where p1 = p0)
__global__ void __launch_bounds__(256) k_net_pairs(const NetPairSrc *__restrict__ src, const int *__restrict__ woff,
                                                   const long long *__restrict__ slot, int wpc, int len,
                                                   __bf16 *__restrict__ dx, float *__restrict__ dt)
{
    const NetPairSrc s = src[blockIdx.x];
    double sr, sb;
    net_scales(s.a, s.b, s.c, s.g, s.p0, s.n, &sr, &sb, nullptr);
    for (int w = 0; w < wpc; w++) {
        const int o = woff[blockIdx.x * wpc + w];
        if (o < 0) continue;
        double em[1] = {0.0};
        for (int c = threadIdx.x; c < len; c += blockDim.x) em[0] = fmax(em[0], fabs(s.p1[1 + o + c] - s.p0[1 + o + c]));
        net_reduce<1, false>(em);
        const double se = (em[0] > 0.0) ? em[0] : 1.0;
        const long long base = slot[blockIdx.x * wpc + w] * len;
        for (int c = threadIdx.x; c < len; c += blockDim.x) {
            const int k = 1 + o + c;
            net_row(s.a, s.b, s.c, s.g, s.p0, s.n, k, sr, sb, dx + (base + c) * NET_NF);
            dt[base + c] = (float)((s.p1[k] - s.p0[k]) / se);
        }
    }
}

// parameters of the layers: BF16 weights (cast from the FP32 masters by
// k_net_cast) and FP32 biases
struct NetParams {
    const __bf16 *w[NET_LAYERS];
    const float *b[NET_LAYERS];
};

// One layer of the per-case network in LDS: the layer's BF16 weights pass
// through ws NET_TGRP taps at a time, the FP32 sums over the tile stay in
// registers, then bias (+ ReLU) goes back into the same activation rows
// (BF16; the halo and the rows beyond the window stay zero) or, for the
// last layer, channel 0's FP32 pre-activation into ys.
// The weights of the next tap group (the next layer's first group after a
// layer's last, when next is given) are loaded into registers (pre) while
// the current group's products run, and stored into ws after them.
#define NET_TGRP 7
#define NET_CASE_THREADS 256
static_assert(NET_TAPS % NET_TGRP == 0, "tap groups");
#define NET_PRE ((NET_TGRP * NET_C * NET_C / 8 + NET_CASE_THREADS - 1) / NET_CASE_THREADS)
struct NetPre {
    bf16x8 v[NET_PRE];
};
template <int COUT>
__device__ __forceinline__ void net_group_load(NetPre &pre, const __bf16 *__restrict__ src)
{
    constexpr int UNITS = NET_TGRP * COUT * NET_C / 8;
#pragma unroll
    for (int j = 0; j < NET_PRE; j++) {
        const int i = (int)threadIdx.x + j * NET_CASE_THREADS;
        if (i < UNITS) pre.v[j] = *(const bf16x8 *)(src + i * 8);
    }
}
template <int COUT>
__device__ __forceinline__ void net_group_store(const NetPre &pre, __bf16 *ws)
{
    constexpr int UNITS = NET_TGRP * COUT * NET_C / 8;
#pragma unroll
    for (int j = 0; j < NET_PRE; j++) {
        const int i = (int)threadIdx.x + j * NET_CASE_THREADS;
        if (i < UNITS) *(bf16x8 *)(ws + i * 8) = pre.v[j];
    }
}
template <int COUT, int NEXT_COUT>
__device__ __forceinline__ void net_case_layer(__bf16 *xs, __bf16 *ws, float *ys, const __bf16 *__restrict__ wl,
                                               const float *__restrict__ bl, int w, bool last, NetPre &pre,
                                               const __bf16 *__restrict__ next)
{
    const int tid = threadIdx.x;
    constexpr int NT = COUT / 16;
    f32x4 acc[4][NT];
    for (int t0 = 0; t0 < NET_TAPS; t0 += NET_TGRP) {
        net_group_store<COUT>(pre, ws);
        __syncthreads();
        if (t0 + NET_TGRP < NET_TAPS) net_group_load<COUT>(pre, wl + (size_t)(t0 + NET_TGRP) * COUT * NET_C);
        else if (next) net_group_load<NEXT_COUT>(pre, next);
        net_tile_mfma<COUT>(xs, ws, acc, t0, t0 + NET_TGRP);
        __syncthreads();   // ws (and, after the last group, xs) free
    }
    const int wave = tid >> 6, lane = tid & 63, m = lane & 15, kq = lane >> 4, r0 = wave * 64;
#pragma unroll
    for (int rb = 0; rb < 4; rb++)
#pragma unroll
        for (int nt = 0; nt < NT; nt++) {
            const int co = nt * 16 + m;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int cell = r0 + rb * 16 + kq * 4 + i;
                if (cell >= w) continue;
                const float v = acc[rb][nt][i] + bl[co];
                if (last) {
                    if (co == 0) ys[cell] = v;
                } else {
                    xs[(NET_HALF + cell) * NET_C + co] = to_bf16(fmaxf(v, 0.0f));
                }
            }
        }
    __syncthreads();   // the next layer's input / ys complete
}

// The network's correction of a batch's warm-start guesses, one workgroup
// per case (the window, <= NET_TILE rows, is one tile): the features of the
// window's rows from B (scales in info, written by the warm-start base pass),
// the NET_LAYERS layers with activations and weights in LDS, then
// guess = B + alpha d over the window with alpha = <A d, r> / <A d, A d> (the
// rows where A d is nonzero).  A case whose base already meets rtol keeps B.
// The arithmetic is k_net_conv's layer by layer (BF16 activations and
// weights, FP32 sums in the same order).  info[LICA_WARM_INFO i]: window
// start (written), backward error, s_r, s_b (read), alpha (written), the
// multigrid start verdict (cleared when the guess moves).
__global__ void __launch_bounds__(NET_CASE_THREADS) __attribute__((amdgpu_waves_per_eu(4))) k_net_case(const NetCase *__restrict__ cases, const int *__restrict__ order,
                                                  NetParams prm, double *__restrict__ info, double rtol)
{
    const int ci = order ? order[blockIdx.x] : (int)blockIdx.x;   // the case
    constexpr int ROWS = NET_TILE + NET_TAPS - 1;
    __shared__ __attribute__((aligned(16))) __bf16 xs[ROWS * NET_C];
    __shared__ __attribute__((aligned(16))) __bf16 ws[NET_TGRP * NET_C * NET_C];
    __shared__ float ys[NET_TILE];
    const NetCase cs = cases[ci];
    const int tid = threadIdx.x;
    double *inf = info + (size_t)LICA_WARM_INFO * ci;
    const double berr = inf[1], sr = inf[2], sb = inf[3];
    const int k0 = net_k0(cs), n = cs.n, w = cs.win;
    if (tid == 0) { inf[0] = (double)k0; inf[4] = 0.0; }
    if (!(berr > rtol)) return;
    if (tid == 0) inf[6] = 0.0;   // the guess moves: the solve tests it afresh
    NetPre pre;
    net_group_load<NET_C>(pre, prm.w[0]);   // in flight while the features are built
    // layer 0 input (features, channels NET_NF.. zero), zero rows outside
    // the window
    for (int i = tid; i < ROWS; i += blockDim.x) {
        const int c = i - NET_HALF;
        __bf16 *row = xs + i * NET_C;
        int ch0 = 0;
        if (c >= 0 && c < w) {
            net_row(cs.a, cs.b, cs.c, cs.g, cs.guess, n, k0 + c, sr, sb, row);
            ch0 = NET_NF;
        }
        for (int ch = ch0; ch < NET_C; ch++) row[ch] = to_bf16(0.0f);
    }
    __syncthreads();
    static_assert(NET_LAYERS == 4, "the layer sequence below");
    net_case_layer<NET_C, NET_C>(xs, ws, ys, prm.w[0], prm.b[0], w, false, pre, prm.w[1]);
    net_case_layer<NET_C, NET_C>(xs, ws, ys, prm.w[1], prm.b[1], w, false, pre, prm.w[2]);
    net_case_layer<NET_C, 16>(xs, ws, ys, prm.w[2], prm.b[2], w, false, pre, prm.w[3]);
    net_case_layer<16, 16>(xs, ws, ys, prm.w[3], prm.b[3], w, true, pre, nullptr);
    auto d = [&](int k) -> double { return (k >= k0 && k < k0 + w) ? (double)ys[k - k0] : 0.0; };
    double v[2] = {0.0, 0.0};
    for (int k = max(k0 - 1, 1) + tid; k <= min(k0 + w, n); k += blockDim.x) {
        double ad = cs.b[k] * d(k);
        if (k > 1) ad = ad + cs.a[k] * d(k - 1);
        if (k < n) ad = ad + cs.c[k] * d(k + 1);
        const double r = net_res(cs.a, cs.b, cs.c, cs.g, cs.guess, n, k);
        v[0] += ad * r;
        v[1] += ad * ad;
    }
    net_reduce<2, true>(v);
    const double alpha = (v[1] > 0.0) ? v[0] / v[1] : 0.0;
    for (int c = tid; c < w; c += blockDim.x) cs.guess[k0 + c] = cs.guess[k0 + c] + alpha * (double)ys[c];
    if (tid == 0) inf[4] = alpha;
}

// ---------------------------------------------------------------------
// host side of the network
#include "../../include/lica_mi355x.h"
#include <vector>
#include <random>
#include <cstring>
#include <cmath>

namespace {
const int NET_COUT[NET_LAYERS] = {NET_C, NET_C, NET_C, 16};
const int NET_CIN_REAL[NET_LAYERS] = {4, NET_C, NET_C, NET_C};

size_t wsize(int l) { return (size_t)NET_TAPS * NET_COUT[l] * NET_C; }

int hip_ok(hipError_t e) { return e == hipSuccess ? 0 : LICA_HIP_ERROR; }
}  // namespace

struct lica_net {
    float *w[NET_LAYERS] = {}, *b[NET_LAYERS] = {};                 // FP32 master
    float *gw[NET_LAYERS] = {}, *gb[NET_LAYERS] = {};               // gradients
    float *mw[NET_LAYERS] = {}, *vw[NET_LAYERS] = {}, *mb[NET_LAYERS] = {}, *vb[NET_LAYERS] = {};
    __bf16 *wb[NET_LAYERS] = {}, *wt[NET_LAYERS] = {};              // BF16 forward / flipped
    float *zero_bias = nullptr;
    long long adam_t = 0;
    hipStream_t stream = nullptr;
    // activation arena
    long long cap_cells = 0;
    int cap_tiles = 0, cap_seqs = 0;
    __bf16 *act[NET_LAYERS + 1] = {};   // act[0] features, act[l+1] output of layer l
    __bf16 *dz[2] = {};
    float *yf = nullptr, *target = nullptr, *loss = nullptr;
    NetSeq *seqs = nullptr;
    int *seq_of_tile = nullptr;
    int nseq = 0, ntiles = 0;
    long long ncell = 0;
    // training set (device resident, fixed-length windows)
    __bf16 *data_x = nullptr;           // [window][cell][NET_NF]
    float *data_t = nullptr;            // [window][cell]
    long long *idx = nullptr;
    long long data_n = 0, data_cap = 0; // windows held / room for
    int data_len = 0;
};

namespace {
// zeroed device memory; the fill (null stream, asynchronous to the host) is
// waited for, as the network's streams are not ordered after the null stream
int net_alloc(void **p, size_t bytes)
{
    if (hipMalloc(p, bytes ? bytes : 1) != hipSuccess) { *p = nullptr; return LICA_HIP_ERROR; }
    int rc = hip_ok(hipMemset(*p, 0, bytes ? bytes : 1));
    return rc ? rc : hip_ok(hipStreamSynchronize(nullptr));
}

// room for ncell cells in nseq sequences
int net_reserve(lica_net *t, long long ncell, int nseq, int ntiles)
{
    if (ncell <= t->cap_cells && nseq <= t->cap_seqs && ntiles <= t->cap_tiles) return 0;
    long long cells = std::max(ncell, t->cap_cells);
    int seqs = std::max(nseq, t->cap_seqs), tiles = std::max(ntiles, t->cap_tiles);
    for (int l = 0; l <= NET_LAYERS; l++) { (void)hipFree(t->act[l]); t->act[l] = nullptr; }
    for (int i = 0; i < 2; i++) { (void)hipFree(t->dz[i]); t->dz[i] = nullptr; }
    (void)hipFree(t->yf); (void)hipFree(t->target); (void)hipFree(t->seqs); (void)hipFree(t->seq_of_tile);
    int rc = 0;
    const size_t ab = sizeof(__bf16) * (size_t)cells * NET_C;
    for (int l = 0; l <= NET_LAYERS && !rc; l++) rc = net_alloc((void **)&t->act[l], ab);
    for (int i = 0; i < 2 && !rc; i++) rc = net_alloc((void **)&t->dz[i], ab);
    rc = rc ? rc : net_alloc((void **)&t->yf, sizeof(float) * (size_t)cells);
    rc = rc ? rc : net_alloc((void **)&t->target, sizeof(float) * (size_t)cells);
    rc = rc ? rc : net_alloc((void **)&t->seqs, sizeof(NetSeq) * (size_t)seqs);
    rc = rc ? rc : net_alloc((void **)&t->seq_of_tile, sizeof(int) * (size_t)tiles);
    if (rc) return rc;
    t->cap_cells = cells; t->cap_seqs = seqs; t->cap_tiles = tiles;
    return 0;
}

// sequence and tile tables for the given lengths (cells contiguous in order)
int net_layout(lica_net *t, const int64_t *lens, int nseq)
{
    std::vector<NetSeq> sq((size_t)nseq);
    std::vector<int> sot;
    long long off = 0;
    for (int s = 0; s < nseq; s++) {
        sq[(size_t)s].off = off;
        sq[(size_t)s].len = (int)lens[s];
        sq[(size_t)s].tile0 = (int)sot.size();
        const int nt = (int)((lens[s] + NET_TILE - 1) / NET_TILE);
        for (int i = 0; i < nt; i++) sot.push_back(s);
        off += lens[s];
    }
    int rc = net_reserve(t, off, nseq, (int)sot.size());
    if (rc) return rc;
    rc = hip_ok(hipMemcpyAsync(t->seqs, sq.data(), sizeof(NetSeq) * sq.size(), hipMemcpyHostToDevice, t->stream));
    rc = rc ? rc : hip_ok(hipMemcpyAsync(t->seq_of_tile, sot.data(), sizeof(int) * sot.size(), hipMemcpyHostToDevice, t->stream));
    t->nseq = nseq;
    t->ntiles = (int)sot.size();
    t->ncell = off;
    return rc ? rc : hip_ok(hipStreamSynchronize(t->stream));
}

int net_cast_weights(lica_net *t)
{
    for (int l = 0; l < NET_LAYERS; l++) {
        const int n = (int)wsize(l);
        hipLaunchKernelGGL(k_net_cast, dim3((n + 255) / 256), dim3(256), 0, t->stream, t->w[l], t->wb[l], n);
    }
    return hip_ok(hipGetLastError());
}

// forward over the current layout; act[0] must hold the features
int net_forward(lica_net *t)
{
    int rc = net_cast_weights(t);
    for (int l = 0; l < NET_LAYERS && !rc; l++) {
        const bool last = (l == NET_LAYERS - 1);
        if (last)
            hipLaunchKernelGGL(k_net_conv<16>, dim3(t->ntiles), dim3(256), 0, t->stream, t->act[l], t->wb[l], t->b[l],
                               t->act[l + 1], t->yf, t->seqs, t->seq_of_tile, 0);
        else
            hipLaunchKernelGGL(k_net_conv<NET_C>, dim3(t->ntiles), dim3(256), 0, t->stream, t->act[l], t->wb[l],
                               t->b[l], t->act[l + 1], (float *)nullptr, t->seqs, t->seq_of_tile, 1);
        rc = hip_ok(hipGetLastError());
    }
    return rc;
}

// loss and gradients of mean (y - target)^2 for the current layout
int net_backward(lica_net *t, double *loss_out)
{
    int rc = hip_ok(hipMemsetAsync(t->loss, 0, sizeof(float), t->stream));
    for (int l = 0; l < NET_LAYERS && !rc; l++) {
        rc = hip_ok(hipMemsetAsync(t->gw[l], 0, sizeof(float) * wsize(l), t->stream));
        rc = rc ? rc : hip_ok(hipMemsetAsync(t->gb[l], 0, sizeof(float) * NET_COUT[l], t->stream));
    }
    if (rc) return rc;
    const long long nc = t->ncell;
    hipLaunchKernelGGL(k_net_loss, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, t->stream, t->yf, t->target,
                       t->dz[0], nc, (float)(1.0 / (double)nc), t->loss);
    int cur = 0;
    for (int l = NET_LAYERS - 1; l >= 0 && !rc; l--) {
        hipLaunchKernelGGL(k_net_wgrad, dim3(t->ntiles), dim3(256), 0, t->stream, t->act[l], t->dz[cur], t->gw[l],
                           t->gb[l], t->seqs, t->seq_of_tile, NET_COUT[l]);
        if (l > 0) {
            const int n = NET_TAPS * NET_C * NET_C;
            hipLaunchKernelGGL(k_net_flip, dim3((n + 255) / 256), dim3(256), 0, t->stream, t->w[l], t->wt[l],
                               NET_COUT[l], NET_C);
            hipLaunchKernelGGL(k_net_conv<NET_C>, dim3(t->ntiles), dim3(256), 0, t->stream, t->dz[cur], t->wt[l],
                               t->zero_bias, t->dz[1 - cur], (float *)nullptr, t->seqs, t->seq_of_tile, 0);
            const long long ne = nc * NET_C;
            hipLaunchKernelGGL(k_net_relu_mask, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, t->stream,
                               t->dz[1 - cur], t->act[l], ne);
            cur = 1 - cur;
        }
        rc = hip_ok(hipGetLastError());
    }
    if (!rc && loss_out) {
        float l = 0.0f;
        rc = hip_ok(hipMemcpyAsync(&l, t->loss, sizeof(float), hipMemcpyDeviceToHost, t->stream));
        rc = rc ? rc : hip_ok(hipStreamSynchronize(t->stream));
        *loss_out = l;
    }
    return rc;
}

// host features [cells][4] (FP32) -> act[0] (BF16, 32 channels)
int net_upload_features(lica_net *t, const float *feat, long long ncell)
{
    std::vector<__bf16> x((size_t)ncell * NET_C);
    for (long long i = 0; i < ncell; i++)
        for (int c = 0; c < NET_C; c++) x[(size_t)i * NET_C + c] = (__bf16)(c < 4 ? feat[i * 4 + c] : 0.0f);
    int rc = hip_ok(hipMemcpyAsync(t->act[0], x.data(), sizeof(__bf16) * x.size(), hipMemcpyHostToDevice, t->stream));
    return rc ? rc : hip_ok(hipStreamSynchronize(t->stream));
}
// room for n windows of t->data_len cells in the data set (keeps the windows held)
int net_data_reserve(lica_net *t, long long n)
{
    if (n <= t->data_cap) return 0;
    const long long cap = std::max(n, 2 * t->data_cap);
    const size_t cells = (size_t)cap * (size_t)t->data_len;
    __bf16 *x = nullptr;
    float *tg = nullptr;
    int rc = net_alloc((void **)&x, sizeof(__bf16) * cells * NET_NF);
    rc = rc ? rc : net_alloc((void **)&tg, sizeof(float) * cells);
    const size_t held = (size_t)t->data_n * (size_t)t->data_len;
    if (!rc && held) {
        rc = hip_ok(hipMemcpy(x, t->data_x, sizeof(__bf16) * held * NET_NF, hipMemcpyDeviceToDevice));
        rc = rc ? rc : hip_ok(hipMemcpy(tg, t->data_t, sizeof(float) * held, hipMemcpyDeviceToDevice));
    }
    if (rc) { (void)hipFree(x); (void)hipFree(tg); return rc; }
    (void)hipFree(t->data_x); (void)hipFree(t->data_t);
    t->data_x = x;
    t->data_t = tg;
    t->data_cap = cap;
    return 0;
}
}  // namespace

extern "C" {

int64_t lica_net_param_count(void)
{
    int64_t n = 0;
    for (int l = 0; l < NET_LAYERS; l++) n += (int64_t)wsize(l) + NET_COUT[l];
    return n;
}

int lica_net_create(uint64_t seed, lica_net **out)
{
    if (!out) return LICA_BAD_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return LICA_NO_DEVICE;
    lica_net *t = new lica_net();
    int rc = hip_ok(hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking));
    std::mt19937_64 rng(seed);
    for (int l = 0; l < NET_LAYERS && !rc; l++) {
        const size_t n = wsize(l);
        rc = net_alloc((void **)&t->w[l], sizeof(float) * n);
        rc = rc ? rc : net_alloc((void **)&t->gw[l], sizeof(float) * n);
        rc = rc ? rc : net_alloc((void **)&t->mw[l], sizeof(float) * n);
        rc = rc ? rc : net_alloc((void **)&t->vw[l], sizeof(float) * n);
        rc = rc ? rc : net_alloc((void **)&t->b[l], sizeof(float) * NET_C);
        rc = rc ? rc : net_alloc((void **)&t->gb[l], sizeof(float) * NET_C);
        rc = rc ? rc : net_alloc((void **)&t->mb[l], sizeof(float) * NET_C);
        rc = rc ? rc : net_alloc((void **)&t->vb[l], sizeof(float) * NET_C);
        rc = rc ? rc : net_alloc((void **)&t->wb[l], sizeof(__bf16) * n);
        rc = rc ? rc : net_alloc((void **)&t->wt[l], sizeof(__bf16) * (size_t)NET_TAPS * NET_C * NET_C);
        if (rc) break;
        // He initialisation over the real fan-in; padded inputs and outputs stay zero
        std::vector<float> w(n, 0.0f);
        const double sd = std::sqrt(2.0 / (double)(NET_TAPS * NET_CIN_REAL[l]));
        std::normal_distribution<double> nd(0.0, sd);
        const int creal = (l == NET_LAYERS - 1) ? 1 : NET_COUT[l];
        for (int tp = 0; tp < NET_TAPS; tp++)
            for (int co = 0; co < creal; co++)
                for (int ci = 0; ci < NET_CIN_REAL[l]; ci++)
                    w[((size_t)tp * NET_COUT[l] + co) * NET_C + ci] = (float)(nd(rng) * (l == NET_LAYERS - 1 ? 0.1 : 1.0));
        rc = hip_ok(hipMemcpy(t->w[l], w.data(), sizeof(float) * n, hipMemcpyHostToDevice));
    }
    rc = rc ? rc : net_alloc((void **)&t->zero_bias, sizeof(float) * NET_C);
    rc = rc ? rc : net_alloc((void **)&t->loss, sizeof(float));
    if (rc) { lica_net_destroy(t); return rc; }
    *out = t;
    return LICA_OK;
}

void lica_net_destroy(lica_net *t)
{
    if (!t) return;
    if (t->stream) (void)hipStreamSynchronize(t->stream);
    for (int l = 0; l < NET_LAYERS; l++) {
        (void)hipFree(t->w[l]); (void)hipFree(t->b[l]); (void)hipFree(t->gw[l]); (void)hipFree(t->gb[l]);
        (void)hipFree(t->mw[l]); (void)hipFree(t->vw[l]); (void)hipFree(t->mb[l]); (void)hipFree(t->vb[l]);
        (void)hipFree(t->wb[l]); (void)hipFree(t->wt[l]);
    }
    for (int l = 0; l <= NET_LAYERS; l++) (void)hipFree(t->act[l]);
    for (int i = 0; i < 2; i++) (void)hipFree(t->dz[i]);
    (void)hipFree(t->yf); (void)hipFree(t->target); (void)hipFree(t->loss); (void)hipFree(t->zero_bias);
    (void)hipFree(t->seqs); (void)hipFree(t->seq_of_tile);
    (void)hipFree(t->data_x); (void)hipFree(t->data_t); (void)hipFree(t->idx);
    if (t->stream) (void)hipStreamDestroy(t->stream);
    delete t;
}

// parameters, layer by layer: W[tap][co][ci] (padded to 32 in/out, 16 out for
// the last layer), then the bias (cout values)
int lica_net_get_params(lica_net *t, float *out)
{
    if (!t || !out) return LICA_BAD_ARG;
    size_t o = 0;
    int rc = 0;
    for (int l = 0; l < NET_LAYERS && !rc; l++) {
        rc = hip_ok(hipMemcpy(out + o, t->w[l], sizeof(float) * wsize(l), hipMemcpyDeviceToHost));
        o += wsize(l);
        rc = rc ? rc : hip_ok(hipMemcpy(out + o, t->b[l], sizeof(float) * NET_COUT[l], hipMemcpyDeviceToHost));
        o += NET_COUT[l];
    }
    return rc;
}

int lica_net_set_params(lica_net *t, const float *in)
{
    if (!t || !in) return LICA_BAD_ARG;
    size_t o = 0;
    int rc = 0;
    for (int l = 0; l < NET_LAYERS && !rc; l++) {
        rc = hip_ok(hipMemcpy(t->w[l], in + o, sizeof(float) * wsize(l), hipMemcpyHostToDevice));
        o += wsize(l);
        rc = rc ? rc : hip_ok(hipMemcpy(t->b[l], in + o, sizeof(float) * NET_COUT[l], hipMemcpyHostToDevice));
        o += NET_COUT[l];
    }
    return rc;
}

// y[cells] = network(features[cells][4]) for nseq sequences of lens[s] cells
int lica_net_forward(lica_net *t, const float *feat, const int64_t *lens, int64_t nseq, float *y)
{
    if (!t || !feat || !lens || nseq <= 0 || !y) return LICA_BAD_ARG;
    int rc = net_layout(t, lens, (int)nseq);
    rc = rc ? rc : net_upload_features(t, feat, t->ncell);
    rc = rc ? rc : net_forward(t);
    rc = rc ? rc : hip_ok(hipMemcpyAsync(y, t->yf, sizeof(float) * (size_t)t->ncell, hipMemcpyDeviceToHost, t->stream));
    return rc ? rc : hip_ok(hipStreamSynchronize(t->stream));
}

// loss = mean (y - target)^2 and its gradient (parameter layout of get_params)
int lica_net_grad(lica_net *t, const float *feat, const float *target, const int64_t *lens, int64_t nseq,
                  float *grad, double *loss)
{
    if (!t || !feat || !target || !lens || nseq <= 0 || !grad) return LICA_BAD_ARG;
    int rc = net_layout(t, lens, (int)nseq);
    rc = rc ? rc : net_upload_features(t, feat, t->ncell);
    rc = rc ? rc : hip_ok(hipMemcpyAsync(t->target, target, sizeof(float) * (size_t)t->ncell, hipMemcpyHostToDevice, t->stream));
    rc = rc ? rc : net_forward(t);
    rc = rc ? rc : net_backward(t, loss);
    size_t o = 0;
    for (int l = 0; l < NET_LAYERS && !rc; l++) {
        rc = hip_ok(hipMemcpy(grad + o, t->gw[l], sizeof(float) * wsize(l), hipMemcpyDeviceToHost));
        o += wsize(l);
        rc = rc ? rc : hip_ok(hipMemcpy(grad + o, t->gb[l], sizeof(float) * NET_COUT[l], hipMemcpyDeviceToHost));
        o += NET_COUT[l];
    }
    return rc;
}

// training windows: n windows of len cells, features [n*len][4], target [n*len]
// (replaces the data set)
int lica_net_set_data(lica_net *t, const float *feat, const float *target, int64_t n, int64_t len)
{
    if (!t || !feat || !target || n <= 0 || len <= 0) return LICA_BAD_ARG;
    t->data_n = 0;
    t->data_len = (int)len;
    int rc = net_data_reserve(t, n);
    if (rc) return rc;
    const size_t cells = (size_t)n * (size_t)len;
    std::vector<__bf16> x(cells * NET_NF);
    for (size_t i = 0; i < cells * NET_NF; i++) x[i] = (__bf16)feat[i];
    rc = hip_ok(hipMemcpy(t->data_x, x.data(), sizeof(__bf16) * x.size(), hipMemcpyHostToDevice));
    rc = rc ? rc : hip_ok(hipMemcpy(t->data_t, target, sizeof(float) * cells, hipMemcpyHostToDevice));
    if (rc) return rc;
    t->data_n = n;
    return LICA_OK;
}

int lica_net_data_size(lica_net *t, int64_t *n, int64_t *len)
{
    if (!t || !n || !len) return LICA_BAD_ARG;
    *n = t->data_n;
    *len = t->data_len;
    return LICA_OK;
}

// the data set back on the host: features [n*len][4] (BF16 values), targets
int lica_net_get_data(lica_net *t, float *feat, float *target)
{
    if (!t || !feat || !target) return LICA_BAD_ARG;
    const size_t cells = (size_t)t->data_n * (size_t)t->data_len;
    if (!cells) return LICA_OK;
    std::vector<__bf16> x(cells * NET_NF);
    int rc = hip_ok(hipMemcpy(x.data(), t->data_x, sizeof(__bf16) * x.size(), hipMemcpyDeviceToHost));
    rc = rc ? rc : hip_ok(hipMemcpy(target, t->data_t, sizeof(float) * cells, hipMemcpyDeviceToHost));
    if (rc) return rc;
    for (size_t i = 0; i < x.size(); i++) feat[i] = (float)x[i];
    return LICA_OK;
}

// Adam steps on random minibatches of `batch` windows; loss_out gets the mean
// minibatch loss of the last 10 % of the steps
int lica_net_train(lica_net *t, int64_t iters, int64_t batch, double lr, uint64_t seed, double *loss_out)
{
    if (!t || !t->data_x || iters < 1 || batch < 1) return LICA_BAD_ARG;
    const int len = t->data_len;
    std::vector<int64_t> lens((size_t)batch, len);
    int rc = net_layout(t, lens.data(), (int)batch);
    if (rc) return rc;
    std::mt19937_64 rng(seed);
    std::uniform_int_distribution<int64_t> pick(0, t->data_n - 1);
    std::vector<long long> hidx((size_t)batch);
    (void)hipFree(t->idx);
    rc = net_alloc((void **)&t->idx, sizeof(long long) * (size_t)batch);
    if (rc) return rc;
    double acc = 0.0;
    int64_t nacc = 0;
    const double b1 = 0.9, b2 = 0.999;
    for (int64_t it = 0; it < iters && !rc; it++) {
        for (int64_t s = 0; s < batch; s++) hidx[(size_t)s] = (long long)pick(rng);
        rc = hip_ok(hipMemcpyAsync(t->idx, hidx.data(), sizeof(long long) * hidx.size(), hipMemcpyHostToDevice, t->stream));
        const long long nc = batch * (long long)len;
        hipLaunchKernelGGL(k_net_gather, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, t->stream, t->data_x,
                           t->data_t, t->idx, t->act[0], t->target, len, (int)batch);
        // the host index buffer is reused next iteration: wait for the copy
        rc = rc ? rc : hip_ok(hipStreamSynchronize(t->stream));
        double l = 0.0;
        rc = rc ? rc : net_forward(t);
        rc = rc ? rc : net_backward(t, (it >= iters - std::max<int64_t>(1, iters / 10)) ? &l : nullptr);
        if (it >= iters - std::max<int64_t>(1, iters / 10)) { acc += l; nacc++; }
        t->adam_t++;
        const float c1 = (float)(1.0 - std::pow(b1, (double)t->adam_t)), c2 = (float)(1.0 - std::pow(b2, (double)t->adam_t));
        for (int ly = 0; ly < NET_LAYERS && !rc; ly++) {
            const int n = (int)wsize(ly);
            hipLaunchKernelGGL(k_net_adam, dim3((n + 255) / 256), dim3(256), 0, t->stream, t->w[ly], t->mw[ly],
                               t->vw[ly], t->gw[ly], n, (float)lr, (float)b1, (float)b2, 1e-8f, c1, c2);
            hipLaunchKernelGGL(k_net_adam, dim3(1), dim3(64), 0, t->stream, t->b[ly], t->mb[ly], t->vb[ly],
                               t->gb[ly], NET_COUT[ly], (float)lr, (float)b1, (float)b2, 1e-8f, c1, c2);
            rc = hip_ok(hipGetLastError());
        }
    }
    rc = rc ? rc : hip_ok(hipStreamSynchronize(t->stream));
    if (loss_out) *loss_out = nacc ? acc / (double)nacc : 0.0;
    return rc;
}

}  // extern "C"

// ---------------------------------------------------------------------
// warm start inside a batch: the network's correction of every case's
// guess, one workgroup per case (called by the batch driver before the
// multigrid kernel of a substep; no allocation, so it can be graph-captured)
int lica_net_guess(lica_net *t, const NetCase *dcases, const int *order, int ncase, double *dinfo, double rtol,
                   void *stream)
{
    NetParams prm;
    for (int l = 0; l < NET_LAYERS; l++) { prm.w[l] = t->wb[l]; prm.b[l] = t->b[l]; }
    hipLaunchKernelGGL(k_net_case, dim3(ncase), dim3(NET_CASE_THREADS), 0, (hipStream_t)stream, dcases, order, prm, dinfo, rtol);
    return hip_ok(hipGetLastError());
}

// the BF16 copies of the weights from the FP32 masters, one launch for all
// layers (the batch driver calls it once per step, before the step's
// launches, so that every set of cases reads the same copies)
int lica_net_cast(lica_net *t, void *stream)
{
    NetCast nc;
    long long tot = 0;
    for (int l = 0; l < NET_LAYERS; l++) { nc.src[l] = t->w[l]; nc.dst[l] = t->wb[l]; nc.end[l] = (tot += (long long)wsize(l)); }
    hipLaunchKernelGGL(k_net_cast_all, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, (hipStream_t)stream, nc);
    return hip_ok(hipGetLastError());
}

int64_t lica_net_window() { return NET_WIN; }

// training windows from a batch's captured systems (see lica_net_add_pairs)
int lica_net_append(lica_net *t, const NetPairSrc *hsrc, int ncase, int64_t wpc, int64_t len, uint64_t seed,
                    int64_t *added)
{
    if (!t || (!hsrc && ncase > 0) || ncase < 0 || wpc < 1 || len < 1 || len > (1 << 30)) return LICA_BAD_ARG;
    if (t->data_n > 0 && len != t->data_len) return LICA_BAD_ARG;
    if (added) *added = 0;
    if (ncase == 0) return LICA_OK;
    std::mt19937_64 rng(seed);
    std::vector<int> woff((size_t)ncase * (size_t)wpc, -1);
    std::vector<long long> slot((size_t)ncase * (size_t)wpc, 0);
    long long n_new = 0;
    std::uniform_int_distribution<int64_t> jit(-(int64_t)len / 4, (int64_t)len / 4);
    for (int i = 0; i < ncase; i++) {
        if (hsrc[i].n < len) continue;
        std::uniform_int_distribution<int64_t> pick(0, hsrc[i].n - len);
        for (int64_t w = 0; w < wpc; w++) {
            int64_t o = (hsrc[i].kif >= 1) ? hsrc[i].kif - 1 - len / 2 + jit(rng) : pick(rng);
            o = std::min<int64_t>(std::max<int64_t>(o, 0), hsrc[i].n - len);
            woff[(size_t)i * wpc + w] = (int)o;
            slot[(size_t)i * wpc + w] = t->data_n + n_new++;
        }
    }
    if (!n_new) return LICA_OK;
    t->data_len = (int)len;
    int rc = net_data_reserve(t, t->data_n + n_new);
    NetPairSrc *dsrc = nullptr;
    int *dwoff = nullptr;
    long long *dslot = nullptr;
    rc = rc ? rc : net_alloc((void **)&dsrc, sizeof(NetPairSrc) * (size_t)ncase);
    rc = rc ? rc : net_alloc((void **)&dwoff, sizeof(int) * woff.size());
    rc = rc ? rc : net_alloc((void **)&dslot, sizeof(long long) * slot.size());
    rc = rc ? rc : hip_ok(hipMemcpy(dsrc, hsrc, sizeof(NetPairSrc) * (size_t)ncase, hipMemcpyHostToDevice));
    rc = rc ? rc : hip_ok(hipMemcpy(dwoff, woff.data(), sizeof(int) * woff.size(), hipMemcpyHostToDevice));
    rc = rc ? rc : hip_ok(hipMemcpy(dslot, slot.data(), sizeof(long long) * slot.size(), hipMemcpyHostToDevice));
    if (!rc) {
        hipLaunchKernelGGL(k_net_pairs, dim3(ncase), dim3(256), 0, t->stream, dsrc, dwoff, dslot, (int)wpc,
                           (int)len, t->data_x, t->data_t);
        rc = hip_ok(hipGetLastError());
        rc = rc ? rc : hip_ok(hipStreamSynchronize(t->stream));
    }
    (void)hipFree(dsrc); (void)hipFree(dwoff); (void)hipFree(dslot);
    if (rc) return rc;
    t->data_n += n_new;
    if (added) *added = n_new;
    return LICA_OK;
}
