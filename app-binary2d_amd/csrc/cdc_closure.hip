// cdc_closure.hip -- learned flux closure for coarse runs: a small periodic
// 3x3 convolutional net (4 -> 32 -> 32 -> 32 -> 6 channels, ReLU) evaluated
// and trained on gfx950 matrix cores (bf16 operands, fp32 accumulation),
// whose 6 output channels are corrections to the x- and y-face fluxes of
// (Sigma, Sx, Sy).  No counterpart in the reference.
//
// Features (clamped to [-16, 16]): a0 ln Sigma, a1 vx, a1 vy, a2 ln(p/Sigma).
// A corrected coarse step: features of the state u_n at the step's start ->
// net -> Y; the plain RK step u_n -> u* (cdc_solver_advance); then
//   u_v = u*_v + dt r_v,  r_v = -s_v ((Y_v[i,j] - Y_v[i-1,j]) / dx
//                                   + (Y_{3+v}[i,j] - Y_{3+v}[i,j-1]) / dy)
// for v = Sigma, Sx, Sy: a flux divergence, so mass and momentum are
// conserved; each face's corrections are scaled down where they would move
// more than a fraction kappa of a neighbour's mass (limit_kernel).  Training fits r to (u_fine - u_plain) / dt on pairs made from a
// down-sampled fine run.
//
// Layouts: activations [pixel][32] bf16 (channel fastest, 64 B per pixel),
// pixel = frame * n^2 + i * n + j; the net's 6 outputs [pixel][8] fp32.
// Every layer is stored padded to 32 x 32 channels (zeros for the first
// layer's inputs 4..31 and the last layer's outputs 6..31, which stay zero
// under training because their gradients vanish).  Master weights, gradients
// and Adam moments are fp32 [layer][co][ci][tap] + [layer][co] (tap = 3 ky + kx);
// the matrix-core operands are bf16 copies [layer][n][k], k = tap * 32 + c:
// forward n = co, c = ci; input gradient n = ci, c = co with the tap flipped.
//
// MFMA 32x32x8 bf16 (v_mfma_f32_32x32x8_bf16): lane l supplies row l % 32 of
// A and column l % 32 of B over k = 4 (l / 32) .. + 3, and holds D rows
// 8 (r / 4) + 4 (l / 32) + r % 4, r = 0..15, of column l % 32.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <random>
#include <string>
#include <vector>

#include "../../include/cdc_solver.h"
#include "cdc_device.h"
#include "cdc_host.h"
#include "cdc_internal.h"

using cdc::host::set_msg;

namespace cdc {
namespace nn {

constexpr int C = 32;          // padded channels
constexpr int K = 9 * C;       // taps x channels
constexpr int NL = 4;          // conv layers
constexpr int LW = C * C * 9;  // padded weights per layer
constexpr int LP = LW + C;     // + bias
constexpr int NPAR = NL * LP;
constexpr int YS = 8;          // output stride (6 used)
constexpr int CIN[NL] = {4, 32, 32, 32};
constexpr int COUT[NL] = {32, 32, 32, 6};

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__host__ __device__ inline uint16_t f2bf(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__host__ __device__ inline float bf2f(uint16_t h) {
    const uint32_t u = (uint32_t)h << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

__device__ inline int wrap(int a, int n) { return a < 0 ? a + n : (a >= n ? a - n : a); }

struct ConvArgs {
    const uint16_t *x;      // [P][32]
    const uint16_t *w;      // [32][K] bf16
    const float *bias;      // [32] or null
    const uint16_t *mask;   // [P][32]: output zeroed where mask <= 0 (or null)
    uint16_t *out_bf;       // [P][32] (or null)
    float *out_f;           // [P][YS] first out_ch channels (when out_bf is null)
    int n, P, relu, out_ch;
};

// One wave: 32 pixels x 32 output channels; four waves per workgroup.
__global__ void __launch_bounds__(256) conv3x3_kernel(const ConvArgs a) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int p0 = (blockIdx.x * 4 + wave) * 32;
    if (p0 >= a.P) return;
    const int m = lane & 31, half = lane >> 5;
    const int nn = a.n * a.n;
    const int p = p0 + m;
    const int f = p / nn, rr = p - f * nn;
    const int i = rr / a.n, j = rr - (rr / a.n) * a.n;
    f32x16 acc;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
    const uint16_t *wrow = a.w + (size_t)m * K + half * 4;
#pragma unroll
    for (int tap = 0; tap < 9; ++tap) {
        const int ii = wrap(i + tap / 3 - 1, a.n), jj = wrap(j + tap % 3 - 1, a.n);
        const uint16_t *xr = a.x + (size_t)(f * nn + ii * a.n + jj) * C + half * 4;
#pragma unroll
        for (int c = 0; c < C / 8; ++c) {
            const bf16x4 av = *(const bf16x4 *)(xr + c * 8);
            const bf16x4 bv = *(const bf16x4 *)(wrow + tap * C + c * 8);
            acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(av, bv, acc, 0, 0, 0);
        }
    }
    const int ch = lane & 31;
    const float b = a.bias ? a.bias[ch] : 0.0f;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int pr = p0 + 8 * (r >> 2) + 4 * half + (r & 3);
        float v = acc[r] + b;
        if (a.relu) v = fmaxf(v, 0.0f);
        if (a.mask && !(bf2f(a.mask[(size_t)pr * C + ch]) > 0.0f)) v = 0.0f;
        if (a.out_bf) a.out_bf[(size_t)pr * C + ch] = f2bf(v);
        else if (ch < a.out_ch) a.out_f[(size_t)pr * YS + ch] = v;
    }
}

// The same convolution for grids whose size is a multiple of 16: a workgroup
// takes a 16 x 16 pixel tile of one frame, staged with its periodic 1-pixel
// halo in LDS (80 B per pixel, so that the 16 B operand reads of 8
// consecutive lanes hit distinct banks); each wave computes 4 rows x 16
// columns as two 32-pixel x 32-channel tiles with v_mfma_f32_32x32x16_bf16
// (lane l supplies row l % 32 of A and column l % 32 of B over
// k = 8 (l / 32) .. + 7; D as for 32x32x8), B (the layer's weights) held in
// registers for the whole tile.  K runs over (tap, 16-channel half) blocks.
// FIRST: the first layer, whose inputs have 4 channels: a block of 16 k
// covers 4 taps x 4 channels, so the 9 taps take 3 blocks instead of 18.
constexpr int TT = 16, TH = TT + 2, TPS = 40;   // tile side, halo side, LDS pixel stride (bf16)
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4b __attribute__((ext_vector_type(4)));

__device__ inline bf16x8 cat4(bf16x4b lo, bf16x4b hi) {
    return __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7);
}

template <bool FIRST>
__global__ void __launch_bounds__(256) conv3x3_tile_kernel(const ConvArgs a) {
    __shared__ __attribute__((aligned(16))) uint16_t tile[TH * TH * TPS];
    constexpr int NB = FIRST ? 3 : 18;   // k blocks of 16
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m = lane & 31, h = lane >> 5;
    const int n = a.n, nn = n * n, tpr = n / TT, tpf = tpr * tpr;
    const int f = blockIdx.x / tpf, t = blockIdx.x - f * tpf;
    const int i0 = (t / tpr) * TT, j0 = (t - (t / tpr) * tpr) * TT;
    const size_t fb = (size_t)f * nn;

    // weights: column m of B, k = 8 h .. + 7 of every block
    bf16x8 wb[NB];
    const uint16_t *wrow = a.w + (size_t)m * K;
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        if constexpr (FIRST) {
            const bf16x4b z = {};
            const int t0 = 4 * b + 2 * h;
            const bf16x4b lo = t0 < 9 ? *(const bf16x4b *)(wrow + t0 * C) : z;
            const bf16x4b hi = t0 + 1 < 9 ? *(const bf16x4b *)(wrow + (t0 + 1) * C) : z;
            wb[b] = cat4(lo, hi);
        } else {
            wb[b] = *(const bf16x8 *)(wrow + (b >> 1) * C + (b & 1) * 16 + 8 * h);
        }
    }
    // the halo tile, 16 B per thread and step (FIRST: the 4 input channels only)
    constexpr int PARTS = FIRST ? 1 : 4;
    for (int q = tid; q < TH * TH * PARTS; q += 256) {
        const int hp = FIRST ? q : q >> 2, part = FIRST ? 0 : q & 3;
        const int hi = hp / TH, hj = hp - hi * TH;
        const int gi = wrap(i0 - 1 + hi, n), gj = wrap(j0 - 1 + hj, n);
        const uint16_t *src = a.x + (fb + (size_t)gi * n + gj) * C + part * 8;
        if constexpr (FIRST) *(uint2 *)(tile + hp * TPS) = *(const uint2 *)src;
        else *(uint4 *)(tile + hp * TPS + part * 8) = *(const uint4 *)src;
    }
    __syncthreads();

    const int ch = lane & 31;
    const float bias = a.bias ? a.bias[ch] : 0.0f;
#pragma unroll
    for (int half = 0; half < 2; ++half) {
        // A rows: pixel m of rows 4 wave + 2 half .. + 1
        const int pr = 4 * wave + 2 * half + (m >> 4), pc = m & 15;
        f32x16 acc;
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            bf16x8 av;
            if constexpr (FIRST) {
                const bf16x4b z = {};
                const int t0 = 4 * b + 2 * h, t1 = t0 + 1;
                const bf16x4b lo = t0 < 9 ? *(const bf16x4b *)(tile + ((pr + t0 / 3) * TH + pc + t0 % 3) * TPS) : z;
                const bf16x4b hi = t1 < 9 ? *(const bf16x4b *)(tile + ((pr + t1 / 3) * TH + pc + t1 % 3) * TPS) : z;
                av = cat4(lo, hi);
            } else {
                const int tap = b >> 1;
                av = *(const bf16x8 *)(tile + ((pr + tap / 3) * TH + pc + tap % 3) * TPS + (b & 1) * 16 + 8 * h);
            }
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, wb[b], acc, 0, 0, 0);
        }
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int mr = 8 * (r >> 2) + 4 * h + (r & 3);
            const int gi = i0 + 4 * wave + 2 * half + (mr >> 4), gj = j0 + (mr & 15);
            const size_t pg = fb + (size_t)gi * n + gj;
            float v = acc[r] + bias;
            if (a.relu) v = fmaxf(v, 0.0f);
            if (a.mask && !(bf2f(a.mask[pg * C + ch]) > 0.0f)) v = 0.0f;
            if (a.out_bf) a.out_bf[pg * C + ch] = f2bf(v);
            else if (ch < a.out_ch) a.out_f[pg * YS + ch] = v;
        }
    }
}

// The whole forward pass over one 16 x 16 output tile (inference: no
// activations kept), for grids whose size is a multiple of 16.  The tile's
// features with a 4-pixel periodic halo (24 x 24 pixels, 4 channels) go to
// LDS; layer 1 computes 22 x 22 pixels, layer 2 20 x 20, layer 3 18 x 18 and
// layer 4 the tile itself, each layer reading the previous one's output from
// LDS (80 B per pixel), so that no intermediate activation leaves the CU.
// Operands and layouts as in conv3x3_tile_kernel; 32-pixel MFMA row tiles
// run over each layer's region in raster order, dealt round-robin to the
// four waves; the last tile of a region is partly padding (computed, not
// stored).  Layer 4's 6 outputs are written as fp32 [pixel][YS].
struct FusedArgs {
    const uint16_t *x;     // features [P][32] (channels 0..3 used)
    const uint16_t *w;     // bf16 forward weights [layer][32][K]
    const float *par;      // fp32 parameters (biases at layer * LP + LW)
    float *y;              // [P][YS]
    int n;
};

constexpr int FR0 = TT + 8, FR1 = TT + 6, FR2 = TT + 4, FR3 = TT + 2;   // region sides
constexpr int FUSED_LDS = FR0 * FR0 * 4 + FR1 * FR1 * TPS + FR2 * FR2 * TPS;  // bf16 elements

// Column m of B, k = 8 h .. + 7 of every k block, for one layer's weights.
template <bool FIRST>
__device__ __forceinline__ void load_weights(const FusedArgs &a, int layer, int lane, bf16x8 *wb) {
    constexpr int NB = FIRST ? 3 : 18;
    const int m = lane & 31, h = lane >> 5;
    const uint16_t *wrow = a.w + (size_t)layer * C * K + (size_t)m * K;
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        if constexpr (FIRST) {
            const bf16x4b z = {};
            const int t0 = 4 * b + 2 * h;
            const bf16x4b lo = t0 < 9 ? *(const bf16x4b *)(wrow + t0 * C) : z;
            const bf16x4b hi = t0 + 1 < 9 ? *(const bf16x4b *)(wrow + (t0 + 1) * C) : z;
            wb[b] = cat4(lo, hi);
        } else {
            wb[b] = *(const bf16x8 *)(wrow + (b >> 1) * C + (b & 1) * 16 + 8 * h);
        }
    }
}

// One layer over its output region: each wave takes row tiles mt and mt + 4
// together (two independent accumulator chains), mt = wave, wave + 8, ...
template <bool FIRST, bool LAST, int rin, int rout>
__device__ __forceinline__ void fused_layer(const FusedArgs &a, int layer, const bf16x8 *wb, const uint16_t *in,
                                            uint16_t *out, int i0, int j0, int wave, int lane) {
    constexpr int NB = FIRST ? 3 : 18;
    constexpr int IPS = FIRST ? 4 : TPS;   // input pixel stride (bf16)
    const int m = lane & 31, h = lane >> 5, ch = lane & 31;
    const float bias = a.par[(size_t)layer * LP + LW + ch];
    constexpr int npx = rout * rout, ntile = (npx + 31) / 32;
    for (int mt = wave; mt < ntile; mt += 8) {
        int base[2];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            // A rows: pixel (mt + 4 u) * 32 + m of the output region (padding clamped to 0)
            int p = (mt + 4 * u) * 32 + m;
            p = p < npx ? p : 0;
            const int pr = p / rout, pc = p - (p / rout) * rout;
            base[u] = pr * rin + pc;
        }
        f32x16 acc[2];
#pragma unroll
        for (int u = 0; u < 2; ++u)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[u][r] = 0.0f;
#pragma unroll
        for (int b = 0; b < NB; ++b) {
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                bf16x8 av;
                if constexpr (FIRST) {
                    const bf16x4b z = {};
                    const int t0 = 4 * b + 2 * h, t1 = t0 + 1;
                    const bf16x4b lo = t0 < 9 ? *(const bf16x4b *)(in + (base[u] + (t0 / 3) * rin + t0 % 3) * IPS) : z;
                    const bf16x4b hi = t1 < 9 ? *(const bf16x4b *)(in + (base[u] + (t1 / 3) * rin + t1 % 3) * IPS) : z;
                    av = cat4(lo, hi);
                } else {
                    const int tap = b >> 1;
                    av = *(const bf16x8 *)(in + (base[u] + (tap / 3) * rin + tap % 3) * IPS + (b & 1) * 16 + 8 * h);
                }
                acc[u] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, wb[b], acc[u], 0, 0, 0);
            }
        }
#pragma unroll
        for (int u = 0; u < 2; ++u) {
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int q = (mt + 4 * u) * 32 + 8 * (r >> 2) + 4 * h + (r & 3);
                if (q >= npx) continue;
                const int qr = q / rout, qc = q - (q / rout) * rout;
                const float v = acc[u][r] + bias;
                if constexpr (LAST) {
                    if (ch < 6) a.y[((size_t)(i0 + qr) * a.n + (j0 + qc)) * YS + ch] = v;
                } else {
                    out[(qr * rout + qc) * TPS + ch] = f2bf(fmaxf(v, 0.0f));
                }
            }
        }
    }
}

// The fused forward pass of one 16 x 16 output tile; feat(F, q, gi, gj)
// writes the 4 bf16 features of grid cell (gi, gj) to F + 4 q.
template <typename Feat>
__device__ __forceinline__ void fused_forward(const FusedArgs &a, Feat feat) {
    __shared__ __attribute__((aligned(16))) uint16_t lds[FUSED_LDS];
    uint16_t *F = lds;                           // features, 24 x 24 x 4
    uint16_t *A1 = F + FR0 * FR0 * 4;            // layer 1 (22 x 22), then layer 3 (18 x 18)
    uint16_t *A2 = A1 + FR1 * FR1 * TPS;         // layer 2 (20 x 20)
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int n = a.n, tpr = n / TT;
    const int i0 = (blockIdx.x / tpr) * TT, j0 = (blockIdx.x - (blockIdx.x / tpr) * tpr) * TT;
    // two register sets of weights: the next layer's load while a layer computes
    bf16x8 wa[18], wb[18];
    load_weights<true>(a, 0, lane, wa);
    load_weights<false>(a, 1, lane, wb);
    for (int q = tid; q < FR0 * FR0; q += 256) {
        const int fi = q / FR0, fj = q - fi * FR0;
        feat(F, q, wrap(i0 - 4 + fi, n), wrap(j0 - 4 + fj, n));
    }
    __syncthreads();
    fused_layer<true, false, FR0, FR1>(a, 0, wa, F, A1, i0, j0, wave, lane);
    load_weights<false>(a, 2, lane, wa);
    __syncthreads();
    fused_layer<false, false, FR1, FR2>(a, 1, wb, A1, A2, i0, j0, wave, lane);
    load_weights<false>(a, 3, lane, wb);
    __syncthreads();
    fused_layer<false, false, FR2, FR3>(a, 2, wa, A2, A1, i0, j0, wave, lane);
    __syncthreads();
    fused_layer<false, true, FR3, TT>(a, 3, wb, A1, nullptr, i0, j0, wave, lane);
}

// the features from the feature buffer a.x
__global__ void __launch_bounds__(256) net_forward_fused_kernel(const FusedArgs a) {
    fused_forward(a, [&](uint16_t *F, int q, int gi, int gj) {
        *(uint2 *)(F + q * 4) = *(const uint2 *)(a.x + ((size_t)gi * a.n + gj) * C);
    });
}

struct WgradArgs {
    const uint16_t *g;   // [P][32] output gradient
    const uint16_t *x;   // [P][32] layer input
    float *part_w;       // [nwg][9][32][32]
    float *part_b;       // [nwg][32]
    int n, P, chunk;     // chunk: pixels per workgroup, a multiple of 8
};

// dW[co][ci][tap] = sum_p g[p][co] x[p + off(tap)][ci] over one chunk of
// pixels, one wave per tap (A = g^T, B = shifted x, K = pixels); wave 0 also
// sums g for the bias.
__global__ void __launch_bounds__(576) wgrad_kernel(const WgradArgs a) {
    const int lane = threadIdx.x & 63, tap = threadIdx.x >> 6;
    const int half = lane >> 5, col = lane & 31;
    const int nn = a.n * a.n;
    const int start = blockIdx.x * a.chunk;
    const int end = min(a.P, start + a.chunk);
    const int di = tap / 3 - 1, dj = tap % 3 - 1;
    f32x16 acc;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
    float bsum = 0.0f;
    for (int p8 = start; p8 < end; p8 += 8) {
        bf16x4 av, bv;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int p = p8 + half * 4 + q;
            const int f = p / nn, rr = p - f * nn;
            const int i = rr / a.n, j = rr - (rr / a.n) * a.n;
            const int src = f * nn + wrap(i + di, a.n) * a.n + wrap(j + dj, a.n);
            const uint16_t gv = a.g[(size_t)p * C + col];
            av[q] = (short)gv;
            bv[q] = (short)a.x[(size_t)src * C + col];
            if (tap == 0) bsum += bf2f(gv);
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(av, bv, acc, 0, 0, 0);
    }
    float *pw = a.part_w + ((size_t)blockIdx.x * 9 + tap) * C * C;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int co = 8 * (r >> 2) + 4 * half + (r & 3);
        pw[co * C + col] = acc[r];
    }
    if (tap == 0) {
        const float other = __shfl_down(bsum, 32, 64);
        if (half == 0) a.part_b[(size_t)blockIdx.x * C + col] = bsum + other;
    }
}

// grad[layer] = sum over workgroups of the partials, in workgroup order.
__global__ void __launch_bounds__(256) wgrad_reduce_kernel(const float *part_w, const float *part_b, int nwg,
                                                           float *grad_layer) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t < LW) {
        // canonical index t = (co * 32 + ci) * 9 + tap
        const int tap = t % 9, cc = t / 9, ci = cc % C, co = cc / C;
        float s = 0.0f;
        for (int w = 0; w < nwg; ++w) s += part_w[(((size_t)w * 9 + tap) * C + co) * C + ci];
        grad_layer[t] = s;
    } else if (t < LP) {
        const int co = t - LW;
        float s = 0.0f;
        for (int w = 0; w < nwg; ++w) s += part_b[(size_t)w * C + co];
        grad_layer[t] = s;
    }
}

struct LossArgs {
    const float *y;        // [P][YS]
    const float *t;        // [P][4] target
    uint16_t *gy;          // [P][32] dL/dY
    float *part;           // [nblocks] per-block loss sums
    int n, P;
    float idx, idy, inv_p;
    float s[3], w[3];
};

__device__ inline float corr_r(const LossArgs &a, int f, int i, int j, int v) {
    const int nn = a.n * a.n;
    const size_t c = (size_t)(f * nn + i * a.n + j);
    const size_t xm = (size_t)(f * nn + wrap(i - 1, a.n) * a.n + j);
    const size_t ym = (size_t)(f * nn + i * a.n + wrap(j - 1, a.n));
    return -a.s[v] * ((a.y[c * YS + v] - a.y[xm * YS + v]) * a.idx + (a.y[c * YS + 3 + v] - a.y[ym * YS + 3 + v]) * a.idy);
}

__global__ void __launch_bounds__(256) loss_kernel(const LossArgs a) {
    __shared__ float red[256];
    const int p = blockIdx.x * 256 + threadIdx.x;
    float l = 0.0f;
    if (p < a.P) {
        const int nn = a.n * a.n;
        const int f = p / nn, rr = p - f * nn;
        const int i = rr / a.n, j = rr - (rr / a.n) * a.n;
        const int ip = wrap(i + 1, a.n), jp = wrap(j + 1, a.n);
        uint16_t *g = a.gy + (size_t)p * C;
        // cell weights (target channel 3) scale each cell's error
        const size_t px = (size_t)(f * nn + ip * a.n + j), py = (size_t)(f * nn + i * a.n + jp);
        const float c0 = a.t[(size_t)p * 4 + 3], cx = a.t[px * 4 + 3], cy = a.t[py * 4 + 3];
#pragma unroll
        for (int v = 0; v < 3; ++v) {
            const float e0 = c0 * (corr_r(a, f, i, j, v) - a.t[(size_t)p * 4 + v]);
            const float ex = cx * (corr_r(a, f, ip, j, v) - a.t[px * 4 + v]);
            const float ey = cy * (corr_r(a, f, i, jp, v) - a.t[py * 4 + v]);
            l += a.w[v] * e0 * e0;
            const float k = 2.0f * a.inv_p * a.w[v] * a.s[v];
            g[v] = f2bf(-k * a.idx * (e0 - ex));
            g[3 + v] = f2bf(-k * a.idy * (e0 - ey));
        }
        for (int c = 6; c < C; ++c) g[c] = 0;
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) a.part[blockIdx.x] = red[0];
}

// loss of step `step` = (sum of block partials in order) / P
__global__ void loss_sum_kernel(const float *part, int nb, float inv_p, float *hist, int step) {
    if (threadIdx.x != 0) return;
    float s = 0.0f;
    for (int b = 0; b < nb; ++b) s += part[b];
    hist[step] = s * inv_p;
}

struct AdamArgs {
    float *p, *m, *v;
    const float *g;
    uint16_t *wf, *wb;   // bf16 operand copies [NL][32][K]
    float lr, b1, b2, eps, c1, c2;   // c1 = 1 / (1 - b1^t), c2 = 1 / (1 - b2^t)
    int update;          // 0: only refresh the bf16 copies
};

__global__ void __launch_bounds__(256) adam_kernel(const AdamArgs a) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= NPAR) return;
    float p = a.p[t];
    if (a.update) {
        const float g = a.g[t];
        const float m = a.b1 * a.m[t] + (1.0f - a.b1) * g;
        const float v = a.b2 * a.v[t] + (1.0f - a.b2) * g * g;
        a.m[t] = m;
        a.v[t] = v;
        p -= a.lr * (m * a.c1) / (sqrtf(v * a.c2) + a.eps);
        a.p[t] = p;
    }
    const int l = t / LP, o = t - l * LP;
    if (o < LW) {
        const int tap = o % 9, cc = o / 9, ci = cc % C, co = cc / C;
        const uint16_t h = f2bf(p);
        a.wf[((size_t)l * C + co) * K + tap * C + ci] = h;
        a.wb[((size_t)l * C + ci) * K + (8 - tap) * C + co] = h;
    }
}

struct FeatArgs {
    const double *u;     // device state (sidx layout)
    uint16_t *x;         // [N*N][32]
    int N, nv;
    double gm1, a0, a1, a2;
    // isothermal: the bodies at the state's time, the softening length
    // squared and the grid's geometry (cell i, j centred at -R + (i + 1/2) dx)
    double x1, y1, m1, x2, y2, m2, rs2, R, dx;
};

// the 4 features of cell p, as stored (bf16, clamped to [-16, 16])
__device__ __forceinline__ void feature_values(const FeatArgs &a, int64_t p, uint16_t *o) {
    const int64_t NN = (int64_t)a.N * a.N;
    const double sig = a.u[cdc::sidx(a.nv, 0, NN, p)], sx = a.u[cdc::sidx(a.nv, 1, NN, p)],
                 sy = a.u[cdc::sidx(a.nv, 2, NN, p)];
    const double vx = sx / sig, vy = sy / sig;
    float f[4];
    f[0] = (float)(a.a0 * log(fmax(sig, 1e-30)));
    f[1] = (float)(a.a1 * vx);
    f[2] = (float)(a.a1 * vy);
    f[3] = 0.0f;
    if (a.nv == 4) {
        const double pr = (a.u[cdc::sidx(a.nv, 3, NN, p)] - 0.5 * sig * (vx * vx + vy * vy)) * a.gm1;
        f[3] = (float)(a.a2 * log(fmax(pr / sig, 1e-30)));
    } else {
        // ln(-phi), phi the softened two-body potential at the cell centre
        const int i = (int)(p / a.N), j = (int)(p - (p / a.N) * a.N);
        const double x = -a.R + (i + 0.5) * a.dx, y = -a.R + (j + 0.5) * a.dx;
        const double d1x = x - a.x1, d1y = y - a.y1, d2x = x - a.x2, d2y = y - a.y2;
        const double mphi = a.m1 / sqrt(d1x * d1x + d1y * d1y + a.rs2) + a.m2 / sqrt(d2x * d2x + d2y * d2y + a.rs2);
        f[3] = (float)(a.a2 * log(fmax(mphi, 1e-30)));
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) o[c] = f2bf(fminf(fmaxf(f[c], -16.0f), 16.0f));
}

__device__ __forceinline__ void features_at(const FeatArgs &a) {
    const int64_t NN = (int64_t)a.N * a.N;
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= NN) return;
    uint16_t *o = a.x + p * C;
    feature_values(a, p, o);
    for (int c = 4; c < C; ++c) o[c] = 0;
}

__global__ void __launch_bounds__(256) features_kernel(const FeatArgs a) { features_at(a); }

// The state pointers and dt of step f of a corrected advance, read from
// device memory so that a captured graph serves every call.
struct StepIO {
    const double *u_in;   // the state the step starts from (the net's input)
    double *u_out;        // the step's result, corrected in place
    double dt;            // the correction's time: the group's steps x dt
    double x1, y1, m1, x2, y2, m2;   // the bodies at the step's start
    double *e_out;                   // the strip-edge columns of u_out (null: not marching)
};

// The strip-edge columns of step f's corrected result (the next step's
// column-marching stage reads them).
__global__ void __launch_bounds__(256) edge_step_kernel(const StepIO *io, int f, int N, int nv) {
    const int64_t total = (int64_t)nv * N * (N / 64) * 4;
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e < total) io[f].e_out[e] = cdc::edge_entry(io[f].u_out, e, N, nv);
}

// Per-workgroup maxima of the signal speed of the last step's corrected
// result (grid-stride), for the next call's dt.
template <int NV>
__global__ void __launch_bounds__(256) max_speed_step_kernel(const StepIO *io, int f, int64_t ncell, double gm,
                                                             double *partial) {
    __shared__ double red[256];
    const double *u = io[f].u_out;
    double m = 0.0;
    for (int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x; c < ncell; c += (int64_t)gridDim.x * 256) {
        double w[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) w[v] = u[cdc::sidx(NV, v, ncell, c)];
        m = fmax(m, cdc::dev::max_signal_speed<NV>(w, gm));
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(256) features_step_kernel(FeatArgs a, const StepIO *io, int f) {
    const StepIO &o = io[f];
    a.u = o.u_in;
    a.x1 = o.x1; a.y1 = o.y1; a.m1 = o.m1;
    a.x2 = o.x2; a.y2 = o.y2; a.m2 = o.m2;
    features_at(a);
}

// The fused forward pass of step f of a corrected advance with the features
// computed from the step's starting state in the kernel (the same values as
// features_step_kernel's, for the tile and its halo), so that no feature
// buffer is written and read back.
__global__ void __launch_bounds__(256) net_forward_fused_step_kernel(const FusedArgs a, FeatArgs fa,
                                                                     const StepIO *io, int f) {
    const StepIO &o = io[f];
    fa.u = o.u_in;
    fa.x1 = o.x1; fa.y1 = o.y1; fa.m1 = o.m1;
    fa.x2 = o.x2; fa.y2 = o.y2; fa.m2 = o.m2;
    fused_forward(a, [&](uint16_t *F, int q, int gi, int gj) {
        uint16_t v[4];
        feature_values(fa, (int64_t)gi * a.n + gj, v);
        *(uint2 *)(F + q * 4) = make_uint2((unsigned)v[0] | ((unsigned)v[1] << 16),
                                           (unsigned)v[2] | ((unsigned)v[3] << 16));
    });
}

struct TargetArgs {
    const double *plain, *truth;   // device states (sidx layout)
    const double *base;            // the state the features come from
    float *t;                      // [N*N][4]: 3 targets and the cell weight
    float *part;                   // [nblocks][3] weighted sums of t_v^2
    int N, nv;
    double inv_dt, mask;           // cell weight 1 where base Sigma >= mask, else 0
};

__global__ void __launch_bounds__(256) target_kernel(const TargetArgs a) {
    __shared__ float red[3][256];
    const int64_t NN = (int64_t)a.N * a.N;
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    float sq[3] = {0.0f, 0.0f, 0.0f};
    if (p < NN) {
        const float wc = a.base[cdc::sidx(a.nv, 0, NN, p)] >= a.mask ? 1.0f : 0.0f;
        for (int v = 0; v < 3; ++v) {
            const int64_t q = cdc::sidx(a.nv, v, NN, p);
            const float d = (float)((a.truth[q] - a.plain[q]) * a.inv_dt);
            a.t[p * 4 + v] = d;
            sq[v] = wc * d * d;
        }
        a.t[p * 4 + 3] = wc;
    }
    for (int v = 0; v < 3; ++v) red[v][threadIdx.x] = sq[v];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int v = 0; v < 3; ++v) red[v][threadIdx.x] += red[v][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x < 3) a.part[blockIdx.x * 3 + threadIdx.x] = red[threadIdx.x][0];
}

struct ApplyArgs {
    double *u;            // device state (sidx layout), corrected in place
    const float *y;       // [N*N][YS]
    double *lam;          // [N*N][2] face limiter factors (x+, y+ face of each cell)
    int N, nv;
    double dt, idx, idy, kappa, gm, mask;
    double s[3];
};

// Limiter, one pass before the update (so that it reads the uncorrected
// state): the three corrections of a face are scaled by the largest
// lambda <= 1 with
//   dt |s_0 Y_Sigma| / dx        <= kappa min(Sigma_L, Sigma_R)
//   dt |s_v Y_Sv| / dx (v = 1,2) <= kappa min(Sigma_L, Sigma_R) max(a_L, a_R)
// (a = |v| + c_s, the cell's signal speed; c_s = 0 for isothermal), so that
// no face moves more than the fraction kappa of a neighbour's mass, or of its
// mass times the local signal speed in momentum.  kappa <= 0 disables it.
// Faces where either density is below mask_density get no correction.
__device__ inline double signal_speed(const ApplyArgs &a, int64_t NN, int64_t q) {
    const double sig = a.u[cdc::sidx(a.nv, 0, NN, q)];
    const double vx = a.u[cdc::sidx(a.nv, 1, NN, q)] / sig, vy = a.u[cdc::sidx(a.nv, 2, NN, q)] / sig;
    double c2 = 0.0;
    if (a.nv == 4)
        c2 = a.gm * (a.u[cdc::sidx(a.nv, 3, NN, q)] - 0.5 * sig * (vx * vx + vy * vy)) * (a.gm - 1.0) / sig;
    return sqrt(vx * vx + vy * vy) + sqrt(fmax(c2, 0.0));
}

__device__ __forceinline__ void limit_at(const ApplyArgs &a) {
    const int64_t NN = (int64_t)a.N * a.N;
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= NN) return;
    const int i = (int)(p / a.N), j = (int)(p - (p / a.N) * a.N);
    const int64_t nb[2] = {(int64_t)wrap(i + 1, a.N) * a.N + j, (int64_t)i * a.N + wrap(j + 1, a.N)};
    const double id[2] = {a.idx, a.idy};
    double l[2] = {1.0, 1.0};
    for (int d = 0; d < 2; ++d)
        if (fmin(a.u[2 * p], a.u[2 * nb[d]]) < a.mask) l[d] = 0.0;   // Sigma: sidx(nv, 0, NN, c) = 2 c
    if (a.kappa > 0.0) {
        const double sig = a.u[2 * p];
        const double ap = signal_speed(a, NN, p);
        for (int d = 0; d < 2; ++d) {
            if (l[d] == 0.0) continue;
            const double cap = a.kappa * fmax(fmin(sig, a.u[2 * nb[d]]), 0.0);
            const double capm = cap * fmax(ap, signal_speed(a, NN, nb[d]));
            double lam = 1.0;
            for (int v = 0; v < 3; ++v) {
                const double m = a.dt * fabs(a.s[v] * (double)a.y[p * YS + 3 * d + v]) * id[d];
                const double c = v == 0 ? cap : capm;
                if (m > c) lam = fmin(lam, c / m);
            }
            l[d] = lam;
        }
    }
    a.lam[p * 2] = l[0];
    a.lam[p * 2 + 1] = l[1];
}

__device__ __forceinline__ void apply_at(const ApplyArgs &a) {
    const int64_t NN = (int64_t)a.N * a.N;
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= NN) return;
    const int i = (int)(p / a.N), j = (int)(p - (p / a.N) * a.N);
    const int64_t xm = (int64_t)wrap(i - 1, a.N) * a.N + j, ym = (int64_t)i * a.N + wrap(j - 1, a.N);
    const double lxp = a.lam[p * 2], lxm = a.lam[xm * 2], lyp = a.lam[p * 2 + 1], lym = a.lam[ym * 2 + 1];
    for (int v = 0; v < 3; ++v) {
        const double fx = lxp * (double)a.y[p * YS + v] - lxm * (double)a.y[xm * YS + v];
        const double fy = lyp * (double)a.y[p * YS + 3 + v] - lym * (double)a.y[ym * YS + 3 + v];
        const double r = -a.s[v] * (fx * a.idx + fy * a.idy);
        a.u[cdc::sidx(a.nv, v, NN, p)] += a.dt * r;
    }
}

__global__ void __launch_bounds__(256) limit_kernel(const ApplyArgs a) { limit_at(a); }
__global__ void __launch_bounds__(256) apply_kernel(const ApplyArgs a) { apply_at(a); }

__global__ void __launch_bounds__(256) limit_step_kernel(ApplyArgs a, const StepIO *io, int f) {
    a.u = io[f].u_out;
    a.dt = io[f].dt;
    limit_at(a);
}

__global__ void __launch_bounds__(256) apply_step_kernel(ApplyArgs a, const StepIO *io, int f) {
    a.u = io[f].u_out;
    a.dt = io[f].dt;
    apply_at(a);
}

}  // namespace nn
}  // namespace cdc

using namespace cdc::nn;

struct cdc_closure {
    cdc_closure_config_t cfg;
    int n = 0, P = 0;
    hipStream_t stream = nullptr;
    float *d_par = nullptr, *d_grad = nullptr, *d_m = nullptr, *d_v = nullptr;
    uint16_t *d_wf = nullptr, *d_wb = nullptr;
    uint16_t *d_act[4] = {};      // X0, A1, A2, A3
    uint16_t *d_g[2] = {};        // gradient ping-pong [P][32]
    float *d_y = nullptr;         // [P][YS]
    double *d_lam = nullptr;      // [P][2] limiter factors
    float *d_part_w = nullptr, *d_part_b = nullptr, *d_part_l = nullptr;
    int nwg = 0, chunk = 0, nlb = 0;
    // dataset
    int cap = 0, count = 0;
    uint16_t *d_feat = nullptr;   // [cap][P][32]
    float *d_tgt = nullptr;       // [cap][P][4]
    double tsq[3] = {0, 0, 0};    // sum over samples and cells of target^2
    float *d_hist = nullptr;
    int hist_cap = 0;
    int64_t adam_step = 0;
    // corrected advances of `fold` steps: per-step state pointers and dt
    // (pinned staging and device copy), and the captured graphs, each valid
    // for one solver object, its graph epoch, fold and tile shape
    StepIO *h_io = nullptr, *d_io = nullptr;
    int cap_io = 0;
    struct FoldGraph {
        uint64_t uid, epoch;
        int fold, nstages;
        cdc::TileShape ts;
        hipGraphExec_t exec;
    };
    std::vector<FoldGraph> graphs;
    bool use_graph = true;
    int interval = 1;   // steps per correction (cdc_closure_set_interval)
};

namespace {

void free_closure(cdc_closure *c) {
    float *fl[] = {c->d_par, c->d_grad, c->d_m, c->d_v, c->d_y, c->d_part_w, c->d_part_b, c->d_part_l, c->d_tgt,
                   c->d_hist};
    for (float *p : fl) (void)hipFree(p);
    (void)hipFree(c->d_lam);
    uint16_t *hl[] = {c->d_wf, c->d_wb, c->d_act[0], c->d_act[1], c->d_act[2], c->d_act[3], c->d_g[0], c->d_g[1],
                      c->d_feat};
    for (uint16_t *p : hl) (void)hipFree(p);
    for (auto &g : c->graphs) (void)hipGraphExecDestroy(g.exec);
    c->graphs.clear();
    (void)hipFree(c->d_io);
    if (c->h_io) (void)hipHostFree(c->h_io);
    if (c->stream) (void)hipStreamDestroy(c->stream);
}

// The captured corrected advances hold the output scales by value.
void drop_fold_graphs(cdc_closure *c) {
    for (auto &g : c->graphs) (void)hipGraphExecDestroy(g.exec);
    c->graphs.clear();
}

int refresh_weights(cdc_closure *c, hipStream_t s) {
    AdamArgs a{};
    a.p = c->d_par;
    a.m = c->d_m;
    a.v = c->d_v;
    a.g = c->d_grad;
    a.wf = c->d_wf;
    a.wb = c->d_wb;
    a.update = 0;
    hipLaunchKernelGGL(adam_kernel, dim3((NPAR + 255) / 256), dim3(256), 0, s, a);
    HIP_TRY(hipGetLastError());
    return CDC_OK;
}

int conv(cdc_closure *c, const uint16_t *x, int layer, bool backward, const uint16_t *mask, uint16_t *out_bf,
         float *out_f, hipStream_t s) {
    ConvArgs a{};
    a.x = x;
    a.w = (backward ? c->d_wb : c->d_wf) + (size_t)layer * C * K;
    a.bias = backward ? nullptr : c->d_par + (size_t)layer * LP + LW;
    a.mask = mask;
    a.out_bf = out_bf;
    a.out_f = out_f;
    a.n = c->n;
    a.P = c->P;
    a.relu = (!backward && layer < NL - 1) ? 1 : 0;
    a.out_ch = 6;
    if (c->n % TT == 0) {
        const dim3 grid((unsigned)(c->P / (TT * TT)));
        if (layer == 0 && !backward) hipLaunchKernelGGL(conv3x3_tile_kernel<true>, grid, dim3(256), 0, s, a);
        else hipLaunchKernelGGL(conv3x3_tile_kernel<false>, grid, dim3(256), 0, s, a);
    } else {
        hipLaunchKernelGGL(conv3x3_kernel, dim3((c->P + 127) / 128), dim3(256), 0, s, a);
    }
    HIP_TRY(hipGetLastError());
    return CDC_OK;
}

// the layer-by-layer forward pass (training: backprop reads the activations)
int forward_layers(cdc_closure *c, hipStream_t s) {
    int rc;
    for (int l = 0; l < NL - 1; ++l)
        if ((rc = conv(c, c->d_act[l], l, false, nullptr, c->d_act[l + 1], nullptr, s))) return rc;
    return conv(c, c->d_act[NL - 1], NL - 1, false, nullptr, nullptr, c->d_y, s);
}

// the net's outputs only (inference): fused over 16 x 16 tiles where the
// grid allows, layer by layer otherwise
bool fused_grid(const cdc_closure *c) { return c->n % TT == 0 && c->P == c->n * c->n; }

FusedArgs fused_args(const cdc_closure *c) {
    FusedArgs a{};
    a.x = c->d_act[0];
    a.w = c->d_wf;
    a.par = c->d_par;
    a.y = c->d_y;
    a.n = c->n;
    return a;
}

int forward(cdc_closure *c, hipStream_t s) {
    if (!fused_grid(c)) return forward_layers(c, s);
    hipLaunchKernelGGL(net_forward_fused_kernel, dim3((unsigned)(c->P / (TT * TT))), dim3(256), 0, s,
                       fused_args(c));
    HIP_TRY(hipGetLastError());
    return CDC_OK;
}

int wgrad(cdc_closure *c, const uint16_t *g, const uint16_t *x, int layer, hipStream_t s) {
    WgradArgs a{};
    a.g = g;
    a.x = x;
    a.part_w = c->d_part_w;
    a.part_b = c->d_part_b;
    a.n = c->n;
    a.P = c->P;
    a.chunk = c->chunk;
    hipLaunchKernelGGL(wgrad_kernel, dim3(c->nwg), dim3(576), 0, s, a);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(wgrad_reduce_kernel, dim3((LP + 255) / 256), dim3(256), 0, s, (const float *)c->d_part_w,
                       (const float *)c->d_part_b, c->nwg, c->d_grad + (size_t)layer * LP);
    HIP_TRY(hipGetLastError());
    return CDC_OK;
}

// loss of the current forward pass against target t; dY into d_g[0]; the
// step's loss into d_hist[step]
int loss(cdc_closure *c, const float *t, int step, hipStream_t s) {
    LossArgs a{};
    a.y = c->d_y;
    a.t = t;
    a.gy = c->d_g[0];
    a.part = c->d_part_l;
    a.n = c->n;
    a.P = c->P;
    const double dx = 2.0 * c->cfg.domain_radius / c->n;
    a.idx = (float)(1.0 / dx);
    a.idy = (float)(1.0 / dx);
    a.inv_p = 1.0f / (float)c->P;
    for (int v = 0; v < 3; ++v) {
        a.s[v] = (float)c->cfg.output_scale[v];
        a.w[v] = (float)c->cfg.loss_weight[v];
    }
    hipLaunchKernelGGL(loss_kernel, dim3(c->nlb), dim3(256), 0, s, a);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(loss_sum_kernel, dim3(1), dim3(64), 0, s, (const float *)c->d_part_l, c->nlb, a.inv_p,
                       c->d_hist, step);
    HIP_TRY(hipGetLastError());
    return CDC_OK;
}

// gradient of the loss of sample k into d_grad (forward included)
int backprop(cdc_closure *c, int k, int step, hipStream_t s) {
    HIP_TRY(hipMemcpyAsync(c->d_act[0], c->d_feat + (size_t)k * c->P * C, (size_t)c->P * C * sizeof(uint16_t),
                           hipMemcpyDeviceToDevice, s));
    int rc;
    if ((rc = forward_layers(c, s))) return rc;
    if ((rc = loss(c, c->d_tgt + (size_t)k * c->P * 4, step, s))) return rc;
    int cur = 0;
    for (int l = NL - 1; l >= 0; --l) {
        if ((rc = wgrad(c, c->d_g[cur], c->d_act[l], l, s))) return rc;
        if (l > 0) {
            // dA_l = conv^T(g, W_l) masked by A_l > 0
            if ((rc = conv(c, c->d_g[cur], l, true, c->d_act[l], c->d_g[1 - cur], nullptr, s))) return rc;
            cur = 1 - cur;
        }
    }
    return CDC_OK;
}

int adam_update(cdc_closure *c, double lr, hipStream_t s) {
    c->adam_step += 1;
    AdamArgs a{};
    a.p = c->d_par;
    a.m = c->d_m;
    a.v = c->d_v;
    a.g = c->d_grad;
    a.wf = c->d_wf;
    a.wb = c->d_wb;
    a.lr = (float)lr;
    a.b1 = 0.9f;
    a.b2 = 0.999f;
    a.eps = 1e-8f;
    a.c1 = (float)(1.0 / (1.0 - pow(0.9, (double)c->adam_step)));
    a.c2 = (float)(1.0 / (1.0 - pow(0.999, (double)c->adam_step)));
    a.update = 1;
    hipLaunchKernelGGL(adam_kernel, dim3((NPAR + 255) / 256), dim3(256), 0, s, a);
    HIP_TRY(hipGetLastError());
    return CDC_OK;
}

int ensure_hist(cdc_closure *c, int n) {
    if (n <= c->hist_cap) return CDC_OK;
    (void)hipFree(c->d_hist);
    c->d_hist = nullptr;
    c->hist_cap = 0;
    HIP_TRY(hipMalloc((void **)&c->d_hist, (size_t)n * sizeof(float)));
    c->hist_cap = n;
    return CDC_OK;
}

// flat canonical (logical) index <-> padded device index
template <typename F>
void for_each_param(F fn) {
    int64_t k = 0;
    for (int l = 0; l < NL; ++l) {
        for (int co = 0; co < COUT[l]; ++co)
            for (int ci = 0; ci < CIN[l]; ++ci)
                for (int tap = 0; tap < 9; ++tap) fn(k++, (int64_t)l * LP + ((int64_t)co * C + ci) * 9 + tap);
        for (int co = 0; co < COUT[l]; ++co) fn(k++, (int64_t)l * LP + LW + co);
    }
}

int64_t logical_params() {
    int64_t n = 0;
    for_each_param([&](int64_t, int64_t) { ++n; });
    return n;
}

// the softening and grid geometry of the isothermal potential feature
void feature_geometry(const cdc_solver *s, FeatArgs &a) {
    const cdc_physics_t &p = s->phys;
    const double rs = p.has_softening_length ? p.softening_length : p.sink_radius;   // src/physics.rs:399-405
    a.rs2 = rs * rs;
    a.R = s->mesh.domain_radius;
    a.dx = 2.0 * s->mesh.domain_radius / (double)s->N;
}

int features(cdc_closure *c, const cdc_solver *s, uint16_t *x, hipStream_t st) {
    FeatArgs a{};
    a.u = s->d_u[s->cur];
    a.x = x;
    a.N = s->N;
    a.nv = s->nv;
    a.gm1 = s->hydro.kind == CDC_EULER ? s->hydro.gamma_law_index - 1.0 : 0.0;
    a.a0 = c->cfg.feature_scale[0];
    a.a1 = c->cfg.feature_scale[1];
    a.a2 = c->cfg.feature_scale[2];
    feature_geometry(s, a);
    double el[4];
    cdc::orbital_elements(s->phys, el);
    const cdc::Orbit o = cdc::orbital_state(el, s->time);
    a.x1 = o.x1; a.y1 = o.y1; a.m1 = o.m1;
    a.x2 = o.x2; a.y2 = o.y2; a.m2 = o.m2;
    hipLaunchKernelGGL(features_kernel, dim3((unsigned)(((int64_t)s->N * s->N + 255) / 256)), dim3(256), 0, st, a);
    HIP_TRY(hipGetLastError());
    return CDC_OK;
}

int check_solver(const cdc_closure *c, const cdc_solver *s) {
    if (!s) return set_msg(CDC_ERR_INVALID_ARGUMENT, "null solver");
    if (s->N != c->n) return set_msg(CDC_ERR_INVALID_ARGUMENT, "solver grid size differs from the closure's");
    if (s->mesh.domain_radius != c->cfg.domain_radius)
        return set_msg(CDC_ERR_INVALID_ARGUMENT, "solver domain differs from the closure's");
    return CDC_OK;
}

}  // namespace

extern "C" {

int32_t cdc_closure_create(const cdc_closure_config_t *cfg, cdc_closure_t **out) {
    if (!out) return set_msg(CDC_ERR_INVALID_ARGUMENT, "null output handle");
    *out = nullptr;
    if (!cfg) return set_msg(CDC_ERR_INVALID_ARGUMENT, "null config");
    if (cfg->grid < 8 || cfg->grid % 8 != 0 || cfg->grid > 4096)
        return set_msg(CDC_ERR_INVALID_ARGUMENT, "grid must be a multiple of 8 in [8, 4096]");
    if (cfg->max_samples < 0) return set_msg(CDC_ERR_INVALID_ARGUMENT, "max_samples must be >= 0");
    if (!(cfg->domain_radius > 0.0)) return set_msg(CDC_ERR_INVALID_ARGUMENT, "domain_radius must be > 0");
    if (cdc_device_count() < 1) return set_msg(CDC_ERR_NO_DEVICE, "no HIP device visible");
    cdc_closure *c = new cdc_closure();
    c->cfg = *cfg;
    c->n = cfg->grid;
    c->P = cfg->grid * cfg->grid;
    c->chunk = std::max(8, ((c->P / 256) + 7) / 8 * 8);
    c->nwg = (c->P + c->chunk - 1) / c->chunk;
    c->nlb = (c->P + 255) / 256;
    c->cap = cfg->max_samples;
    auto fail = [&](int code, const char *m) {
        free_closure(c);
        delete c;
        return set_msg(code, m);
    };
    bool ok = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) == hipSuccess;
    const size_t act = (size_t)c->P * C * sizeof(uint16_t);
    ok = ok && hipMalloc((void **)&c->d_par, NPAR * sizeof(float)) == hipSuccess;
    ok = ok && hipMalloc((void **)&c->d_grad, NPAR * sizeof(float)) == hipSuccess;
    ok = ok && hipMalloc((void **)&c->d_m, NPAR * sizeof(float)) == hipSuccess;
    ok = ok && hipMalloc((void **)&c->d_v, NPAR * sizeof(float)) == hipSuccess;
    ok = ok && hipMalloc((void **)&c->d_wf, (size_t)NL * C * K * sizeof(uint16_t)) == hipSuccess;
    ok = ok && hipMalloc((void **)&c->d_wb, (size_t)NL * C * K * sizeof(uint16_t)) == hipSuccess;
    for (int i = 0; i < 4 && ok; ++i) ok = hipMalloc((void **)&c->d_act[i], act) == hipSuccess;
    for (int i = 0; i < 2 && ok; ++i) ok = hipMalloc((void **)&c->d_g[i], act) == hipSuccess;
    ok = ok && hipMalloc((void **)&c->d_y, (size_t)c->P * YS * sizeof(float)) == hipSuccess;
    ok = ok && hipMalloc((void **)&c->d_lam, (size_t)c->P * 2 * sizeof(double)) == hipSuccess;
    ok = ok && hipMalloc((void **)&c->d_part_w, (size_t)c->nwg * 9 * C * C * sizeof(float)) == hipSuccess;
    ok = ok && hipMalloc((void **)&c->d_part_b, (size_t)c->nwg * C * sizeof(float)) == hipSuccess;
    ok = ok && hipMalloc((void **)&c->d_part_l, (size_t)c->nlb * 3 * sizeof(float)) == hipSuccess;
    if (ok && c->cap > 0) {
        ok = hipMalloc((void **)&c->d_feat, (size_t)c->cap * act) == hipSuccess &&
             hipMalloc((void **)&c->d_tgt, (size_t)c->cap * c->P * 4 * sizeof(float)) == hipSuccess;
    }
    if (!ok) {
        (void)hipGetLastError();
        return fail(CDC_ERR_OUT_OF_MEMORY, "closure buffers");
    }
    // He-normal initial weights from the seed (zero biases, zero padding)
    std::vector<float> h(NPAR, 0.0f);
    std::mt19937_64 rng(cfg->seed);
    std::normal_distribution<double> nd(0.0, 1.0);
    for_each_param([&](int64_t, int64_t d) {
        const int l = (int)(d / LP), o = (int)(d - (int64_t)l * LP);
        if (o < LW) h[(size_t)d] = (float)(nd(rng) * sqrt(2.0 / (9.0 * CIN[l])) * (l == NL - 1 ? cfg->init_output_gain : 1.0));
    });
    std::vector<float> z(NPAR, 0.0f);
    ok = hipMemcpy(c->d_par, h.data(), NPAR * sizeof(float), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(c->d_m, z.data(), NPAR * sizeof(float), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(c->d_v, z.data(), NPAR * sizeof(float), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(c->d_grad, z.data(), NPAR * sizeof(float), hipMemcpyHostToDevice) == hipSuccess;
    if (!ok) return fail(CDC_ERR_HIP, "closure parameter upload");
    if (refresh_weights(c, c->stream) || hipStreamSynchronize(c->stream) != hipSuccess)
        return fail(CDC_ERR_HIP, "closure weight conversion");
    *out = c;
    return CDC_OK;
}

void cdc_closure_destroy(cdc_closure_t *c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    free_closure(c);
    delete c;
}

int64_t cdc_closure_num_params(void) { return logical_params(); }

int32_t cdc_closure_get_params(cdc_closure_t *c, float *out) {
    if (!c || !out) return set_msg(CDC_ERR_INVALID_ARGUMENT, "null pointer");
    std::vector<float> h(NPAR);
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipMemcpy(h.data(), c->d_par, NPAR * sizeof(float), hipMemcpyDeviceToHost));
    for_each_param([&](int64_t k, int64_t d) { out[k] = h[(size_t)d]; });
    return CDC_OK;
}

int32_t cdc_closure_set_params(cdc_closure_t *c, const float *in) {
    if (!c || !in) return set_msg(CDC_ERR_INVALID_ARGUMENT, "null pointer");
    std::vector<float> h(NPAR, 0.0f);
    for_each_param([&](int64_t k, int64_t d) { h[(size_t)d] = in[k]; });
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipMemcpy(c->d_par, h.data(), NPAR * sizeof(float), hipMemcpyHostToDevice));
    std::vector<float> z(NPAR, 0.0f);
    HIP_TRY(hipMemcpy(c->d_m, z.data(), NPAR * sizeof(float), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->d_v, z.data(), NPAR * sizeof(float), hipMemcpyHostToDevice));
    c->adam_step = 0;
    int rc = refresh_weights(c, c->stream);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    return CDC_OK;
}

int32_t cdc_closure_get_config(cdc_closure_t *c, cdc_closure_config_t *out) {
    if (!c || !out) return set_msg(CDC_ERR_INVALID_ARGUMENT, "null pointer");
    *out = c->cfg;
    return CDC_OK;
}

int32_t cdc_closure_set_interval(cdc_closure_t *c, int32_t interval) {
    if (!c) return set_msg(CDC_ERR_INVALID_ARGUMENT, "null closure");
    if (interval < 1) return set_msg(CDC_ERR_INVALID_ARGUMENT, "interval must be >= 1");
    if (interval != c->interval) {
        if (c->stream) HIP_TRY(hipStreamSynchronize(c->stream));
        drop_fold_graphs(c);
        c->interval = interval;
    }
    return CDC_OK;
}

int32_t cdc_closure_set_scales(cdc_closure_t *c, const double *output_scale, const double *loss_weight) {
    if (!c) return set_msg(CDC_ERR_INVALID_ARGUMENT, "null closure");
    for (int v = 0; v < 3; ++v) {
        if (output_scale) c->cfg.output_scale[v] = output_scale[v];
        if (loss_weight) c->cfg.loss_weight[v] = loss_weight[v];
    }
    if (output_scale) {
        if (c->stream) HIP_TRY(hipStreamSynchronize(c->stream));
        drop_fold_graphs(c);
    }
    return CDC_OK;
}

int32_t cdc_closure_forward(cdc_closure_t *c, const cdc_solver_t *s, float *y_out) {
    if (!c) return set_msg(CDC_ERR_INVALID_ARGUMENT, "null closure");
    int rc = check_solver(c, s);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(s->stream));
    if ((rc = features(c, s, c->d_act[0], c->stream))) return rc;
    if ((rc = forward(c, c->stream))) return rc;
    if (y_out) {
        std::vector<float> h((size_t)c->P * YS);
        HIP_TRY(hipMemcpyAsync(h.data(), c->d_y, h.size() * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        for (int p = 0; p < c->P; ++p)
            for (int k = 0; k < 6; ++k) y_out[(size_t)p * 6 + k] = h[(size_t)p * YS + k];
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    return CDC_OK;
}

int32_t cdc_closure_features(cdc_closure_t *c, const cdc_solver_t *s, float *x_out) {
    if (!c || !x_out) return set_msg(CDC_ERR_INVALID_ARGUMENT, "null pointer");
    int rc = check_solver(c, s);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(s->stream));
    if ((rc = features(c, s, c->d_act[0], c->stream))) return rc;
    std::vector<uint16_t> h((size_t)c->P * C);
    HIP_TRY(hipMemcpyAsync(h.data(), c->d_act[0], h.size() * sizeof(uint16_t), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (int p = 0; p < c->P; ++p)
        for (int k = 0; k < 4; ++k) x_out[(size_t)p * 4 + k] = bf2f(h[(size_t)p * C + k]);
    return CDC_OK;
}

int32_t cdc_closure_add_sample(cdc_closure_t *c, const cdc_solver_t *base, const cdc_solver_t *plain,
                               const cdc_solver_t *truth, double dt) {
    if (!c) return set_msg(CDC_ERR_INVALID_ARGUMENT, "null closure");
    int rc;
    if ((rc = check_solver(c, base)) || (rc = check_solver(c, plain)) || (rc = check_solver(c, truth))) return rc;
    if (c->count >= c->cap) return set_msg(CDC_ERR_INVALID_ARGUMENT, "closure dataset is full");
    if (!(dt > 0.0)) return set_msg(CDC_ERR_INVALID_ARGUMENT, "dt must be > 0");
    HIP_TRY(hipStreamSynchronize(base->stream));
    HIP_TRY(hipStreamSynchronize(plain->stream));
    HIP_TRY(hipStreamSynchronize(truth->stream));
    const int k = c->count;
    if ((rc = features(c, base, c->d_feat + (size_t)k * c->P * C, c->stream))) return rc;
    TargetArgs a{};
    a.plain = plain->d_u[plain->cur];
    a.truth = truth->d_u[truth->cur];
    a.base = base->d_u[base->cur];
    a.mask = c->cfg.mask_density;
    a.t = c->d_tgt + (size_t)k * c->P * 4;
    a.part = c->d_part_l;
    a.N = c->n;
    a.nv = base->nv;
    a.inv_dt = 1.0 / dt;
    hipLaunchKernelGGL(target_kernel, dim3(c->nlb), dim3(256), 0, c->stream, a);
    HIP_TRY(hipGetLastError());
    std::vector<float> part((size_t)c->nlb * 3);
    HIP_TRY(hipMemcpyAsync(part.data(), c->d_part_l, part.size() * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (int b = 0; b < c->nlb; ++b)
        for (int v = 0; v < 3; ++v) c->tsq[v] += part[(size_t)b * 3 + v];
    c->count += 1;
    return CDC_OK;
}

int32_t cdc_closure_num_samples(cdc_closure_t *c, int32_t *count, double *target_ms) {
    if (!c) return set_msg(CDC_ERR_INVALID_ARGUMENT, "null closure");
    if (count) *count = c->count;
    if (target_ms)
        for (int v = 0; v < 3; ++v) target_ms[v] = c->count ? c->tsq[v] / ((double)c->count * c->P) : 0.0;
    return CDC_OK;
}

int32_t cdc_closure_clear_samples(cdc_closure_t *c) {
    if (!c) return set_msg(CDC_ERR_INVALID_ARGUMENT, "null closure");
    c->count = 0;
    for (int v = 0; v < 3; ++v) c->tsq[v] = 0.0;
    return CDC_OK;
}

int32_t cdc_closure_sample_grad(cdc_closure_t *c, int32_t k, double *loss_out, float *grad_out) {
    if (!c) return set_msg(CDC_ERR_INVALID_ARGUMENT, "null closure");
    if (k < 0 || k >= c->count) return set_msg(CDC_ERR_INVALID_ARGUMENT, "no such sample");
    int rc = ensure_hist(c, 1);
    if (rc) return rc;
    if ((rc = backprop(c, k, 0, c->stream))) return rc;
    std::vector<float> g(NPAR);
    float l = 0.0f;
    HIP_TRY(hipMemcpyAsync(g.data(), c->d_grad, NPAR * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(&l, c->d_hist, sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (loss_out) *loss_out = l;
    if (grad_out) for_each_param([&](int64_t i, int64_t d) { grad_out[i] = g[(size_t)d]; });
    return CDC_OK;
}

int32_t cdc_closure_train(cdc_closure_t *c, int32_t steps, double lr, uint64_t seed, float *loss_out) {
    if (!c) return set_msg(CDC_ERR_INVALID_ARGUMENT, "null closure");
    if (steps < 0) return set_msg(CDC_ERR_INVALID_ARGUMENT, "steps must be >= 0");
    if (c->count < 1) return set_msg(CDC_ERR_INVALID_ARGUMENT, "closure dataset is empty");
    int rc = ensure_hist(c, std::max(1, steps));
    if (rc) return rc;
    // samples in shuffled epochs
    std::mt19937_64 rng(seed);
    std::vector<int> order;
    for (int t = 0; t < steps; ++t) {
        if (order.empty()) {
            for (int k = 0; k < c->count; ++k) order.push_back(k);
            std::shuffle(order.begin(), order.end(), rng);
        }
        const int k = order.back();
        order.pop_back();
        if ((rc = backprop(c, k, t, c->stream))) return rc;
        if ((rc = adam_update(c, lr, c->stream))) return rc;
    }
    if (loss_out && steps > 0)
        HIP_TRY(hipMemcpyAsync(loss_out, c->d_hist, (size_t)steps * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return CDC_OK;
}

}  // extern "C"

namespace {

// The per-closure part of the limiter / correction arguments (the state
// pointer and dt come per step).
ApplyArgs apply_base(const cdc_closure *c, const cdc_solver *s) {
    ApplyArgs a{};
    a.y = c->d_y;
    a.lam = c->d_lam;
    a.N = s->N;
    const double dx = 2.0 * s->mesh.domain_radius / (double)s->N;
    a.idx = 1.0 / dx;
    a.idy = 1.0 / dx;
    a.kappa = c->cfg.limiter;
    a.mask = c->cfg.mask_density;
    a.nv = s->nv;
    a.gm = s->hydro.kind == CDC_EULER ? s->hydro.gamma_law_index : 1.0;
    for (int v = 0; v < 3; ++v) a.s[v] = c->cfg.output_scale[v];
    return a;
}

// The device work of a corrected advance of `fold` steps of rk stages each
// (arguments already staged): per step the net on its starting state, the
// RK stages, the limiter and the correction of the result; then the max
// signal speed of the final state into d_amax_adv and the read-back of it
// and of the error flag.  Fused, each stage launch also finalizes the stage
// before it; unfused (after an error), every stage is finalized before the
// next starts, so that a failed stage skips the rest.
int enqueue_fold(cdc_closure *c, cdc_solver *s, int fold, int rk, bool fuse) {
    const size_t n = (size_t)fold * rk;
    const unsigned grid = (unsigned)(((int64_t)s->N * s->N + 255) / 256);
    HIP_TRY(cdc::launch_reset_advance(s->d_err_flag, s->d_amax_adv, s->stream));
    HIP_TRY(hipMemcpyAsync(c->d_io, c->h_io, (size_t)fold * sizeof(StepIO), hipMemcpyHostToDevice, s->stream));
    int rc = cdc::host::upload_stages(s, n);
    if (rc) return rc;
    FeatArgs fa{};
    fa.x = c->d_act[0];
    fa.N = s->N;
    fa.nv = s->nv;
    fa.gm1 = s->hydro.kind == CDC_EULER ? s->hydro.gamma_law_index - 1.0 : 0.0;
    fa.a0 = c->cfg.feature_scale[0];
    fa.a1 = c->cfg.feature_scale[1];
    fa.a2 = c->cfg.feature_scale[2];
    feature_geometry(s, fa);
    const ApplyArgs ab = apply_base(c, s);
    for (int f = 0; f < fold; ++f) {
        // steps in groups of c->interval: the net on a group's starting
        // state, the correction after its last step
        const bool first = f % c->interval == 0;
        const bool last = f % c->interval == c->interval - 1 || f == fold - 1;
        if (first && fused_grid(c)) {
            hipLaunchKernelGGL(net_forward_fused_step_kernel, dim3((unsigned)(c->P / (TT * TT))), dim3(256), 0,
                               s->stream, fused_args(c), fa, c->d_io, f);
            HIP_TRY(hipGetLastError());
        } else if (first) {
            hipLaunchKernelGGL(features_step_kernel, dim3(grid), dim3(256), 0, s->stream, fa, c->d_io, f);
            HIP_TRY(hipGetLastError());
            if ((rc = forward_layers(c, s->stream))) return rc;
        }
        if ((rc = cdc::host::enqueue_stage_range(s, n, (size_t)f * rk, (size_t)(f + 1) * rk, fuse))) return rc;
        if (!last) continue;   // the stages wrote the result's edge columns themselves
        hipLaunchKernelGGL(limit_step_kernel, dim3(grid), dim3(256), 0, s->stream, ab, c->d_io, f);
        HIP_TRY(hipGetLastError());
        hipLaunchKernelGGL(apply_step_kernel, dim3(grid), dim3(256), 0, s->stream, ab, c->d_io, f);
        HIP_TRY(hipGetLastError());
        if (cdc::host::marching(s)) {
            const int64_t ne = (int64_t)s->nv * s->N * (s->N / 64) * 4;
            hipLaunchKernelGGL(edge_step_kernel, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, s->stream, c->d_io,
                               f, s->N, s->nv);
            HIP_TRY(hipGetLastError());
        }
    }
    const double gm = s->hydro.kind == CDC_EULER ? s->hydro.gamma_law_index : 1.0;
    if (s->nv == 3)
        hipLaunchKernelGGL(max_speed_step_kernel<3>, dim3(s->namax), dim3(256), 0, s->stream, c->d_io, fold - 1,
                           (int64_t)s->ncell(), gm, s->d_amax);
    else
        hipLaunchKernelGGL(max_speed_step_kernel<4>, dim3(s->namax), dim3(256), 0, s->stream, c->d_io, fold - 1,
                           (int64_t)s->ncell(), gm, s->d_amax);
    HIP_TRY(hipGetLastError());
    HIP_TRY(cdc::launch_column_max(s->d_amax, s->namax, 1, s->d_amax_adv, s->stream));
    HIP_TRY(hipMemcpyAsync(s->h_flag, s->d_err_flag, sizeof(int32_t), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipMemcpyAsync(s->h_amax, s->d_amax_adv, sizeof(double), hipMemcpyDeviceToHost, s->stream));
    return CDC_OK;
}

}  // namespace

extern "C" {

int32_t cdc_closure_advance_fold(cdc_closure_t *c, cdc_solver_t *s, int32_t fold, double *dt_out,
                                 cdc_error_t *err) {
    if (err) memset(err, 0, sizeof(*err));
    if (!c) return set_msg(CDC_ERR_INVALID_ARGUMENT, "null closure");
    int rc = check_solver(c, s);
    if (rc) return rc;
    if (s->nv < 3) return set_msg(CDC_ERR_INVALID_ARGUMENT, "unsupported hydro");
    if (fold < 0) return set_msg(CDC_ERR_INVALID_ARGUMENT, "fold must be >= 0");
    // one dt for the call, as cdc_solver_advance (src/scheme.rs:336-340)
    double amax = 0.0;
    if ((rc = cdc::host::compute_amax(s, &amax))) return rc;
    const double dt = cdc::host::step_dt(s, amax);
    if (dt_out) *dt_out = dt;
    if (fold == 0) return CDC_OK;

    cdc::host::StepPlan plan;
    cdc::host::plan_steps(s, fold, dt, plan);
    const int rk = s->phys.rk_order;
    const size_t n = plan.args.size();
    for (auto &a : plan.args) {
        a.want_amax = 0;   // the result's max speed is taken after its correction
        a.amax_out = nullptr;
        a.cool_factor = s->naux > 2 ? 1 : 0;
    }
    if ((rc = cdc::host::stage_args(s, plan.args.data(), n))) return rc;
    if ((rc = cdc::host::ensure_edges(s, s->stream))) return rc;
    if (fold > c->cap_io) {
        HIP_TRY(hipStreamSynchronize(s->stream));
        drop_fold_graphs(c);
        (void)hipFree(c->d_io);
        if (c->h_io) (void)hipHostFree(c->h_io);
        c->d_io = nullptr;
        c->h_io = nullptr;
        c->cap_io = 0;
        HIP_TRY(hipHostMalloc((void **)&c->h_io, (size_t)fold * sizeof(StepIO), hipHostMallocDefault));
        HIP_TRY(hipMalloc((void **)&c->d_io, (size_t)fold * sizeof(StepIO)));
        c->cap_io = fold;
    }
    for (int f = 0; f < fold; ++f) {
        c->h_io[f].u_in = plan.args[(size_t)f * rk].uc;
        c->h_io[f].u_out = plan.args[(size_t)f * rk + rk - 1].uo;
        // a group's last step corrects over the group's steps
        const int g0 = f - f % c->interval;
        const bool last = f % c->interval == c->interval - 1 || f == fold - 1;
        c->h_io[f].dt = last ? (double)(f - g0 + 1) * dt : dt;
        const cdc::Orbit &o = plan.args[(size_t)f * rk].orb;
        c->h_io[f].x1 = o.x1; c->h_io[f].y1 = o.y1; c->h_io[f].m1 = o.m1;
        c->h_io[f].x2 = o.x2; c->h_io[f].y2 = o.y2; c->h_io[f].m2 = o.m2;
        c->h_io[f].e_out = plan.args[(size_t)f * rk + rk - 1].eo;
    }

    // replay the captured call, or capture it
    bool done = false;
    if (c->use_graph) {
        for (const auto &g : c->graphs)
            if (g.uid == s->uid && g.epoch == s->graph_epoch && g.fold == fold && g.nstages == (int)n &&
                g.ts.tx == s->ts.tx &&
                g.ts.ty == s->ts.ty && g.ts.nt == s->ts.nt) {
                HIP_TRY(hipGraphLaunch(g.exec, s->stream));
                done = true;
                break;
            }
        if (!done) {
            if (hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
                (void)hipGetLastError();
                c->use_graph = false;
            } else {
                rc = enqueue_fold(c, s, fold, rk, true);
                hipGraph_t graph = nullptr;
                const hipError_t ec = hipStreamEndCapture(s->stream, &graph);
                hipGraphExec_t exec = nullptr;
                if (!rc && ec == hipSuccess && graph &&
                    hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess) {
                    (void)hipGraphDestroy(graph);
                    if (c->graphs.size() >= 4) {
                        (void)hipGraphExecDestroy(c->graphs.front().exec);
                        c->graphs.erase(c->graphs.begin());
                    }
                    c->graphs.push_back({s->uid, s->graph_epoch, fold, (int)n, s->ts, exec});
                    HIP_TRY(hipGraphLaunch(exec, s->stream));
                    done = true;
                } else {
                    if (graph) (void)hipGraphDestroy(graph);
                    (void)hipGetLastError();
                    if (rc) return rc;
                    c->use_graph = false;
                }
            }
        }
    }
    if (!done && (rc = enqueue_fold(c, s, fold, rk, true))) return rc;
    HIP_TRY(hipStreamSynchronize(s->stream));
    int32_t flag = *s->h_flag;
    if (flag != 0 && n > 1) {
        // repeated unfused from the untouched committed state, so that the
        // flag names the first failing stage (as cdc_solver_advance does)
        if ((rc = enqueue_fold(c, s, fold, rk, false))) return rc;
        HIP_TRY(hipStreamSynchronize(s->stream));
        flag = *s->h_flag;
    }
    if (flag != 0) return cdc::host::resolve_error(s, flag, plan, err);
    cdc::host::commit(s, plan);
    // the committed state is the corrected one: its max signal speed is the
    // one taken after the correction, not the last stage's
    s->amax_valid = false;
    s->amax_host = *s->h_amax;
    s->amax_host_version = s->version;
    return CDC_OK;
}

int32_t cdc_closure_advance(cdc_closure_t *c, cdc_solver_t *s, double *dt_out, cdc_error_t *err) {
    return cdc_closure_advance_fold(c, s, 1, dt_out, err);
}

}  // extern "C"
