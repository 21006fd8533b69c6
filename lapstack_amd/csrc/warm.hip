// warm.hip -- the learned warm start of the WLS solve: a small convolutional
// encoder-decoder that predicts the solution's means over 8 x 8 blocks of the
// WLS grid from the system's own planes averaged over the same blocks (peak
// depth as differences from the block's weighted mean, data weight, edge
// conductances, and coarser push-pull means); the
// start MG-PCG runs from is their bilinear prolongation, in place of the FGS
// sweeps.  Trained and run on gfx950 by the kernels here: implicit-GEMM 3x3
// convolutions on bf16 MFMA with f32 accumulation for the forward pass and
// both halves of the backward pass (input gradient, weight gradient), 2x2
// average pooling, nearest upsampling, a per-image relative squared error and
// Adam on f32 master weights.  No reference counterpart: the reference starts
// PCG from its FGS sweeps (depth.rs:1019-1034).
//
// Layout.  Activations are bf16, channel-last ([image][y][x][channel]) with a
// channel stride that may exceed the channel count, so the decoder's
// concatenations are views of one buffer (the upsampled half written at
// channel 0, the encoder's skip tensor at its offset).  Weights are bf16
// [tap][out channel][in channel] (tap = 3*(dy+1) + dx+1), channel counts padded
// to multiples of 16; the input-gradient pass runs the same kernel over the
// flipped, transposed copy [8-tap][in][out].
//
// Network (channels, on the coarse grid): 10 -> e0 16,16 -> pool -> e1 32,32 ->
// pool -> e2 64,64 -> pool -> e3 64,64 -> up ++ e2 -> d2 64 -> up ++ e1 -> d1
// 32 -> up ++ e0 -> d0 16 -> out 1; ReLU after every conv but the last; the
// predicted block mean (depth / dscale) is the weighted mean dm of the block's
// peak depths plus 1/32 of the output.  255,489 parameters.
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "api.hpp"
#include "common.hpp"

namespace lsh {

typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(uint16_t v) { return __uint_as_float((uint32_t)v << 16); }
__device__ __forceinline__ uint16_t f2bf(float f) {  // round to nearest even
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
static uint16_t f2bf_host(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// a channel-last bf16 tensor view: image b, pixel (y, x), channel c at
// p[((b*H + y)*W + x)*cs + c]
struct View {
    uint16_t *p;
    int H, W, C, cs;
};

constexpr int WB = 8;                  // the coarse grid's block edge, in WLS cells
constexpr int WNFEAT = 10;             // input channels
constexpr int WPP[3] = {4, 16, 64};    // the push-pull levels, in coarse cells
constexpr float WKAPPA = 1.0f / 32.0f; // the output's scale, in depth / dscale

// ------------------------------------------------------------------ conv
// Implicit GEMM on a 4-row x 64-pixel tile per workgroup: wave w owns row w of
// the tile (4 groups of 16 pixels) and NT*16 output channels (the output
// channel group blockIdx.z % groups).  The input channels go through in slices
// of KC: the (4+2) x 66 window of the slice is staged in LDS once for the 9
// taps and 4 waves; per tap, each B fragment (KC x 16 weights, read from
// global/L1) serves the wave's 4 pixel groups and each A fragment (16 pixels x
// KC channels, from LDS) serves the NT channel tiles.  KC = 32 runs the
// 16x16x32 bf16 MFMA, KC = 16 the 16x16x16 one.  Zero padding at the borders.
#define CONV_TR 4
#define CONV_PW 66
template <int KC, int NT>
__global__ __launch_bounds__(256) void k_conv(View in, const uint16_t *__restrict__ wt, const float *__restrict__ bias,
                                              int cop, int groups, View out, float *__restrict__ out32, View mask,
                                              int relu) {
    constexpr int PER = KC / 4, PS = KC + 8;  // bf16 per lane per MFMA; LDS pixel stride (padded)
    __shared__ __attribute__((aligned(16))) uint16_t Xs[(CONV_TR + 2) * CONV_PW * PS];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, m = lane & 15, kg = lane >> 4;
    const int cgi = blockIdx.z % groups, b = blockIdx.z / groups;
    const int y0 = blockIdx.y * CONV_TR, x0 = blockIdx.x * 64;
    const int H = in.H, W = in.W, cin = in.C;
    f32x4 acc[4][NT];
#pragma unroll
    for (int g = 0; g < 4; g++)
#pragma unroll
        for (int t = 0; t < NT; t++) acc[g][t] = f32x4{0.f, 0.f, 0.f, 0.f};
    const uint16_t *wbase = wt + ((size_t)cgi * NT * 16 + m) * cin + PER * kg;
    for (int ci0 = 0; ci0 < cin; ci0 += KC) {
        __syncthreads();
        for (int i = threadIdx.x; i < (CONV_TR + 2) * CONV_PW * (KC / 4); i += 256) {
            int r = i / (CONV_PW * (KC / 4)), q = i - r * (CONV_PW * (KC / 4)), px = q / (KC / 4), c4 = q - px * (KC / 4);
            int yy = y0 + r - 1, xx = x0 + px - 1;
            uint2 v = {0u, 0u};
            if (yy >= 0 && yy < H && xx >= 0 && xx < W)
                v = *(const uint2 *)(in.p + (((size_t)b * H + yy) * W + xx) * in.cs + ci0 + 4 * c4);
            *(uint2 *)(Xs + (r * CONV_PW + px) * PS + 4 * c4) = v;
        }
        __syncthreads();
        for (int tap = 0; tap < 9; tap++) {
            const int dy = tap / 3, dx = tap % 3;  // window offsets (0..2)
            const uint16_t *wsrc = wbase + (size_t)tap * cop * cin + ci0;
            const uint16_t *xrow = Xs + ((wv + dy) * CONV_PW + m + dx) * PS + PER * kg;
            if constexpr (KC == 16) {
                s16x4 bw[NT];
#pragma unroll
                for (int t = 0; t < NT; t++) bw[t] = *(const s16x4 *)(wsrc + (size_t)t * 16 * cin);
#pragma unroll
                for (int g = 0; g < 4; g++) {
                    s16x4 a = *(const s16x4 *)(xrow + g * 16 * PS);
#pragma unroll
                    for (int t = 0; t < NT; t++)
                        acc[g][t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bw[t], acc[g][t], 0, 0, 0);
                }
            } else {
                bf16x8 bw[NT];
#pragma unroll
                for (int t = 0; t < NT; t++) bw[t] = *(const bf16x8 *)(wsrc + (size_t)t * 16 * cin);
#pragma unroll
                for (int g = 0; g < 4; g++) {
                    bf16x8 a = *(const bf16x8 *)(xrow + g * 16 * PS);
#pragma unroll
                    for (int t = 0; t < NT; t++)
                        acc[g][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bw[t], acc[g][t], 0, 0, 0);
                }
            }
        }
    }
    // lane holds D[pixel g*16 + 4*kg + j][channel t*16 + m]
    const int y = y0 + wv;
    if (y >= H) return;
#pragma unroll
    for (int g = 0; g < 4; g++)
#pragma unroll
        for (int t = 0; t < NT; t++) {
            const int co = (cgi * NT + t) * 16 + m;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int x = x0 + g * 16 + 4 * kg + j;
                if (x >= W) continue;
                float v = acc[g][t][j] + (bias ? bias[co] : 0.0f);
                if (relu) v = fmaxf(v, 0.0f);
                const size_t pix = ((size_t)b * H + y) * W + x;
                if (out32) {
                    if (co == 0) out32[pix] = v;
                    continue;
                }
                if (co >= out.C) continue;
                if (mask.p && !(bf2f(mask.p[pix * mask.cs + co]) > 0.0f)) v = 0.0f;
                out.p[pix * out.cs + co] = f2bf(v);
            }
        }
}

// Weight gradient dW[tap][co][ci] = sum over pixels dz[p][co] * in[p + tap][ci]:
// a GEMM with the pixels as the reduction.  A workgroup stages a 64-pixel row
// segment of dz and the 3 x 66 input window in LDS and accumulates up to 16
// output tiles (4 per wave) over its share of the segments; each workgroup
// writes its partial tiles, a second kernel sums them in a fixed order.
__global__ __launch_bounds__(256) void k_conv_dw(View in, View dz, int nimg, int ntiles, int ncit, float *__restrict__ part) {
    extern __shared__ uint16_t sm[];
    const int cin = in.C, cop = dz.C, H = in.H, W = in.W;
    uint16_t *Zs = sm, *Xs = sm + 64 * cop;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, m = lane & 15, kg = lane >> 4;
    const int nxc = (W + 63) / 64;
    const long nch = (long)nimg * H * nxc;
    const int ncot = cop / 16;
    f32x4 acc[4];
    int tl[4], tdy[4], tdx[4], tco[4], tci[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        acc[i] = f32x4{0.f, 0.f, 0.f, 0.f};
        int t = blockIdx.y * 16 + wv * 4 + i;
        tl[i] = t;
        int tap = t / (ncot * ncit), r = t - tap * ncot * ncit;
        tdy[i] = tap / 3 - 1;
        tdx[i] = tap % 3 - 1;
        tco[i] = r / ncit;
        tci[i] = r - (r / ncit) * ncit;
    }
    for (long c = blockIdx.x; c < nch; c += gridDim.x) {
        const int b = (int)(c / ((long)H * nxc)), rr = (int)(c - (long)b * H * nxc), y = rr / nxc,
                  x0 = (rr - y * nxc) * 64;
        __syncthreads();
        for (int i = threadIdx.x; i < 64 * cop / 4; i += 256) {
            int px = i / (cop / 4), c4 = i - px * (cop / 4), x = x0 + px;
            uint2 v = {0u, 0u};
            if (x < W) v = *(const uint2 *)(dz.p + (((size_t)b * H + y) * W + x) * dz.cs + 4 * c4);
            *(uint2 *)(Zs + px * cop + 4 * c4) = v;
        }
        for (int i = threadIdx.x; i < 3 * 66 * cin / 4; i += 256) {
            int r = i / (66 * cin / 4), q = i - r * (66 * cin / 4), px = q / (cin / 4), c4 = q - px * (cin / 4);
            int yy = y + r - 1, xx = x0 + px - 1;
            uint2 v = {0u, 0u};
            if (yy >= 0 && yy < H && xx >= 0 && xx < W)
                v = *(const uint2 *)(in.p + (((size_t)b * H + yy) * W + xx) * in.cs + 4 * c4);
            *(uint2 *)(Xs + (r * 66 + px) * cin + 4 * c4) = v;
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 4; i++) {
            if (tl[i] >= ntiles) continue;
            for (int kk = 0; kk < 64; kk += 16) {
                s16x4 a, bb;
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    int px = kk + 4 * kg + j;
                    a[j] = (short)Zs[px * cop + tco[i] * 16 + m];
                    bb[j] = (short)Xs[((tdy[i] + 1) * 66 + px + tdx[i] + 1) * cin + tci[i] * 16 + m];
                }
                acc[i] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bb, acc[i], 0, 0, 0);
            }
        }
    }
    // lane holds D[co 4*kg + j][ci m] of each tile
#pragma unroll
    for (int i = 0; i < 4; i++) {
        if (tl[i] >= ntiles) continue;
        float *o = part + ((size_t)blockIdx.x * gridDim.y * 16 + tl[i]) * 256;
#pragma unroll
        for (int j = 0; j < 4; j++) o[(4 * kg + j) * 16 + m] = acc[i][j];
    }
}

// sum the partial tiles over the workgroups (in order) into g[tap][co][ci]
__global__ void k_dw_reduce(const float *__restrict__ part, int nparts, int tpad, int ntiles, int ncot, int ncit,
                            int cop, int cip, float *__restrict__ g) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ntiles * 256) return;
    int t = i >> 8, e = i & 255;
    float s = 0.0f;
    for (int p = 0; p < nparts; p++) s += part[((size_t)p * tpad + t) * 256 + e];
    int tap = t / (ncot * ncit), r = t - tap * ncot * ncit, cot = r / ncit, cit = r - cot * ncit;
    g[((size_t)tap * cop + cot * 16 + (e >> 4)) * cip + cit * 16 + (e & 15)] = s;
}

// bias gradient: per-workgroup channel sums of dz, then the ordered sum
__global__ __launch_bounds__(256) void k_bias_part(View dz, int nimg, float *__restrict__ part) {
    __shared__ float sh[256];
    const int cop = dz.C, per = 256 / cop, c = threadIdx.x % cop, lane = threadIdx.x / cop;
    const size_t npx = (size_t)nimg * dz.H * dz.W;
    float s = 0.0f;
    if (lane < per)
        for (size_t p = (size_t)blockIdx.x * per + lane; p < npx; p += (size_t)gridDim.x * per)
            s += bf2f(dz.p[p * dz.cs + c]);
    sh[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x < cop) {
        float t = 0.0f;
        for (int l = 0; l < per; l++) t += sh[l * cop + threadIdx.x];
        part[(size_t)blockIdx.x * cop + threadIdx.x] = t;
    }
}

__global__ void k_bias_reduce(const float *__restrict__ part, int nparts, int cop, float *__restrict__ g) {
    int c = threadIdx.x;
    if (c >= cop) return;
    float s = 0.0f;
    for (int p = 0; p < nparts; p++) s += part[(size_t)p * cop + c];
    g[c] = s;
}

// ------------------------------------------------------- pool, upsample, pack
// 2x2 average over the cells inside the finer grid (ceil sizes)
__global__ void k_pool(View in, View out, int nimg) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int C = out.C;
    size_t n = (size_t)nimg * out.H * out.W * C;
    if (i >= n) return;
    int c = (int)(i % C);
    size_t q = i / C;
    int x = (int)(q % out.W), y = (int)((q / out.W) % out.H), b = (int)(q / ((size_t)out.W * out.H));
    float s = 0.0f;
    int cnt = 0;
    for (int dy = 0; dy < 2; dy++)
        for (int dx = 0; dx < 2; dx++) {
            int yy = 2 * y + dy, xx = 2 * x + dx;
            if (yy < in.H && xx < in.W) {
                s += bf2f(in.p[(((size_t)b * in.H + yy) * in.W + xx) * in.cs + c]);
                cnt++;
            }
        }
    out.p[(((size_t)b * out.H + y) * out.W + x) * out.cs + c] = f2bf(s / (float)cnt);
}

// d(fine) += d(pooled) / count, kept where the fine activation is positive
__global__ void k_pool_bwd(View dp, View dfine, View fine, int nimg) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int C = dfine.C;
    size_t n = (size_t)nimg * dfine.H * dfine.W * C;
    if (i >= n) return;
    int c = (int)(i % C);
    size_t q = i / C;
    int x = (int)(q % dfine.W), y = (int)((q / dfine.W) % dfine.H), b = (int)(q / ((size_t)dfine.W * dfine.H));
    int Y = y / 2, X = x / 2;
    int cnt = (min(2 * Y + 2, dfine.H) - 2 * Y) * (min(2 * X + 2, dfine.W) - 2 * X);
    size_t fo = (((size_t)b * dfine.H + y) * dfine.W + x);
    float v = bf2f(dfine.p[fo * dfine.cs + c]);
    if (bf2f(fine.p[fo * fine.cs + c]) > 0.0f)
        v += bf2f(dp.p[(((size_t)b * dp.H + Y) * dp.W + X) * dp.cs + c]) / (float)cnt;
    dfine.p[fo * dfine.cs + c] = f2bf(v);
}

// nearest upsampling into the finer grid
__global__ void k_up(View in, View out, int nimg) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int C = in.C;
    size_t n = (size_t)nimg * out.H * out.W * C;
    if (i >= n) return;
    int c = (int)(i % C);
    size_t q = i / C;
    int x = (int)(q % out.W), y = (int)((q / out.W) % out.H), b = (int)(q / ((size_t)out.W * out.H));
    out.p[(((size_t)b * out.H + y) * out.W + x) * out.cs + c] =
        in.p[(((size_t)b * in.H + y / 2) * in.W + x / 2) * in.cs + c];
}

// its adjoint: the sum of the finer cells of each coarse cell
__global__ void k_up_bwd(View dfine, View dcoarse, int nimg) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int C = dcoarse.C;
    size_t n = (size_t)nimg * dcoarse.H * dcoarse.W * C;
    if (i >= n) return;
    int c = (int)(i % C);
    size_t q = i / C;
    int x = (int)(q % dcoarse.W), y = (int)((q / dcoarse.W) % dcoarse.H), b = (int)(q / ((size_t)dcoarse.W * dcoarse.H));
    float s = 0.0f;
    for (int dy = 0; dy < 2; dy++)
        for (int dx = 0; dx < 2; dx++) {
            int yy = 2 * y + dy, xx = 2 * x + dx;
            if (yy < dfine.H && xx < dfine.W) s += bf2f(dfine.p[(((size_t)b * dfine.H + yy) * dfine.W + xx) * dfine.cs + c]);
        }
    dcoarse.p[(((size_t)b * dcoarse.H + y) * dcoarse.W + x) * dcoarse.cs + c] = f2bf(s);
}

// ------------------------------------------------------------------ ends
// The network runs on the coarse grid of WB x WB blocks of the WLS grid and
// predicts the block means of the solution; the warm start is their bilinear
// prolongation.  Partial blocks at the right and bottom edges average the
// cells that exist.

// Per coarse cell, over the WLS cells of its block: dm = sum(w d) / sum(w) /
// dscale with the data weight w = max(conf, 0) + 1e-4, the mean of d / dscale,
// the mean of w, the means of the edge conductances exp(-|g - g_right| /
// sigma_c) (0 past the last column) and exp(-|g - g_below| / sigma_c) (0 past
// the last row) as the solver builds them, and with a target t the mean of
// t / dscale - dm.  cf holds the planes [dm, dmean, w, ax, ay, tc], each
// nimg*ch*cw, in f64: the inputs are rounded once, to bf16, as they enter the
// network.  A workgroup takes 256 columns of one block row of one image
// (grid: column groups, block rows, images): each thread sums its column over
// the block's rows, coalesced across the wave, and the 8 lanes of each block
// add their sums by shuffles.
__global__ __launch_bounds__(256) void k_feat_blocks(const float *__restrict__ d, const float *__restrict__ conf,
                                                     const float *__restrict__ g, const float *__restrict__ t, int nimg,
                                                     int h, int w, int ch, int cw, double inv_scale, float inv_sigma,
                                                     double *__restrict__ cf) {
    const size_t nc = (size_t)nimg * ch * cw;
    const int x = blockIdx.x * 256 + threadIdx.x, cy = blockIdx.y, b = blockIdx.z;
    const int y0 = cy * WB, y1 = min(y0 + WB, h);
    double s[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};  // w, w d, d, ax, ay, t
    if (x < w)
        for (int y = y0; y < y1; y++) {
            const size_t o = ((size_t)b * h + y) * w + x;
            const double wi = fmax((double)conf[o], 0.0) + 1e-4, di = d[o];
            s[0] += wi;
            s[1] += wi * di;
            s[2] += di;
            if (x + 1 < w) s[3] += expf(fabsf(g[o] - g[o + 1]) * inv_sigma);
            if (y + 1 < h) s[4] += expf(fabsf(g[o] - g[o + w]) * inv_sigma);
            if (t) s[5] += t[o];
        }
#pragma unroll
    for (int m = 1; m < WB; m <<= 1)
#pragma unroll
        for (int k = 0; k < 6; k++) s[k] += __shfl_xor(s[k], m);
    const int cx = x / WB;
    if ((threadIdx.x & (WB - 1)) || cx >= cw) return;
    const size_t i = ((size_t)b * ch + cy) * cw + cx;
    const int x0 = cx * WB, x1 = min(x0 + WB, w);
    const double inv_n = 1.0 / (double)((y1 - y0) * (x1 - x0));
    const double dm = s[1] / s[0] * inv_scale;
    cf[i] = dm;
    cf[nc + i] = s[2] * inv_n * inv_scale;
    cf[2 * nc + i] = s[0] * inv_n;
    cf[3 * nc + i] = s[3] * inv_n;
    cf[4 * nc + i] = s[4] * inv_n;
    if (t) cf[5 * nc + i] = s[5] * inv_n * inv_scale - dm;
}

// A push-pull level: per cell of the grid 4, 16 or 64 times coarser than the
// block grid (4 x 4 cells of the level before, or of the blocks), the sums of
// w and w dm over the blocks inside and their count, then log10 of the mean
// weight and the weighted mean.  src: the level before's planes [sw, swd,
// cnt, ...] (nimg*ph*pw each), or null to sum the blocks (cf).  out: this
// level's planes [sw, swd, cnt, lpk, mk], nimg*kh*kw each.
__global__ void k_feat_pp(const double *__restrict__ cf, const double *__restrict__ src, int nimg, int ch, int cw,
                          int ph, int pw, int kh, int kw, double *__restrict__ out) {
    const size_t np = (size_t)nimg * kh * kw;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const int kx = (int)(i % kw), ky = (int)((i / kw) % kh), b = (int)(i / ((size_t)kw * kh));
    double sw = 0.0, swd = 0.0, cnt = 0.0;
    if (!src) {
        const size_t nc = (size_t)nimg * ch * cw;
        for (int y = ky * 4; y < min(ky * 4 + 4, ch); y++)
            for (int x = kx * 4; x < min(kx * 4 + 4, cw); x++) {
                const size_t j = ((size_t)b * ch + y) * cw + x;
                const double v = cf[2 * nc + j];
                sw += v;
                swd += v * cf[j];
                cnt += 1.0;
            }
    } else {
        const size_t pn = (size_t)nimg * ph * pw;
        for (int y = ky * 4; y < min(ky * 4 + 4, ph); y++)
            for (int x = kx * 4; x < min(kx * 4 + 4, pw); x++) {
                const size_t j = ((size_t)b * ph + y) * pw + x;
                sw += src[j];
                swd += src[pn + j];
                cnt += src[2 * pn + j];
            }
    }
    out[i] = sw;
    out[np + i] = swd;
    out[2 * np + i] = cnt;
    out[3 * np + i] = log10(sw / cnt);
    out[4 * np + i] = swd / sw;
}

// bilinear resampling of the (ih, iw) plane p at the cell (oy, ox) of a grid
// over the same area, ry = ih / oh and rx = iw / ow its scales: source
// position (o + 1/2) r - 1/2, clamped at 0, its right / lower neighbour
// clamped to the last cell
template <class P>
__device__ __forceinline__ double resample(P p, int ih, int iw, double ry, double rx, int oy, int ox) {
    const double sy = fmax(((double)oy + 0.5) * ry - 0.5, 0.0);
    const double sx = fmax(((double)ox + 0.5) * rx - 0.5, 0.0);
    const int y0 = (int)sy, x0 = (int)sx;
    const int y1 = y0 + (y0 < ih - 1 ? 1 : 0), x1 = x0 + (x0 < iw - 1 ? 1 : 0);
    const double ly = sy - (double)y0, lx = sx - (double)x0;
    return (1.0 - ly) * ((1.0 - lx) * p((size_t)y0 * iw + x0) + lx * p((size_t)y0 * iw + x1)) +
           ly * ((1.0 - lx) * p((size_t)y1 * iw + x0) + lx * p((size_t)y1 * iw + x1));
}

// the push-pull levels of a batch: per level l the planes log10 mean weight
// and weighted mean dm, each nimg * kh[l] * kw[l]
struct PushPull {
    const double *lpk[3], *mk[3];
    int kh[3], kw[3];
    double ry[3], rx[3];  // kh / ch, kw / cw
};

// The network's input on the coarse grid.  Depth enters only as differences
// from the block's own weighted mean dm, in units of WKAPPA (the solution's
// deviation from dm does not change when a constant is added to the depth,
// and differences of nearly equal values would not survive bf16): dmean - dm,
// then per push-pull level its weighted mean resampled to the coarse grid
// minus dm; then w, log10(w) / 4 + 1, ax, ay, and per push-pull level log10
// of its mean weight / 4 + 1 resampled.  Channels 10..15 zero.
__global__ void k_feat_pack(const double *__restrict__ cf, PushPull pp, int nimg, int ch, int cw, View x) {
    const size_t nc = (size_t)nimg * ch * cw;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    const int cx = (int)(i % cw), cy = (int)((i / cw) % ch), b = (int)(i / ((size_t)cw * ch));
    const double dm = cf[i], pw = cf[2 * nc + i], r = 1.0 / (double)WKAPPA;
    double v[WNFEAT];
    v[0] = (cf[nc + i] - dm) * r;
    v[4] = pw;
    v[5] = log10(pw) * 0.25 + 1.0;
    v[6] = cf[3 * nc + i];
    v[7] = cf[4 * nc + i];
#pragma unroll
    for (int l = 0; l < 3; l++) {
        const size_t o = (size_t)b * pp.kh[l] * pp.kw[l];
        const double *mk = pp.mk[l] + o, *lpk = pp.lpk[l] + o;
        v[1 + l] = (resample([&](size_t j) { return mk[j]; }, pp.kh[l], pp.kw[l], pp.ry[l], pp.rx[l], cy, cx) - dm) * r;
        if (l < 2)
            v[8 + l] =
                resample([&](size_t j) { return lpk[j]; }, pp.kh[l], pp.kw[l], pp.ry[l], pp.rx[l], cy, cx) * 0.25 + 1.0;
    }
    uint16_t *o = x.p + i * x.cs;
#pragma unroll
    for (int c = 0; c < WNFEAT; c++) o[c] = f2bf((float)v[c]);
#pragma unroll
    for (int c = WNFEAT; c < 16; c++) o[c] = 0;
}

// the warm start on the WLS grid: dscale times the bilinear prolongation of
// the predicted block means dm + WKAPPA out
__global__ void k_prolong(const double *__restrict__ dm, const float *__restrict__ out32, int nimg, int ch, int cw,
                          int h, int w, double ry, double rx, double scale, float *__restrict__ u) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)nimg * h * w) return;
    const int x = (int)(i % w), y = (int)((i / w) % h), b = (int)(i / ((size_t)w * h));
    const size_t o = (size_t)b * ch * cw;
    const double v = resample([&](size_t j) { return dm[o + j] + (double)WKAPPA * (double)out32[o + j]; }, ch, cw, ry, rx,
                              y, x);
    u[i] = (float)(scale * v);
}

// the loss weight of each image: 1 / max(mean((tc / WKAPPA)^2), 1e-4), so
// that every image's error counts relative to its own target's size
__global__ __launch_bounds__(256) void k_loss_weight(const double *__restrict__ tc, int npx, float *__restrict__ wimg) {
    __shared__ double sh[256];
    const double *p = tc + (size_t)blockIdx.x * npx;
    double s = 0.0;
    for (int i = threadIdx.x; i < npx; i += 256) {
        const double v = p[i] * (1.0 / (double)WKAPPA);
        s += v * v;
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) wimg[blockIdx.x] = (float)(1.0 / fmax(sh[0] / npx, 1e-4));
}

// The loss: the mean over the images of wimg * mean((out - tc / WKAPPA)^2),
// its gradient into channel 0 of dout (bf16); per-block f64 partial sums
__global__ __launch_bounds__(256) void k_loss(const float *__restrict__ out32, const double *__restrict__ tc,
                                              const float *__restrict__ wimg, int npx, size_t n, float gscale,
                                              View dout, double *__restrict__ part) {
    __shared__ double sh[256];
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    double s = 0.0;
    if (i < n) {
        const float wi = wimg[i / npx];
        const float e = out32[i] - (float)(tc[i] * (1.0 / (double)WKAPPA));
        s = (double)wi * (double)e * (double)e;
        uint16_t *o = dout.p + i * dout.cs;
        o[0] = f2bf(gscale * wi * e);
#pragma unroll
        for (int c = 1; c < 16; c++) o[c] = 0;
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// Adam on the f32 master weights (bias-corrected step sizes from the host)
__global__ void k_adam(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m, float *__restrict__ v,
                       size_t n, float lr_t, float b1, float b2, float eps_t) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float gi = g[i];
    float mi = b1 * m[i] + (1.0f - b1) * gi;
    float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr_t * mi / (sqrtf(vi) + eps_t);
}

// bf16 copies of one layer's weights: forward [tap][co][ci], and the flipped
// transpose [8 - tap][ci][co] for the input gradient
__global__ void k_cast_w(const float *__restrict__ w, int cop, int cip, uint16_t *__restrict__ wf,
                         uint16_t *__restrict__ wt) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 9 * cop * cip) return;
    int tap = i / (cop * cip), r = i - tap * cop * cip, co = r / cip, ci = r - co * cip;
    uint16_t v = f2bf(w[i]);
    wf[i] = v;
    wt[((size_t)(8 - tap) * cip + ci) * cop + co] = v;
}

// ------------------------------------------------------------------ host
struct LayerSpec {
    int cin, cout;
    bool relu;
};
static const LayerSpec SPECS[12] = {{WNFEAT, 16, true},   {16, 16, true},  {16, 32, true}, {32, 32, true},
                                    {32, 64, true},  {64, 64, true},  {64, 64, true}, {64, 64, true},
                                    {128, 64, true}, {96, 32, true},  {48, 16, true}, {16, 1, false}};
enum { E0A, E0B, E1A, E1B, E2A, E2B, E3A, E3B, D2, D1, D0, OUT, NLAYER };

static int pad16(int c) { return (c + 15) / 16 * 16; }

struct Layer {
    int cin, cout, cip, cop;
    bool relu;
    size_t w_off, b_off;  // into the padded master arrays
    size_t n_nat;         // natural parameters (W[co][ci][ky][kx], then b[co])
    DBuf<uint16_t> wf, wt;
};

struct Acts {
    int nimg = 0, H[4] = {0, 0, 0, 0}, W[4] = {0, 0, 0, 0};
    bool bwd = false;
    DBuf<uint16_t> X0, A0a, C0, P0, A1a, C1, P1, A2a, C2, P2, A3a, B3, D2b, D1b, D0b;
    DBuf<float> out32;
    DBuf<uint16_t> dOUT, dD0, dC0, dD1, dC1, dD2, dC2, dB3, dA3a, dP2, dA2a, dP1, dA1a, dP0, dA0a;
};

}  // namespace lsh

struct lsh_warm {
    std::vector<lsh::Layer> L;
    float sigma_c = 0.04f;  // the guide scale of the WLS systems served
    lsh::DBuf<double> cf, pp;  // block planes, push-pull levels
    lsh::DBuf<float> wimg;     // loss weights
    size_t npad = 0, nnat = 0;
    lsh::DBuf<float> w32, g32, m, v;
    uint64_t step = 0;
    lsh::Acts A;
    lsh::DBuf<float> part;
    lsh::DBuf<double> lpart;
    lsh::DBuf<float> dbuf, cbuf, gbuf, tbuf, ubuf;  // host-memory inputs staged on the device
};

namespace lsh {

static View view(DBuf<uint16_t> &b, int H, int W, int C, int cs, int off = 0) { return View{b.p + off, H, W, C, cs}; }
static View noview() { return View{nullptr, 0, 0, 0, 0}; }

static double splitmix(uint64_t &s) {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

static void recast(lsh_warm &n) {
    for (auto &l : n.L)
        launch("warm_cast", k_cast_w, dim3(cdiv(9 * l.cop * l.cip, 256)), dim3(256), 0, (const float *)(n.w32.p + l.w_off),
               l.cop, l.cip, l.wf.p, l.wt.p);
}

// natural <-> padded parameter layouts (host)
static void nat_to_pad(const lsh_warm &n, const float *nat, std::vector<float> &pad) {
    pad.assign(n.npad, 0.0f);
    size_t o = 0;
    for (auto &l : n.L) {
        for (int co = 0; co < l.cout; co++)
            for (int ci = 0; ci < l.cin; ci++)
                for (int t = 0; t < 9; t++) pad[l.w_off + ((size_t)t * l.cop + co) * l.cip + ci] = nat[o++];
        for (int co = 0; co < l.cout; co++) pad[l.b_off + co] = nat[o++];
    }
}
static void pad_to_nat(const lsh_warm &n, const std::vector<float> &pad, float *nat) {
    size_t o = 0;
    for (auto &l : n.L) {
        for (int co = 0; co < l.cout; co++)
            for (int ci = 0; ci < l.cin; ci++)
                for (int t = 0; t < 9; t++) nat[o++] = pad[l.w_off + ((size_t)t * l.cop + co) * l.cip + ci];
        for (int co = 0; co < l.cout; co++) nat[o++] = pad[l.b_off + co];
    }
}

static void ensure_acts(lsh_warm &n, int nimg, int H, int W, bool bwd) {
    Acts &a = n.A;
    if (a.nimg == nimg && a.H[0] == H && a.W[0] == W && (a.bwd || !bwd)) return;
    a = Acts();
    a.nimg = nimg;
    a.bwd = bwd;
    for (int l = 0; l < 4; l++) {
        a.H[l] = l ? (a.H[l - 1] + 1) / 2 : H;
        a.W[l] = l ? (a.W[l - 1] + 1) / 2 : W;
    }
    auto px = [&](int l) { return (size_t)nimg * a.H[l] * a.W[l]; };
    a.X0.alloc(px(0) * 16);
    a.A0a.alloc(px(0) * 16);
    a.C0.alloc(px(0) * 48);
    a.P0.alloc(px(1) * 16);
    a.A1a.alloc(px(1) * 32);
    a.C1.alloc(px(1) * 96);
    a.P1.alloc(px(2) * 32);
    a.A2a.alloc(px(2) * 64);
    a.C2.alloc(px(2) * 128);
    a.P2.alloc(px(3) * 64);
    a.A3a.alloc(px(3) * 64);
    a.B3.alloc(px(3) * 64);
    a.D2b.alloc(px(2) * 64);
    a.D1b.alloc(px(1) * 32);
    a.D0b.alloc(px(0) * 16);
    a.out32.alloc(px(0));
    if (bwd) {
        a.dOUT.alloc(px(0) * 16);
        a.dD0.alloc(px(0) * 16);
        a.dC0.alloc(px(0) * 48);
        a.dD1.alloc(px(1) * 32);
        a.dC1.alloc(px(1) * 96);
        a.dD2.alloc(px(2) * 64);
        a.dC2.alloc(px(2) * 128);
        a.dB3.alloc(px(3) * 64);
        a.dA3a.alloc(px(3) * 64);
        a.dP2.alloc(px(3) * 64);
        a.dA2a.alloc(px(2) * 64);
        a.dP1.alloc(px(2) * 32);
        a.dA1a.alloc(px(1) * 32);
        a.dP0.alloc(px(1) * 16);
        a.dA0a.alloc(px(0) * 16);
    }
}

template <int KC>
static void conv_kc(const View &in, const uint16_t *w, const float *bias, int cop, const View &out, float *out32,
                    const View &mask, bool relu, int nimg) {
    // output channels in groups of at most 64 (NT <= 4 tiles per wave)
    const int tiles = cop / 16, nt = tiles % 4 == 0 ? 4 : (tiles % 3 == 0 ? 3 : (tiles <= 4 ? tiles : 2));
    const int groups = tiles / nt;
    LSH_REQUIRE(groups * nt == tiles, "warm: unsupported output width");
    dim3 grid(cdiv(in.W, 64), cdiv(in.H, CONV_TR), nimg * groups), block(256);
    double bytes = 2.0 * nimg * in.H * in.W * (in.C + (out32 ? 2 : out.C)) + 2.0 * 9 * cop * in.C;
    switch (nt) {
    case 1: launch_b("warm_conv", bytes, k_conv<KC, 1>, grid, block, 0, in, w, bias, cop, groups, out, out32, mask, (int)relu); break;
    case 2: launch_b("warm_conv", bytes, k_conv<KC, 2>, grid, block, 0, in, w, bias, cop, groups, out, out32, mask, (int)relu); break;
    case 3: launch_b("warm_conv", bytes, k_conv<KC, 3>, grid, block, 0, in, w, bias, cop, groups, out, out32, mask, (int)relu); break;
    default: launch_b("warm_conv", bytes, k_conv<KC, 4>, grid, block, 0, in, w, bias, cop, groups, out, out32, mask, (int)relu); break;
    }
}

static void conv(const View &in, const uint16_t *w, const float *bias, int cop, const View &out, float *out32,
                 const View &mask, bool relu, int nimg) {
    LSH_REQUIRE(in.C % 16 == 0 && in.cs % 16 == 0, "warm: input channels not a multiple of 16");
    if (in.C % 32 == 0 && in.cs % 8 == 0)
        conv_kc<32>(in, w, bias, cop, out, out32, mask, relu, nimg);
    else
        conv_kc<16>(in, w, bias, cop, out, out32, mask, relu, nimg);
}

// an element-wise pass over n outputs of bf16 (2 bytes): bpe the bytes it
// moves per output (pooling and the upsampling's adjoint read four inputs per
// output, 10; the upsampling reads one per four outputs, 2.5)
static void ew(const char *name, size_t n, double bpe, void (*k)(View, View, int), View a, View b, int nimg) {
    launch_b(name, bpe * (double)n, k, dim3(cdiv(n, 256)), dim3(256), 0, a, b, nimg);
}

// the coarse grid of an h x w WLS grid
static int cdim(int v) { return (v + WB - 1) / WB; }

// the block planes and the packed input of a batch of nimg h x w systems (the
// activations sized for the coarse grid first); with a target t also its
// block means
static void features(lsh_warm &n, const float *d, const float *conf, const float *g, const float *t, int nimg, int h,
                     int w, float dscale, float sigma_c) {
    Acts &a = n.A;
    const int ch = a.H[0], cw = a.W[0];
    const size_t nc = (size_t)nimg * ch * cw;
    PushPull pp;
    size_t np[3], tot = 0;
    for (int l = 0; l < 3; l++) {
        pp.kh[l] = (int)cdiv(ch, WPP[l]);
        pp.kw[l] = (int)cdiv(cw, WPP[l]);
        pp.ry[l] = (double)pp.kh[l] / (double)ch;
        pp.rx[l] = (double)pp.kw[l] / (double)cw;
        np[l] = (size_t)nimg * pp.kh[l] * pp.kw[l];
        tot += 5 * np[l];
    }
    n.cf.ensure(6 * nc);
    n.pp.ensure(tot);
    launch_b("warm_feat", (t ? 16.0 : 12.0) * nimg * h * w + (t ? 48.0 : 40.0) * nc, k_feat_blocks, dim3(cdiv(w, 256), ch, nimg),
             dim3(256), 0, d, conf, g, t, nimg, h, w, ch, cw, 1.0 / dscale, -1.0f / sigma_c, n.cf.p);
    double *q = n.pp.p;
    const double *prev = nullptr;
    for (int l = 0; l < 3; l++) {
        static_assert(WPP[0] == 4 && WPP[1] == 16 && WPP[2] == 64, "push-pull levels 4 times apart");
        launch("warm_feat", k_feat_pp, dim3(cdiv(np[l], 256)), dim3(256), 0, (const double *)n.cf.p, prev, nimg, ch, cw,
               l ? pp.kh[l - 1] : 0, l ? pp.kw[l - 1] : 0, pp.kh[l], pp.kw[l], q);
        pp.lpk[l] = q + 3 * np[l];
        pp.mk[l] = q + 4 * np[l];
        prev = q;
        q += 5 * np[l];
    }
    launch("warm_feat", k_feat_pack, dim3(cdiv(nc, 256)), dim3(256), 0, (const double *)n.cf.p, pp, nimg, ch, cw,
           view(a.X0, ch, cw, 16, 16));
}

// the network on the packed input X0
static void forward(lsh_warm &n, int nimg) {
    Acts &a = n.A;
    const int *H = a.H, *W = a.W;
    auto &L = n.L;
    View X0 = view(a.X0, H[0], W[0], 16, 16);
    const float *b32 = n.w32.p;
    auto B = [&](int l) { return b32 + L[l].b_off; };
    View A0a = view(a.A0a, H[0], W[0], 16, 16), S0 = view(a.C0, H[0], W[0], 16, 48, 32);
    View C0 = view(a.C0, H[0], W[0], 48, 48), U0 = view(a.C0, H[0], W[0], 32, 48, 0);
    View P0 = view(a.P0, H[1], W[1], 16, 16), A1a = view(a.A1a, H[1], W[1], 32, 32);
    View S1 = view(a.C1, H[1], W[1], 32, 96, 64), C1 = view(a.C1, H[1], W[1], 96, 96), U1 = view(a.C1, H[1], W[1], 64, 96, 0);
    View P1 = view(a.P1, H[2], W[2], 32, 32), A2a = view(a.A2a, H[2], W[2], 64, 64);
    View S2 = view(a.C2, H[2], W[2], 64, 128, 64), C2 = view(a.C2, H[2], W[2], 128, 128), U2 = view(a.C2, H[2], W[2], 64, 128, 0);
    View P2 = view(a.P2, H[3], W[3], 64, 64), A3a = view(a.A3a, H[3], W[3], 64, 64), B3 = view(a.B3, H[3], W[3], 64, 64);
    View D2v = view(a.D2b, H[2], W[2], 64, 64), D1v = view(a.D1b, H[1], W[1], 32, 32), D0v = view(a.D0b, H[0], W[0], 16, 16);
    View none = noview();
    auto px = [&](int l) { return (size_t)nimg * H[l] * W[l]; };
    conv(X0, L[E0A].wf.p, B(E0A), L[E0A].cop, A0a, nullptr, none, true, nimg);
    conv(A0a, L[E0B].wf.p, B(E0B), L[E0B].cop, S0, nullptr, none, true, nimg);
    ew("warm_pool", px(1) * 16, 10.0, k_pool, S0, P0, nimg);
    conv(P0, L[E1A].wf.p, B(E1A), L[E1A].cop, A1a, nullptr, none, true, nimg);
    conv(A1a, L[E1B].wf.p, B(E1B), L[E1B].cop, S1, nullptr, none, true, nimg);
    ew("warm_pool", px(2) * 32, 10.0, k_pool, S1, P1, nimg);
    conv(P1, L[E2A].wf.p, B(E2A), L[E2A].cop, A2a, nullptr, none, true, nimg);
    conv(A2a, L[E2B].wf.p, B(E2B), L[E2B].cop, S2, nullptr, none, true, nimg);
    ew("warm_pool", px(3) * 64, 10.0, k_pool, S2, P2, nimg);
    conv(P2, L[E3A].wf.p, B(E3A), L[E3A].cop, A3a, nullptr, none, true, nimg);
    conv(A3a, L[E3B].wf.p, B(E3B), L[E3B].cop, B3, nullptr, none, true, nimg);
    ew("warm_up", px(2) * 64, 2.5, k_up, B3, U2, nimg);
    conv(C2, L[D2].wf.p, B(D2), L[D2].cop, D2v, nullptr, none, true, nimg);
    ew("warm_up", px(1) * 64, 2.5, k_up, D2v, U1, nimg);
    conv(C1, L[D1].wf.p, B(D1), L[D1].cop, D1v, nullptr, none, true, nimg);
    ew("warm_up", px(0) * 32, 2.5, k_up, D1v, U0, nimg);
    conv(C0, L[D0].wf.p, B(D0), L[D0].cop, D0v, nullptr, none, true, nimg);
    View dummy = View{nullptr, H[0], W[0], 1, 1};
    conv(D0v, L[OUT].wf.p, B(OUT), L[OUT].cop, dummy, a.out32.p, none, false, nimg);
}

static void weight_grad(lsh_warm &n, int l, const View &in, const View &dz, int nimg) {
    Layer &ly = n.L[l];
    const int ncot = ly.cop / 16, ncit = ly.cip / 16, ntiles = 9 * ncot * ncit, groups = cdiv(ntiles, 16);
    LSH_REQUIRE(in.C == ly.cip && dz.C == ly.cop, "warm: weight-gradient shapes");
    const long nch = (long)nimg * in.H * cdiv(in.W, 64);
    const int parts = (int)std::min<long>(nch, 128);
    const size_t npx = (size_t)nimg * dz.H * dz.W;
    const int bparts = (int)std::min<size_t>(cdiv(npx, 256), 256);
    n.part.ensure(std::max((size_t)parts * groups * 16 * 256, (size_t)bparts * ly.cop));
    size_t shm = (size_t)(64 * ly.cop + 3 * 66 * ly.cip) * sizeof(uint16_t);
    launch_b("warm_conv_dw", 2.0 * nimg * in.H * in.W * (in.C + dz.C), k_conv_dw, dim3(parts, groups), dim3(256), shm,
             in, dz, nimg, ntiles, ncit, n.part.p);
    launch("warm_dw_reduce", k_dw_reduce, dim3(cdiv(ntiles * 256, 256)), dim3(256), 0, (const float *)n.part.p, parts,
           groups * 16, ntiles, ncot, ncit, ly.cop, ly.cip, n.g32.p + ly.w_off);
    launch("warm_bias_grad", k_bias_part, dim3(bparts), dim3(256), 0, dz, nimg, n.part.p);
    launch("warm_bias_grad", k_bias_reduce, dim3(1), dim3(128), 0, (const float *)n.part.p, bparts, ly.cop,
           n.g32.p + ly.b_off);
}

// loss and gradients of the batch (features with the target and forward
// first); returns the loss
static double backward(lsh_warm &n, int nimg) {
    Acts &a = n.A;
    const int *H = a.H, *W = a.W;
    auto &L = n.L;
    const size_t n0 = (size_t)nimg * H[0] * W[0];
    auto px = [&](int l) { return (size_t)nimg * H[l] * W[l]; };
    View dOUT = view(a.dOUT, H[0], W[0], 16, 16);
    const int lb = (int)cdiv(n0, 256);
    const int npx = H[0] * W[0];
    const double *tc = n.cf.p + 5 * n0;
    n.lpart.ensure(lb);
    n.wimg.ensure(nimg);
    launch("warm_loss", k_loss_weight, dim3(nimg), dim3(256), 0, tc, npx, n.wimg.p);
    launch("warm_loss", k_loss, dim3(lb), dim3(256), 0, (const float *)a.out32.p, tc, (const float *)n.wimg.p, npx, n0,
           2.0f / (float)n0, dOUT, n.lpart.p);
    View X0 = view(a.X0, H[0], W[0], 16, 16);
    View A0a = view(a.A0a, H[0], W[0], 16, 16), S0 = view(a.C0, H[0], W[0], 16, 48, 32), C0 = view(a.C0, H[0], W[0], 48, 48);
    View P0 = view(a.P0, H[1], W[1], 16, 16), A1a = view(a.A1a, H[1], W[1], 32, 32);
    View S1 = view(a.C1, H[1], W[1], 32, 96, 64), C1 = view(a.C1, H[1], W[1], 96, 96);
    View P1 = view(a.P1, H[2], W[2], 32, 32), A2a = view(a.A2a, H[2], W[2], 64, 64);
    View S2 = view(a.C2, H[2], W[2], 64, 128, 64), C2 = view(a.C2, H[2], W[2], 128, 128);
    View P2 = view(a.P2, H[3], W[3], 64, 64), A3a = view(a.A3a, H[3], W[3], 64, 64);
    // the decoder's own outputs need no mask here: their ReLU masks act through
    // the concatenations they were upsampled into (C0, C1) and D0 directly
    View D0v = view(a.D0b, H[0], W[0], 16, 16);
    View dD0 = view(a.dD0, H[0], W[0], 16, 16), dC0 = view(a.dC0, H[0], W[0], 48, 48);
    View dU0 = view(a.dC0, H[0], W[0], 32, 48, 0), dS0 = view(a.dC0, H[0], W[0], 16, 48, 32);
    View dD1 = view(a.dD1, H[1], W[1], 32, 32), dC1 = view(a.dC1, H[1], W[1], 96, 96);
    View dU1 = view(a.dC1, H[1], W[1], 64, 96, 0), dS1 = view(a.dC1, H[1], W[1], 32, 96, 64);
    View dD2 = view(a.dD2, H[2], W[2], 64, 64), dC2 = view(a.dC2, H[2], W[2], 128, 128);
    View dU2 = view(a.dC2, H[2], W[2], 64, 128, 0), dS2 = view(a.dC2, H[2], W[2], 64, 128, 64);
    View dB3 = view(a.dB3, H[3], W[3], 64, 64), dA3a = view(a.dA3a, H[3], W[3], 64, 64), dP2 = view(a.dP2, H[3], W[3], 64, 64);
    View dA2a = view(a.dA2a, H[2], W[2], 64, 64), dP1 = view(a.dP1, H[2], W[2], 32, 32);
    View dA1a = view(a.dA1a, H[1], W[1], 32, 32), dP0 = view(a.dP0, H[1], W[1], 16, 16), dA0a = view(a.dA0a, H[0], W[0], 16, 16);
    View none = noview();
    // out layer (no ReLU): dz = dOUT
    weight_grad(n, OUT, D0v, dOUT, nimg);
    conv(dOUT, L[OUT].wt.p, nullptr, L[OUT].cip, dD0, nullptr, D0v, false, nimg);
    weight_grad(n, D0, C0, dD0, nimg);
    conv(dD0, L[D0].wt.p, nullptr, L[D0].cip, dC0, nullptr, C0, false, nimg);
    ew("warm_up_bwd", px(1) * 32, 10.0, k_up_bwd, dU0, dD1, nimg);
    weight_grad(n, D1, C1, dD1, nimg);
    conv(dD1, L[D1].wt.p, nullptr, L[D1].cip, dC1, nullptr, C1, false, nimg);
    ew("warm_up_bwd", px(2) * 64, 10.0, k_up_bwd, dU1, dD2, nimg);
    weight_grad(n, D2, C2, dD2, nimg);
    conv(dD2, L[D2].wt.p, nullptr, L[D2].cip, dC2, nullptr, C2, false, nimg);
    ew("warm_up_bwd", px(3) * 64, 10.0, k_up_bwd, dU2, dB3, nimg);
    weight_grad(n, E3B, A3a, dB3, nimg);
    conv(dB3, L[E3B].wt.p, nullptr, L[E3B].cip, dA3a, nullptr, A3a, false, nimg);
    weight_grad(n, E3A, P2, dA3a, nimg);
    conv(dA3a, L[E3A].wt.p, nullptr, L[E3A].cip, dP2, nullptr, none, false, nimg);
    launch("warm_pool_bwd", k_pool_bwd, dim3(cdiv(px(2) * 64, 256)), dim3(256), 0, dP2, dS2, S2, nimg);
    weight_grad(n, E2B, A2a, dS2, nimg);
    conv(dS2, L[E2B].wt.p, nullptr, L[E2B].cip, dA2a, nullptr, A2a, false, nimg);
    weight_grad(n, E2A, P1, dA2a, nimg);
    conv(dA2a, L[E2A].wt.p, nullptr, L[E2A].cip, dP1, nullptr, none, false, nimg);
    launch("warm_pool_bwd", k_pool_bwd, dim3(cdiv(px(1) * 32, 256)), dim3(256), 0, dP1, dS1, S1, nimg);
    weight_grad(n, E1B, A1a, dS1, nimg);
    conv(dS1, L[E1B].wt.p, nullptr, L[E1B].cip, dA1a, nullptr, A1a, false, nimg);
    weight_grad(n, E1A, P0, dA1a, nimg);
    conv(dA1a, L[E1A].wt.p, nullptr, L[E1A].cip, dP0, nullptr, none, false, nimg);
    launch("warm_pool_bwd", k_pool_bwd, dim3(cdiv(px(0) * 16, 256)), dim3(256), 0, dP0, dS0, S0, nimg);
    weight_grad(n, E0B, A0a, dS0, nimg);
    conv(dS0, L[E0B].wt.p, nullptr, L[E0B].cip, dA0a, nullptr, A0a, false, nimg);
    weight_grad(n, E0A, X0, dA0a, nimg);
    std::vector<double> lp(lb);
    copy_mem(lp.data(), LSH_HOST, n.lpart.p, LSH_DEVICE, lb * sizeof(double));
    double s = 0.0;
    for (double v : lp) s += v;
    return s / (double)n0;
}

struct Inputs {
    const float *d, *conf, *g, *t;
};

// device pointers of the caller's planes (staged when in host memory)
static Inputs stage(lsh_warm &n, const float *d, const float *conf, const float *g, const float *t, lsh_mem mem,
                    size_t count) {
    if (mem == LSH_DEVICE) return Inputs{d, conf, g, t};
    n.dbuf.ensure(count);
    n.cbuf.ensure(count);
    n.gbuf.ensure(count);
    copy_mem(n.dbuf.p, LSH_DEVICE, d, LSH_HOST, count * sizeof(float));
    copy_mem(n.cbuf.p, LSH_DEVICE, conf, LSH_HOST, count * sizeof(float));
    copy_mem(n.gbuf.p, LSH_DEVICE, g, LSH_HOST, count * sizeof(float));
    if (t) {
        n.tbuf.ensure(count);
        copy_mem(n.tbuf.p, LSH_DEVICE, t, LSH_HOST, count * sizeof(float));
    }
    return Inputs{n.dbuf.p, n.cbuf.p, n.gbuf.p, t ? n.tbuf.p : nullptr};
}

// the start for one system of the depth pass, on the library's stream (no
// synchronisation): dscale times the prolongated predicted block means
void warm_predict_dev(lsh_warm *n, const float *d, const float *conf, const float *g, int w, int h, float dscale,
                      float sigma_c, float *u) {
    const int ch = cdim(h), cw = cdim(w);
    ensure_acts(*n, 1, ch, cw, false);
    features(*n, d, conf, g, nullptr, 1, h, w, dscale, sigma_c);
    forward(*n, 1);
    const size_t cnt = (size_t)w * h;
    launch_b("warm_prolong", 4.0 * cnt, k_prolong, dim3(cdiv(cnt, 256)), dim3(256), 0, (const double *)n->cf.p,
             (const float *)n->A.out32.p, 1, ch, cw, h, w, (double)ch / (double)h, (double)cw / (double)w,
             (double)dscale, u);
}

}  // namespace lsh

using namespace lsh;

extern "C" {

char *lsh_warm_new(uint64_t seed, lsh_warm **out) {
    LSH_API_BEGIN
    LSH_REQUIRE(out, "lsh_warm_new: null argument");
    std::unique_ptr<lsh_warm> n(new lsh_warm());
    size_t off = 0, nat = 0;
    for (int i = 0; i < NLAYER; i++) {
        Layer l;
        l.cin = SPECS[i].cin;
        l.cout = SPECS[i].cout;
        l.cip = pad16(l.cin);
        l.cop = pad16(l.cout);
        l.relu = SPECS[i].relu;
        l.w_off = off;
        off += (size_t)9 * l.cop * l.cip;
        l.b_off = off;
        off += l.cop;
        l.n_nat = (size_t)9 * l.cin * l.cout + l.cout;
        nat += l.n_nat;
        l.wf.alloc((size_t)9 * l.cop * l.cip);
        l.wt.alloc((size_t)9 * l.cop * l.cip);
        n->L.push_back(std::move(l));
    }
    n->npad = off;
    n->nnat = nat;
    for (auto *b : {&n->w32, &n->g32, &n->m, &n->v}) {
        b->alloc(off);
        b->zero();
    }
    // He-normal weights (std sqrt(2 / (9 cin))), zero biases; the last layer
    // starts at zero so the first prediction is the blocks' weighted mean depth
    std::vector<float> natw(nat, 0.0f);
    uint64_t s = seed;
    size_t o = 0;
    for (auto &l : n->L) {
        double sd = (&l == &n->L.back()) ? 0.0 : std::sqrt(2.0 / (9.0 * l.cin));
        for (size_t i = 0; i < (size_t)9 * l.cin * l.cout; i++) {
            double u1 = splitmix(s), u2 = splitmix(s);
            natw[o++] = (float)(sd * std::sqrt(-2.0 * std::log(u1 + 1e-300)) * std::cos(6.283185307179586 * u2));
        }
        o += l.cout;
    }
    std::vector<float> pad;
    nat_to_pad(*n, natw.data(), pad);
    copy_mem(n->w32.p, LSH_DEVICE, pad.data(), LSH_HOST, pad.size() * sizeof(float));
    recast(*n);
    sync();
    *out = n.release();
    LSH_API_END
}

void lsh_warm_free(lsh_warm *n) { delete n; }

char *lsh_warm_nparams(lsh_warm *n, uint32_t *count) {
    LSH_API_BEGIN
    LSH_REQUIRE(n && count, "lsh_warm_nparams: null argument");
    *count = (uint32_t)n->nnat;
    LSH_API_END
}

char *lsh_warm_get(lsh_warm *n, float *params, uint32_t count) {
    LSH_API_BEGIN
    LSH_REQUIRE(n && params, "lsh_warm_get: null argument");
    LSH_REQUIRE(count == n->nnat, "warm: parameter count mismatch");
    std::vector<float> pad(n->npad);
    copy_mem(pad.data(), LSH_HOST, n->w32.p, LSH_DEVICE, n->npad * sizeof(float));
    pad_to_nat(*n, pad, params);
    LSH_API_END
}

char *lsh_warm_set(lsh_warm *n, const float *params, uint32_t count) {
    LSH_API_BEGIN
    LSH_REQUIRE(n && params, "lsh_warm_set: null argument");
    LSH_REQUIRE(count == n->nnat, "warm: parameter count mismatch");
    std::vector<float> pad;
    nat_to_pad(*n, params, pad);
    copy_mem(n->w32.p, LSH_DEVICE, pad.data(), LSH_HOST, n->npad * sizeof(float));
    n->m.zero();
    n->v.zero();
    n->step = 0;
    recast(*n);
    sync();
    LSH_API_END
}

char *lsh_warm_set_sigma(lsh_warm *n, float sigma_c) {
    LSH_API_BEGIN
    LSH_REQUIRE(n, "lsh_warm_set_sigma: null handle");
    LSH_REQUIRE(sigma_c > 0.0f && std::isfinite(sigma_c), "warm: sigma_c must be positive");
    n->sigma_c = sigma_c;
    LSH_API_END
}

char *lsh_warm_predict(lsh_warm *n, const float *d, const float *conf, const float *guide, lsh_mem in_mem,
                       uint32_t nimg, uint32_t w, uint32_t h, float dscale, float *u, lsh_mem out_mem) {
    LSH_API_BEGIN
    LSH_REQUIRE(n && d && conf && guide && u, "lsh_warm_predict: null argument");
    LSH_REQUIRE(nimg > 0 && w > 0 && h > 0, "warm: empty batch");
    LSH_REQUIRE(dscale > 0.0f, "warm: dscale must be positive");
    const size_t cnt = (size_t)nimg * w * h;
    Inputs in = stage(*n, d, conf, guide, nullptr, in_mem, cnt);
    const int ch = cdim((int)h), cw = cdim((int)w);
    ensure_acts(*n, (int)nimg, ch, cw, false);
    features(*n, in.d, in.conf, in.g, nullptr, (int)nimg, (int)h, (int)w, dscale, n->sigma_c);
    forward(*n, (int)nimg);
    float *uo = u;
    if (out_mem == LSH_HOST) {
        n->ubuf.ensure(cnt);
        uo = n->ubuf.p;
    }
    launch_b("warm_prolong", 4.0 * cnt, k_prolong, dim3(cdiv(cnt, 256)), dim3(256), 0, (const double *)n->cf.p,
             (const float *)n->A.out32.p, (int)nimg, ch, cw, (int)h, (int)w, (double)ch / (double)h,
             (double)cw / (double)w, (double)dscale, uo);
    if (out_mem == LSH_HOST) copy_mem(u, LSH_HOST, uo, LSH_DEVICE, cnt * sizeof(float));
    sync();
    LSH_API_END
}

char *lsh_warm_grads(lsh_warm *n, const float *d, const float *conf, const float *guide, const float *target,
                     lsh_mem in_mem, uint32_t nimg, uint32_t w, uint32_t h, float dscale, double *loss, float *grads,
                     uint32_t count) {
    LSH_API_BEGIN
    LSH_REQUIRE(n && d && conf && guide && target, "lsh_warm_grads: null argument");
    LSH_REQUIRE(nimg > 0 && w > 0 && h > 0, "warm: empty batch");
    LSH_REQUIRE(dscale > 0.0f, "warm: dscale must be positive");
    LSH_REQUIRE(!grads || count == n->nnat, "warm: parameter count mismatch");
    const size_t cnt = (size_t)nimg * w * h;
    Inputs in = stage(*n, d, conf, guide, target, in_mem, cnt);
    ensure_acts(*n, (int)nimg, cdim((int)h), cdim((int)w), true);
    features(*n, in.d, in.conf, in.g, in.t, (int)nimg, (int)h, (int)w, dscale, n->sigma_c);
    forward(*n, (int)nimg);
    double l = backward(*n, (int)nimg);
    if (loss) *loss = l;
    if (grads) {
        std::vector<float> pad(n->npad);
        copy_mem(pad.data(), LSH_HOST, n->g32.p, LSH_DEVICE, n->npad * sizeof(float));
        pad_to_nat(*n, pad, grads);
    }
    LSH_API_END
}

char *lsh_warm_train_step(lsh_warm *n, const float *d, const float *conf, const float *guide, const float *target,
                          lsh_mem in_mem, uint32_t nimg, uint32_t w, uint32_t h, float dscale, float lr, double *loss) {
    LSH_API_BEGIN
    LSH_REQUIRE(n && d && conf && guide && target, "lsh_warm_train_step: null argument");
    LSH_REQUIRE(nimg > 0 && w > 0 && h > 0, "warm: empty batch");
    LSH_REQUIRE(dscale > 0.0f, "warm: dscale must be positive");
    const size_t cnt = (size_t)nimg * w * h;
    Inputs in = stage(*n, d, conf, guide, target, in_mem, cnt);
    ensure_acts(*n, (int)nimg, cdim((int)h), cdim((int)w), true);
    features(*n, in.d, in.conf, in.g, in.t, (int)nimg, (int)h, (int)w, dscale, n->sigma_c);
    forward(*n, (int)nimg);
    double l = backward(*n, (int)nimg);
    if (loss) *loss = l;
    if (!std::isfinite(l)) return nullptr;  // a non-finite loss leaves the weights and the optimiser as they were
    n->step++;
    const float b1 = 0.9f, b2 = 0.999f;
    double c1 = 1.0 - std::pow((double)b1, (double)n->step), c2 = 1.0 - std::pow((double)b2, (double)n->step);
    float lr_t = (float)(lr * std::sqrt(c2) / c1), eps_t = (float)(1e-8 * std::sqrt(c2));
    launch("warm_adam", k_adam, dim3(cdiv(n->npad, 256)), dim3(256), 0, n->w32.p, (const float *)n->g32.p, n->m.p,
           n->v.p, n->npad, lr_t, b1, b2, eps_t);
    recast(*n);
    LSH_API_END
}

}  // extern "C"
