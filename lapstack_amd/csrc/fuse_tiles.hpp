// fuse_tiles.hpp -- the per-tile bodies of the pyramid REDUCE and the
// region-energy fold, shared by the per-frame kernels (fuse.hip) and the
// batched persistent kernel (batch.hip), so both compute every sample with
// the same operations in the same order.
#pragma once
#include "common.hpp"

namespace lsh {

// pyramid.rs:46, :124-125
__constant__ float KR[5] = {1.0f / 16.0f, 4.0f / 16.0f, 6.0f / 16.0f, 4.0f / 16.0f, 1.0f / 16.0f};
#define EV0 (2.0f * (1.0f / 16.0f))
#define EV1 (2.0f * (6.0f / 16.0f))
#define EV2 (2.0f * (1.0f / 16.0f))
#define ODDK (2.0f * (4.0f / 16.0f))

#define RED_TX 64
#define RED_TY 4
struct ReduceLds {
    float T[2 * RED_TY + 3][2 * RED_TX + 3 + 1];
    float Hs[2 * RED_TY + 3][RED_TX + 1];
};

// Where a tile body reads the samples of a level: float planes (c * n + o), or
// the frame as decoded, interleaved RGB u16 scaled by 1/65535 (io.rs:89), which
// is the float level 0 the ingest would have written, value for value.
struct PlaneSrc {
    const float *__restrict__ G;
    size_t n;
    __device__ __forceinline__ float operator()(size_t o, int c) const { return G[c * n + o]; }
};
struct RgbSrc {
    const uint16_t *__restrict__ p;
    __device__ __forceinline__ float operator()(size_t o, int c) const {
        return (float)p[3 * o + c] * (1.0f / 65535.0f);
    }
};

// A sample index along an axis of n: reflect-101 (pyramid.rs:26-38) on tiles
// that reach past a border, the index itself on interior tiles, where every
// index a tile touches is in range and the reflection is the identity.
template <bool EDGE>
__device__ __forceinline__ int refl(int i, int n) {
    return EDGE ? rf101(i, n) : i;
}

template <bool EDGE, class Src>
__device__ __forceinline__ void reduce_body(const Src &s, int c, int w, int h, float *__restrict__ d, int ow, int oh,
                                            int ox0, int oy0, ReduceLds &S) {
    constexpr int TX = RED_TX, TY = RED_TY, IW = 2 * TX + 3, IH = 2 * TY + 3;
    const int tid = threadIdx.x;
    const int ix0 = 2 * ox0 - 2, iy0 = 2 * oy0 - 2;
    const int nx = EDGE ? min(IW, 2 * (ow - ox0) + 3) : IW, ny = EDGE ? min(IH, 2 * (oh - oy0) + 3) : IH;
    for (int k = tid; k < IH * IW; k += TX * TY) {
        int r = k / IW, x = k - r * IW;
        if (!EDGE || (r < ny && x < nx)) S.T[r][x] = s((size_t)refl<EDGE>(iy0 + r, h) * w + refl<EDGE>(ix0 + x, w), c);
    }
    __syncthreads();
    for (int k = tid; k < IH * TX; k += TX * TY) {
        int r = k / TX, jj = k - r * TX;
        if (!EDGE || (r < ny && ox0 + jj < ow)) {
            float a = 0.0f;
#pragma unroll
            for (int u = 0; u < 5; u++) a += KR[u] * S.T[r][2 * jj + u];
            S.Hs[r][jj] = a;
        }
    }
    __syncthreads();
    const int tx = tid & (TX - 1), ty = tid / TX;
    const int j = ox0 + tx, i = oy0 + ty;
    if (!EDGE || (j < ow && i < oh)) {
        const int r0 = 2 * ty;
        d[(size_t)i * ow + j] = KR[0] * S.Hs[r0][tx] + KR[1] * S.Hs[r0 + 1][tx] + KR[2] * S.Hs[r0 + 2][tx] +
                                KR[3] * S.Hs[r0 + 3][tx] + KR[4] * S.Hs[r0 + 4][tx];
    }
}

// pyramid.rs:86-119: REDUCE of plane c on the 64x4 output tile at (ox0, oy0):
// the 131x11 input window (reflect-101 at the borders) is staged in LDS, the
// horizontal 5-tap sums of its 11 rows at the 64 even columns go to LDS, and
// each of the 256 threads sums its five rows.  Calls may follow each other on
// one ReduceLds with no barrier between them.
template <class Src>
__device__ __forceinline__ void reduce_tile(const Src &s, int c, int w, int h, float *__restrict__ d, int ow, int oh,
                                            int ox0, int oy0, ReduceLds &S) {
    const bool inner = ox0 >= 1 && oy0 >= 1 && 2 * ox0 + 2 * RED_TX <= w - 1 && 2 * oy0 + 2 * RED_TY <= h - 1 &&
                       ox0 + RED_TX <= ow && oy0 + RED_TY <= oh;
    if (inner)
        reduce_body<false>(s, c, w, h, d, ow, oh, ox0, oy0, S);
    else
        reduce_body<true>(s, c, w, h, d, ow, oh, ox0, oy0, S);
}

// pyramid.rs:128-140: the horizontal EXPAND of coarse row cr at fine column x
__device__ __forceinline__ float expand_h(const float *__restrict__ C, int cw, int cr, int x) {
    const float *r = C + (size_t)cr * cw;
    int c = x >> 1;
    if ((x & 1) == 0) return EV0 * r[rf101(c - 1, cw)] + EV1 * r[rf101(c, cw)] + EV2 * r[rf101(c + 1, cw)];
    return ODDK * (r[rf101(c, cw)] + r[rf101(c + 1, cw)]);
}

// pyramid.rs:143-168: EXPAND of the coarse plane at fine pixel (y, x)
__device__ __forceinline__ float expand_at(const float *__restrict__ C, int cw, int ch, int y, int x) {
    int c = y >> 1;
    if ((y & 1) == 0)
        return EV0 * expand_h(C, cw, rf101(c - 1, ch), x) + EV1 * expand_h(C, cw, rf101(c, ch), x) +
               EV2 * expand_h(C, cw, rf101(c + 1, ch), x);
    return ODDK * (expand_h(C, cw, rf101(c, ch), x) + expand_h(C, cw, rf101(c + 1, ch), x));
}

// fuse.rs:152-165
__device__ __forceinline__ float node_energy(float r, float g, float b, int chroma) {
    if (chroma) return r * r + g * g + b * b;
    float y = 0.299f * r + 0.587f * g + 0.114f * b;
    return y * y;
}

struct Wt9 {
    float k[9];
};

template <int R, int TY>
struct FoldLds {
    static constexpr int TX = 64, EW = TX + 2 * R, EH = TY + 2 * R;
    static constexpr int CWM = TX / 2 + R + 4, CHM = TY / 2 + R + 4;
    float Cs[3][CHM][CWM];
    float Hs[3][CHM][EW];
    float Es[EH][EW];
    float Ts[EH][TX];
    float Ls[3][TY][TX];
};

constexpr int floordiv2c(int a) { return a >= 0 ? a / 2 : -((1 - a) / 2); }

template <int R, int TY, bool EDGE, class Src>
__device__ __forceinline__ void fold_body(const Src &G, const float *__restrict__ Cn, int w, int h, int cw, int ch,
                                          const Wt9 &wt, int chroma, int bx, int by, FoldLds<R, TY> &S) {
    constexpr int TX = 64, EW = TX + 2 * R, EH = TY + 2 * R;
    // bx and by are even (multiples of 64 and of TY), so the coarse window's
    // size depends on R and TY only
    constexpr int CW = floordiv2c(TX + R - 1) - floordiv2c(-R) + 3, CH = floordiv2c(TY + R - 1) - floordiv2c(-R) + 3;
    static_assert(CW <= FoldLds<R, TY>::CWM && CH <= FoldLds<R, TY>::CHM, "coarse window exceeds its LDS tile");
    const int tid = threadIdx.x;
    const size_t cn = (size_t)cw * ch;
    const int cx0 = (bx >> 1) + floordiv2c(-R) - 1, cy0 = (by >> 1) + floordiv2c(-R) - 1;
    // the last fine column / row any output of this tile reads (window reach)
    const int xlim = min(bx + TX, w) - 1 + R, ylim = min(by + TY, h) - 1 + R;
    for (int k = tid; k < 3 * CH * CW; k += 256) {
        int c = k / (CH * CW), r = k - c * CH * CW, j = r / CW, i = r - j * CW;
        S.Cs[c][j][i] = Cn[c * cn + (size_t)refl<EDGE>(cy0 + j, ch) * cw + refl<EDGE>(cx0 + i, cw)];
    }
    __syncthreads();
    // pyramid.rs:128-140: horizontal EXPAND of each staged coarse row at each tile column
    for (int k = tid; k < 3 * CH * EW; k += 256) {
        int c = k / (CH * EW), r = k - c * CH * EW, j = r / EW, lx = r - j * EW;
        if (EDGE && bx + lx - R > xlim) continue;
        int gx = refl<EDGE>(bx + lx - R, w), i0 = (gx >> 1) - cx0;
        const float *row = S.Cs[c][j];
        S.Hs[c][j][lx] = (gx & 1) == 0 ? EV0 * row[i0 - 1] + EV1 * row[i0] + EV2 * row[i0 + 1]
                                       : ODDK * (row[i0] + row[i0 + 1]);
    }
    __syncthreads();
    // pyramid.rs:151-166 vertical EXPAND, L = G - EXPAND (pyramid.rs:208-211), node energy
    for (int k = tid; k < EH * EW; k += 256) {
        int ly = k / EW, lx = k - ly * EW;
        if (EDGE && (by + ly - R > ylim || bx + lx - R > xlim)) continue;
        int gy = refl<EDGE>(by + ly - R, h), gx = refl<EDGE>(bx + lx - R, w), j0 = (gy >> 1) - cy0;
        size_t o = (size_t)gy * w + gx;
        float l[3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            float e = (gy & 1) == 0 ? EV0 * S.Hs[c][j0 - 1][lx] + EV1 * S.Hs[c][j0][lx] + EV2 * S.Hs[c][j0 + 1][lx]
                                    : ODDK * (S.Hs[c][j0][lx] + S.Hs[c][j0 + 1][lx]);
            l[c] = G(o, c) - e;
        }
        S.Es[ly][lx] = node_energy(l[0], l[1], l[2], chroma);
        int cy = ly - R, cx = lx - R;
        if (cy >= 0 && cy < TY && cx >= 0 && cx < TX) {
            S.Ls[0][cy][cx] = l[0];
            S.Ls[1][cy][cx] = l[1];
            S.Ls[2][cy][cx] = l[2];
        }
    }
    __syncthreads();
    // fuse.rs:174-183: horizontal window on every tile row
    for (int k = tid; k < EH * TX; k += 256) {
        int ly = k / TX, lx = k - ly * TX;
        if (EDGE && (by + ly - R > ylim || bx + lx >= w)) continue;
        float a = 0.0f;
#pragma unroll
        for (int t = 0; t <= 2 * R; t++) a += wt.k[t] * S.Es[ly][lx + t];
        S.Ts[ly][lx] = a;
    }
    __syncthreads();
}

// One band-pass level of the fold (fuse.rs:401-451) on the 64 x TY tile at
// (bx, by) of the level (both even), up to the energies: L = G - EXPAND(G_next)
// and the node energy on the tile and its window margin, then the horizontal
// binomial window on every tile row.  The coarse window under the tile and its
// horizontal EXPAND at every tile column are staged in LDS first, so each fine
// sample costs three global loads.  LDS rows and columns are indexed by
// unreflected coordinates and hold the reflect-101 sample of each; every index
// a fine sample the window can reach needs lies inside the staged window.
// Leaves the tile's L samples in S.Ls and the horizontal sums in S.Ts, with a
// barrier after.  256 threads.
template <int R, int TY, class Src>
__device__ __forceinline__ void fold_stage(const Src &G, const float *__restrict__ Cn, int w, int h, int cw, int ch,
                                           const Wt9 &wt, int chroma, int bx, int by, FoldLds<R, TY> &S) {
    constexpr int TX = 64, CW = floordiv2c(TX + R - 1) - floordiv2c(-R) + 3,
                  CH = floordiv2c(TY + R - 1) - floordiv2c(-R) + 3;
    const int cx0 = (bx >> 1) + floordiv2c(-R) - 1, cy0 = (by >> 1) + floordiv2c(-R) - 1;
    const bool inner = bx - R >= 0 && by - R >= 0 && bx + TX - 1 + R <= w - 1 && by + TY - 1 + R <= h - 1 && cx0 >= 0 &&
                       cy0 >= 0 && cx0 + CW - 1 <= cw - 1 && cy0 + CH - 1 <= ch - 1;
    if (inner)
        fold_body<R, TY, false>(G, Cn, w, h, cw, ch, wt, chroma, bx, by, S);
    else
        fold_body<R, TY, true>(G, Cn, w, h, cw, ch, wt, chroma, bx, by, S);
}

// fuse.rs:186-195: the vertical window at tile row ty, column tx (after fold_stage)
template <int R, int TY>
__device__ __forceinline__ float fold_energy(const Wt9 &wt, int tx, int ty, const FoldLds<R, TY> &S) {
    float en = 0.0f;
#pragma unroll
    for (int t = 0; t <= 2 * R; t++) en += wt.k[t] * S.Ts[ty + t][tx];
    return en;
}

// One band-pass level of the fold of one frame on the 64 x TY tile at (bx,
// by): the stages above, then the strict-'>' select against the best energies
// so far, with nothing but the accumulator, the best energies and the winner
// map written.  Ends with a barrier (the LDS tile is free for the next one).
template <int R, int TY, class Src>
__device__ __forceinline__ void fold_tile(const Src &G, const float *__restrict__ Cn, int w, int h, int cw, int ch,
                                          float *__restrict__ best, float *__restrict__ acc,
                                          uint16_t *__restrict__ winner, int first, unsigned widx, const Wt9 &wt,
                                          int chroma, float *__restrict__ en_out, int bx, int by, FoldLds<R, TY> &S) {
    fold_stage<R, TY>(G, Cn, w, h, cw, ch, wt, chroma, bx, by, S);
    const size_t n = (size_t)w * h;
    const int tid = threadIdx.x, tx = tid & 63, x = bx + tx;
    if (x < w) {
#pragma unroll
        for (int q = 0; q < TY / 4; q++) {
            int ty = (tid >> 6) + 4 * q, y = by + ty;
            if (y >= h) break;
            const float en = fold_energy<R, TY>(wt, tx, ty, S);
            size_t o = (size_t)y * w + x;
            if (en_out) en_out[o] = en;
            if (first) {
                best[o] = en;
                acc[o] = S.Ls[0][ty][tx];
                acc[n + o] = S.Ls[1][ty][tx];
                acc[2 * n + o] = S.Ls[2][ty][tx];
            } else if (en > best[o]) {  // strict: ties keep the earlier frame (fuse.rs:429, 442)
                best[o] = en;
                acc[o] = S.Ls[0][ty][tx];
                acc[n + o] = S.Ls[1][ty][tx];
                acc[2 * n + o] = S.Ls[2][ty][tx];
                if (winner) winner[o] = (uint16_t)widx;
            }
        }
    }
    __syncthreads();
}

}  // namespace lsh
