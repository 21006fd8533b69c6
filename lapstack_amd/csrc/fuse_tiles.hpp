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
template <int TY>
struct ReduceLdsT {
    float T[2 * TY + 3][2 * RED_TX + 3 + 1];
    float Hs[2 * TY + 3][RED_TX + 1];
};
using ReduceLds = ReduceLdsT<RED_TY>;

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

template <int TY, bool EDGE, bool PRE, class Src>
__device__ __forceinline__ void reduce_body(const Src &s, int c, int w, int h, float *__restrict__ d, int ow, int oh,
                                            int ox0, int oy0, ReduceLdsT<TY> &S) {
    constexpr int TX = RED_TX, IW = 2 * TX + 3, IH = 2 * TY + 3;
    const int tid = threadIdx.x;
    const int ix0 = 2 * ox0 - 2, iy0 = 2 * oy0 - 2;
    const int nx = EDGE ? min(IW, 2 * (ow - ox0) + 3) : IW, ny = EDGE ? min(IH, 2 * (oh - oy0) + 3) : IH;
    auto ld = [&](int k, float &v) {
        const int r = k / IW, x = k - r * IW;
        if (EDGE && !(r < ny && x < nx)) return false;
        v = s((size_t)refl<EDGE>(iy0 + r, h) * w + refl<EDGE>(ix0 + x, w), c);
        return true;
    };
    auto st = [&](int k, float v) {
        const int r = k / IW;
        S.T[r][k - r * IW] = v;
    };
    if constexpr (PRE) {
        lds_stage<IH * IW, 256, float>(ld, st);
    } else {
        for (int k = tid; k < IH * IW; k += 256) {
            float v;
            if (ld(k, v)) st(k, v);
        }
    }
    __syncthreads();
    for (int k = tid; k < IH * TX; k += 256) {
        int r = k / TX, jj = k - r * TX;
        if (!EDGE || (r < ny && ox0 + jj < ow)) {
            float a = 0.0f;
#pragma unroll
            for (int u = 0; u < 5; u++) a += KR[u] * S.T[r][2 * jj + u];
            S.Hs[r][jj] = a;
        }
    }
    __syncthreads();
    const int tx = tid & (TX - 1), j = ox0 + tx;
    if (EDGE && j >= ow) return;
#pragma unroll
    for (int q = 0; q < TY / 4; q++) {
        const int ty = tid / TX + 4 * q, i = oy0 + ty, r0 = 2 * ty;
        if (EDGE && i >= oh) break;
        d[(size_t)i * ow + j] = KR[0] * S.Hs[r0][tx] + KR[1] * S.Hs[r0 + 1][tx] + KR[2] * S.Hs[r0 + 2][tx] +
                                KR[3] * S.Hs[r0 + 3][tx] + KR[4] * S.Hs[r0 + 4][tx];
    }
}

// pyramid.rs:86-119: REDUCE of plane c on the 64 x TY output tile at (ox0,
// oy0): the (2 TY + 3) x 131 input window (reflect-101 at the borders) is
// staged in LDS, the horizontal 5-tap sums of its rows at the 64 even columns
// go to LDS, and each of the 256 threads sums five rows for each of its TY / 4
// outputs.  Calls may follow each other on one LDS tile with no barrier
// between them.  PRE: the window's loads all go out before the first is
// stored (one wait on memory per tile), else in a strided loop (fewer
// registers); the same values either way.
template <bool PRE, int TY, class Src>
__device__ __forceinline__ void reduce_tile(const Src &s, int c, int w, int h, float *__restrict__ d, int ow, int oh,
                                            int ox0, int oy0, ReduceLdsT<TY> &S) {
    const bool inner = ox0 >= 1 && oy0 >= 1 && 2 * ox0 + 2 * RED_TX <= w - 1 && 2 * oy0 + 2 * TY <= h - 1 &&
                       ox0 + RED_TX <= ow && oy0 + TY <= oh;
    if (inner)
        reduce_body<TY, false, PRE>(s, c, w, h, d, ow, oh, ox0, oy0, S);
    else
        reduce_body<TY, true, PRE>(s, c, w, h, d, ow, oh, ox0, oy0, S);
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

template <int R, int TY, bool EDGE, bool PRE, class Src>
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
    constexpr int NC = 3 * CH * CW, ITC = (NC + 255) / 256, NL = EH * EW, ITL = PRE ? (NL + 255) / 256 : 1;
    auto coarse_at = [&](int k) {
        const int c = k / (CH * CW), r = k - c * CH * CW, j = r / CW, i = r - j * CW;
        return Cn[c * cn + (size_t)refl<EDGE>(cy0 + j, ch) * cw + refl<EDGE>(cx0 + i, cw)];
    };
    auto coarse_put = [&](int k, float v) {
        const int c = k / (CH * CW), r = k - c * CH * CW, j = r / CW;
        S.Cs[c][j][r - j * CW] = v;
    };
    auto skip = [&](int ly, int lx) { return EDGE && (by + ly - R > ylim || bx + lx - R > xlim); };
    auto fine_at = [&](int ly, int lx) { return (size_t)refl<EDGE>(by + ly - R, h) * w + refl<EDGE>(bx + lx - R, w); };
    float gv[ITL][3];
    if constexpr (PRE) {
        // Every global load of the tile is issued before the first is used: the
        // coarse window, then the fine samples of the L stage below, held in
        // registers, so the tile waits on memory once.
        float cv[ITC];
#pragma unroll
        for (int it = 0; it < ITC; it++) {
            const int k = tid + 256 * it;
            if (k < NC) cv[it] = coarse_at(k);
        }
#pragma unroll
        for (int it = 0; it < ITL; it++) {
            const int k = tid + 256 * it, ly = k / EW, lx = k - ly * EW;
            if (k < NL && !skip(ly, lx)) {
                const size_t o = fine_at(ly, lx);
#pragma unroll
                for (int c = 0; c < 3; c++) gv[it][c] = G(o, c);
            }
        }
#pragma unroll
        for (int it = 0; it < ITC; it++) {
            const int k = tid + 256 * it;
            if (k < NC) coarse_put(k, cv[it]);
        }
    } else {
        for (int k = tid; k < NC; k += 256) coarse_put(k, coarse_at(k));
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
    auto lstage = [&](int ly, int lx, const float *g) {
        int gy = refl<EDGE>(by + ly - R, h), j0 = (gy >> 1) - cy0;
        float l[3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            float e = (gy & 1) == 0 ? EV0 * S.Hs[c][j0 - 1][lx] + EV1 * S.Hs[c][j0][lx] + EV2 * S.Hs[c][j0 + 1][lx]
                                    : ODDK * (S.Hs[c][j0][lx] + S.Hs[c][j0 + 1][lx]);
            l[c] = g[c] - e;
        }
        S.Es[ly][lx] = node_energy(l[0], l[1], l[2], chroma);
        int cy = ly - R, cx = lx - R;
        if (cy >= 0 && cy < TY && cx >= 0 && cx < TX) {
            S.Ls[0][cy][cx] = l[0];
            S.Ls[1][cy][cx] = l[1];
            S.Ls[2][cy][cx] = l[2];
        }
    };
    if constexpr (PRE) {
#pragma unroll
        for (int it = 0; it < ITL; it++) {
            const int k = tid + 256 * it, ly = k / EW, lx = k - ly * EW;
            if (k < NL && !skip(ly, lx)) lstage(ly, lx, gv[it]);
        }
    } else {
        for (int k = tid; k < NL; k += 256) {
            const int ly = k / EW, lx = k - ly * EW;
            if (skip(ly, lx)) continue;
            const size_t o = fine_at(ly, lx);
            const float g[3] = {G(o, 0), G(o, 1), G(o, 2)};
            lstage(ly, lx, g);
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
// barrier after.  256 threads.  PRE: every global load of the tile is issued
// up front and held in registers (one wait on memory per tile); otherwise the
// loads go out in strided loops, which keeps the register count low where
// residency matters more than latency (the persistent batched kernel).  Both
// compute the same values.
template <int R, int TY, bool PRE, class Src>
__device__ __forceinline__ void fold_stage(const Src &G, const float *__restrict__ Cn, int w, int h, int cw, int ch,
                                           const Wt9 &wt, int chroma, int bx, int by, FoldLds<R, TY> &S) {
    constexpr int TX = 64, CW = floordiv2c(TX + R - 1) - floordiv2c(-R) + 3,
                  CH = floordiv2c(TY + R - 1) - floordiv2c(-R) + 3;
    const int cx0 = (bx >> 1) + floordiv2c(-R) - 1, cy0 = (by >> 1) + floordiv2c(-R) - 1;
    const bool inner = bx - R >= 0 && by - R >= 0 && bx + TX - 1 + R <= w - 1 && by + TY - 1 + R <= h - 1 && cx0 >= 0 &&
                       cy0 >= 0 && cx0 + CW - 1 <= cw - 1 && cy0 + CH - 1 <= ch - 1;
    if (inner)
        fold_body<R, TY, false, PRE>(G, Cn, w, h, cw, ch, wt, chroma, bx, by, S);
    else
        fold_body<R, TY, true, PRE>(G, Cn, w, h, cw, ch, wt, chroma, bx, by, S);
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
    const size_t n = (size_t)w * h;
    const int tid = threadIdx.x, tx = tid & 63, x = bx + tx;
    // the best energies so far at the thread's outputs, loaded ahead of the
    // tile's stages (nothing in them writes best)
    float prev[TY / 4];
    if (!first && x < w) {
#pragma unroll
        for (int q = 0; q < TY / 4; q++) {
            const int y = by + (tid >> 6) + 4 * q;
            if (y < h) prev[q] = best[(size_t)y * w + x];
        }
    }
    fold_stage<R, TY, true>(G, Cn, w, h, cw, ch, wt, chroma, bx, by, S);
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
            } else if (en > prev[q]) {  // strict: ties keep the earlier frame (fuse.rs:429, 442)
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
