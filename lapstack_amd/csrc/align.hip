// align.hip -- registration and resampling on gfx950 (align.rs, brightness.rs,
// the per-frame part of stack.rs), behind the frames part of the C ABI.
//
// A frame arrives as interleaved RGB u16 in HBM and is never expanded to f32
// before the warp: every kernel reads the 16-bit taps and scales each by
// 1/65535 exactly as the decoder does (io.rs:89), so the values match the
// CPU path's planes bit for bit.
//
// The search is the reference's bounded Nelder-Mead, run on the device by one
// cooperative kernel per frame: each cost evaluation warps the target level
// onto the reference level (f64 coordinates, f32 Spline4x4 taps), folds d, d^2
// and the count per thread in f64, reduces them across each block with
// wavefront shuffles and across blocks, after one grid barrier, in a fixed
// order -- no host round trip per evaluation.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <vector>

#include "api.hpp"
#include "common.hpp"
#include "engine.hpp"

namespace lsh {


#define INV16 (1.0f / 65535.0f)

// align.rs:562-571 on the decoded frame, four pixels per thread: their 24
// bytes of interleaved u16 as three 8-byte loads and their lumas as one 16-byte
// store (a frame whose address is not 8-byte aligned, and the last n % 4
// pixels, one pixel at a time)
__device__ __forceinline__ float luma_of(uint32_t r, uint32_t g, uint32_t b) {
    return 0.299f * ((float)r * INV16) + 0.587f * ((float)g * INV16) + 0.114f * ((float)b * INV16);
}
__global__ void k_luma_u16(const FrameRef src, size_t n, float *__restrict__ y) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x, i = 4 * t;
    if (i >= n) return;
    const uint16_t *__restrict__ rgb = src.get();
    if (i + 4 <= n && ((uintptr_t)rgb & 7) == 0 && ((uintptr_t)y & 15) == 0) {
        const uint64_t *q = (const uint64_t *)(rgb + 3 * i);
        const uint64_t a = q[0], b = q[1], c = q[2];
        float4 o;
        o.x = luma_of((uint32_t)(a & 0xffff), (uint32_t)((a >> 16) & 0xffff), (uint32_t)((a >> 32) & 0xffff));
        o.y = luma_of((uint32_t)(a >> 48), (uint32_t)(b & 0xffff), (uint32_t)((b >> 16) & 0xffff));
        o.z = luma_of((uint32_t)((b >> 32) & 0xffff), (uint32_t)(b >> 48), (uint32_t)(c & 0xffff));
        o.w = luma_of((uint32_t)((c >> 16) & 0xffff), (uint32_t)((c >> 32) & 0xffff), (uint32_t)(c >> 48));
        *(float4 *)(y + i) = o;
        return;
    }
    for (size_t k = i; k < i + 4 && k < n; k++) y[k] = luma_of(rgb[3 * k], rgb[3 * k + 1], rgb[3 * k + 2]);
}

// align.rs:650-734: Burt REDUCE (a = 0.33) with the border renormalisation
__device__ __forceinline__ void burt_k(float *k) {
    float a = 0.33f, b = 0.25f, c = 0.25f - a / 2.0f;
    k[0] = c;
    k[1] = b;
    k[2] = a;
    k[3] = b;
    k[4] = c;
}

// The 64 x TY output tile at (ox0, oy0): the (2 TY + 3) x 131 input window
// staged in LDS (in-range samples only), the horizontal 5-tap sums of its rows
// at the 64 even columns into LDS, then each of the 256 threads sums five rows
// for each of its TY / 4 outputs and divides by the in-range weight sums -- per
// output the operations of the per-output loop of align.rs:683-734, in its
// order (taps outside the plane are skipped, not zero-filled).  Interior tiles
// skip the range tests.
#define BT_TX 64
#define BT_TY 4  // taller tiles measured no faster on a 45 MP plane
template <int TY>
struct BurtLds {
    float T[2 * TY + 3][2 * BT_TX + 3 + 1];
    float Hs[2 * TY + 3][BT_TX + 1];
};

// src(gx, gy): the input plane's sample at an in-range position (a stored
// plane, or a value computed on the spot)
template <int TY, bool EDGE, class Src>
__device__ __forceinline__ void burt_body(const Src &src, int w, int h, float *__restrict__ dst, int ow, int oh, int ox0,
                                          int oy0, BurtLds<TY> &S) {
    constexpr int IW = 2 * BT_TX + 3, IH = 2 * TY + 3;
    float k[5];
    burt_k(k);
    const int tid = threadIdx.x, ix0 = 2 * ox0 - 2, iy0 = 2 * oy0 - 2;
    lds_stage<IH * IW, 256, float>(
        [&](int q, float &v) {
            const int r = q / IW, c = q - r * IW, gy = iy0 + r, gx = ix0 + c;
            if (EDGE && !(gy >= 0 && gy < h && gx >= 0 && gx < w)) return false;
            v = src(gx, gy);
            return true;
        },
        [&](int q, float v) {
            const int r = q / IW;
            S.T[r][q - r * IW] = v;
        });
    __syncthreads();
    for (int q = tid; q < IH * BT_TX; q += 256) {
        const int r = q / BT_TX, jj = q - r * BT_TX, gy = iy0 + r;
        if (EDGE && (gy < 0 || gy >= h || ox0 + jj >= ow)) continue;
        const int cx = 2 * (ox0 + jj);
        float a = 0.0f;
#pragma unroll
        for (int u = 0; u < 5; u++) {
            const int i = cx + u - 2;
            if (!EDGE || (i >= 0 && i < w)) a += k[u] * S.T[r][2 * jj + u];
        }
        S.Hs[r][jj] = a;
    }
    __syncthreads();
    const int tx = tid & (BT_TX - 1), oj = ox0 + tx, cx = 2 * oj;
    if (EDGE && oj >= ow) return;
    float cn = 0.0f;
#pragma unroll
    for (int t = 0; t < 5; t++) {
        const int xx = cx + t - 2;
        if (!EDGE || (xx >= 0 && xx < w)) cn += k[t];
    }
#pragma unroll
    for (int qy = 0; qy < TY / 4; qy++) {
        const int ty = tid / BT_TX + 4 * qy, oi = oy0 + ty, cy = 2 * oi;
        if (EDGE && oi >= oh) break;
        float rn = 0.0f;
#pragma unroll
        for (int t = 0; t < 5; t++) {
            const int yy = cy + t - 2;
            if (!EDGE || (yy >= 0 && yy < h)) rn += k[t];
        }
        float acc = 0.0f;
#pragma unroll
        for (int t = 0; t < 5; t++) {
            const int yy = cy + t - 2;
            if (!EDGE || (yy >= 0 && yy < h)) acc += k[t] * S.Hs[2 * ty + t][tx];
        }
        dst[(size_t)oi * ow + oj] = acc / (rn * cn);
    }
}

template <int TY>
__global__ __launch_bounds__(256) void k_burt(const float *__restrict__ src, int w, int h, float *__restrict__ dst,
                                              int ow, int oh) {
    __shared__ BurtLds<TY> S;
    const int ox0 = blockIdx.x * BT_TX, oy0 = blockIdx.y * TY;
    const bool inner = ox0 >= 1 && oy0 >= 1 && 2 * ox0 + 2 * BT_TX <= w - 1 && 2 * oy0 + 2 * TY <= h - 1 &&
                       ox0 + BT_TX <= ow && oy0 + TY <= oh;
    const auto plane = [src, w](int gx, int gy) { return src[(size_t)gy * w + gx]; };
    if (inner)
        burt_body<TY, false>(plane, w, h, dst, ow, oh, ox0, oy0, S);
    else
        burt_body<TY, true>(plane, w, h, dst, ow, oh, ox0, oy0, S);
}

// align.rs:639-646
__device__ __forceinline__ void spline4f(float t, float *w) {
    w[0] = ((-1.0f / 3.0f * t + 0.8f) * t - 0.46666667f) * t;
    w[1] = ((t - 1.8f) * t - 0.2f) * t + 1.0f;
    w[2] = ((1.2f - t) * t + 0.8f) * t;
    w[3] = ((1.0f / 3.0f * t - 0.2f) * t - 0.13333334f) * t;
}

// align.rs:291-294 / :485-487: the source point of destination pixel (x, y)
__device__ __forceinline__ void src_point(const Inv9 &v, int affine, int x, int y, double &sx, double &sy) {
    double yf = (double)y, xf = (double)x;
    if (affine) {
        sx = (v.m[1] * yf + v.m[2]) + v.m[0] * xf;
        sy = (v.m[4] * yf + v.m[5]) + v.m[3] * xf;
    } else {
        double d = v.m[6] * xf + v.m[7] * yf + v.m[8];
        sx = (v.m[0] * xf + v.m[1] * yf + v.m[2]) / d;
        sy = (v.m[3] * xf + v.m[4] * yf + v.m[5]) / d;
    }
}

// threadIdx.x, read afresh where it is used: the addresses derived from it in
// an outer loop are then recomputed there rather than kept live across the
// whole search (where the register allocator spills them to scratch memory,
// and each reload waits on the memory system).
__device__ __forceinline__ int tid_here() {
    int t;
    asm volatile("v_mov_b32 %0, %1" : "=v"(t) : "v"((int)threadIdx.x));
    return t;
}

__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// v from the lane DPP control CTRL names (both halves by v_mov_b32 with DPP)
template <int CTRL>
__device__ __forceinline__ double dpp_get(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ __forceinline__ double lane_get(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(unsigned)b, l), hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// The sum of v over the 64 lanes of a wave (all of them active), the same
// value in every lane: within each row of 16 lanes a butterfly of DPP moves
// (adjacent lanes, then pairs, then rotations of the row by 4 and by 8 lanes),
// then the four rows' sums (lanes 0, 16, 32, 48) in row order.  A fixed tree,
// so every wave and every block forms the same value from the same inputs;
// each step is a register move and an add, where a shuffle through the LDS
// crossbar waits on its latency.
template <class Op>
__device__ __forceinline__ double wave_reduce(double v, Op op) {
    v = op(v, dpp_get<0xB1>(v));   // quad_perm [1, 0, 3, 2]
    v = op(v, dpp_get<0x4E>(v));   // quad_perm [2, 3, 0, 1]
    v = op(v, dpp_get<0x124>(v));  // row_ror:4
    v = op(v, dpp_get<0x128>(v));  // row_ror:8
    return op(op(op(lane_get(v, 0), lane_get(v, 16)), lane_get(v, 32)), lane_get(v, 48));
}
__device__ __forceinline__ double wave_total(double v) {
    return wave_reduce(v, [](double a, double b) { return a + b; });
}

// Sums NV doubles per thread over the grid into out[NV]: per-block partials,
// then the last block to finish folds them in block order (deterministic).
template <int NV>
__device__ void grid_sum(double (&v)[NV], double *__restrict__ part, unsigned *__restrict__ ticket,
                         double *__restrict__ out) {
    __shared__ double sh[NV][4];
    __shared__ bool last;
    int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int k = 0; k < NV; k++) {
        double s = wave_sum(v[k]);
        if (lane == 0) sh[k][wv] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 0; k < NV; k++) part[(size_t)blockIdx.x * NV + k] = ((sh[k][0] + sh[k][1]) + sh[k][2]) + sh[k][3];
        __threadfence();
        unsigned t = atomicAdd(ticket, 1u);
        last = t == gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    double acc[NV];
    for (int k = 0; k < NV; k++) acc[k] = 0.0;
    for (unsigned b = threadIdx.x; b < gridDim.x; b += blockDim.x)
        for (int k = 0; k < NV; k++) acc[k] += part[(size_t)b * NV + k];
    __syncthreads();
    for (int k = 0; k < NV; k++) {
        double s = wave_sum(acc[k]);
        if (lane == 0) sh[k][wv] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 0; k < NV; k++) out[k] = ((sh[k][0] + sh[k][1]) + sh[k][2]) + sh[k][3];
        *ticket = 0;
    }
}

// ------------------------------------------------------- the search on device
// multiscale_align (align.rs:896-942) as one cooperative kernel: every block
// runs the same bounded Nelder-Mead (align.rs:796-883) on the same numbers.
// Each cost evaluation is a grid-strided warp-and-reduce; once every block's
// arrival flag says its partial sums are written, every block folds them in
// the same fixed order, so all blocks hold the identical value and take the
// identical next step, with no host round trip and no broadcast.
#define NM_MAXL 16
#define NM_MAXG 256      // participating blocks at most (one arrival flag each; one per CU)
#define NM_GRID_STEP 16  // the grid is a multiple of this many blocks
// the search's flag words: the error flag, the join count, the hand-off flag
// and the closed participant count each on its own 128-byte line, then one
// arrival flag per participant (the number of passes it has completed)
#define NM_LINE 32
#define NM_FL_ERR (0 * NM_LINE)
#define NM_FL_JOIN (1 * NM_LINE)
#define NM_FL_HAND (2 * NM_LINE)
#define NM_FL_CLOSED (3 * NM_LINE)
#define NM_FL_ARR (4 * NM_LINE)
#define NM_FL_WORDS (4 * NM_LINE + NM_MAXG)
struct NmConst {
    const float *ref[NM_MAXL];
    const float *tgt[NM_MAXL];
    int lw[NM_MAXL], lh[NM_MAXL];
    int nlev, finest, nfree;
    int solo_max;  // levels of at most this many pixels are searched by block 0 alone
    int spec_max;  // grid levels of at most this many pixels evaluate r, e and c in one pass
    int free_idx[8];
    double step[NM_MAXL][8], tol[NM_MAXL][8];
    double span[8];
    // every wait ends (with the error flag raised) after wait_ticks of the
    // device wall clock; the first block closes the set of participants once
    // the whole grid has joined, or when none has joined for join_ticks
    uint64_t wait_ticks, join_ticks;
    int max_part;                    // at most this many blocks take part (0: the grid)
    int phases;                      // time the passes' phases (the clock reads cost a pass about 1 us)
    double *trace;                   // non-null: the counted costs, in the reference's order
    int trace_cap;
    unsigned long long *digest;      // per participant: a digest of every cost it folded
};

// align.rs:241-252 and :258-286 on the device.  CR: the rotation's sin and cos
// correctly rounded (sincos_cr, the host's bits: the maps the warps use);
// otherwise the device library's (the search's cost evaluations, where a last
// bit of the map moves a cost by about 1e-13 of itself and no decision).
template <bool CR>
__device__ void sim_inverse_dev(const double *s, int w, int h, double *inv) {
    double cx = (double)w / 2.0, cy = (double)h / 2.0;
    double c, sn;
    if constexpr (CR) {
        sincos_cr(s[3], &sn, &c);
    } else {
        sincos(s[3], &sn, &c);  // one argument reduction for both, the same bits as sin and cos
    }
    double sc = s[2];
    double a = sc * c, b = sc * (c * s[5] - sn * s[4]), d = sc * sn, e = sc * (sn * s[5] + c * s[4]);
    double gx = s[6] / (double)w, gy = s[7] / (double)h;
    double a2 = a + cx * gx, b2 = b + cx * gy, d2 = d + cy * gx, e2 = e + cy * gy;
    double m[9] = {a2, b2, cx + s[0] * (double)w - (a2 * cx + b2 * cy), d2, e2,
                   cy + s[1] * (double)h - (d2 * cx + e2 * cy), gx, gy, 1.0 - gx * cx - gy * cy};
    if (m[6] == 0.0 && m[7] == 0.0 && m[8] == 1.0) {
        double det = m[0] * m[4] - m[1] * m[3];
        double ia = m[4] / det, ib = -m[1] / det, id = -m[3] / det, ie = m[0] / det;
        inv[0] = ia;
        inv[1] = ib;
        inv[2] = -(ia * m[2] + ib * m[5]);
        inv[3] = id;
        inv[4] = ie;
        inv[5] = -(id * m[2] + ie * m[5]);
        inv[6] = 0.0;
        inv[7] = 0.0;
        inv[8] = 1.0;
        return;
    }
    double t[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            int r0 = (j + 1) % 3, r1 = (j + 2) % 3, c0 = (i + 1) % 3, c1 = (i + 2) % 3;
            t[i * 3 + j] = m[r0 * 3 + c0] * m[r1 * 3 + c1] - m[r0 * 3 + c1] * m[r1 * 3 + c0];
        }
    for (int i = 0; i < 9; i++) inv[i] = t[i] / t[8];
}

#define NM_THREADS 1024
#define NM_WAVES (NM_THREADS / 64)
#define NM_MAXP 9     // cost evaluations per collective pass
#define NM_TILE 64    // a pass walks 64-pixel-wide destination tiles of 16, 32 or 64 rows
#define NM_BOX 8192   // floats of a tile's staged source window
#ifndef NM_HB
#define NM_HB 30688   // floats a block holds a level's tiles in (their windows, and their reference pixels where they fit): the rest of the LDS
#endif
#define NM_HMAXT 8    // tiles a block holds at most
#define NM_HMARGIN 6  // pixels a held window extends past the first pass's bound
// the state words: the Sim, the evaluation total, per level the evaluations and
// block 0's ticks, the four phase tick totals, per level the passes, the
// participants and the traced costs
#define NM_ST_PART (13 + 3 * NM_MAXL)
#define NM_ST_TRACE (14 + 3 * NM_MAXL)
#define NM_ST_HAND (15 + 3 * NM_MAXL)  // block 0's Sim after the solo levels, handed to the grid
#define NM_ST_LPH (23 + 3 * NM_MAXL)   // per level the phase tick totals (NM_NPH of them)
#define NM_NPH 8
#define NM_STATE (23 + (3 + NM_NPH) * NM_MAXL)

struct NmShared {
    double S[9][8], FV[9], S2[9][8], FV2[9], XR[8], XE[8], XC[8], C[8], lo[8], hi[8], cur[8], FR;
    double PT[NM_MAXP][8], PC[NM_MAXP], inv[NM_MAXP][9];
    double red[3 * NM_MAXP][NM_WAVES], sums[3 * NM_MAXP];
    unsigned lev_evals[NM_MAXL], lev_passes[NM_MAXL];
    uint64_t ph[4], tmark;  // block 0's wall-clock ticks per phase: inverse, pixels, combine, steering
    // per level: the four phases, then the combine's parts (publishing the
    // partials, waiting for the other blocks, folding), then the steering's
    // branch (up to an iteration's set-up)
    uint64_t lph[NM_MAXL][NM_NPH];
    int box[5];             // the pass's window margins (x lo, x hi, y lo, y hi) and whether to stage
    int hwin[NM_HMAXT][8];  // per tile of a held level: its window (x0, y0, w, h, staged), held, window and reference offsets in hb
    int hoff;               // the held tiles' floats so far (the hold pass)
    int restage;            // some tile of the block is staged every pass (it did not fit)
    int grid;               // the participating blocks (read through nm_G, as a scalar)
    int rank;               // this block's place among them (nm_B)
    double tolv[8];         // the level's convergence tolerances
    int stop;
    int ctl;                // the next pass's point count (0: the level is done), from wave 0
    unsigned long long dig; // the digest of every cost this block has folded
    int ntrace;             // the costs traced so far (rank 0)
};

// The participating blocks and this block's rank among them, block-uniform
// (scalar registers).  A search takes the blocks that are running when it
// starts: normally the whole grid, but a plain launch has no promise that all
// its blocks are resident at once (other work on the device, a profiler that
// serialises dispatches), so each block takes a ticket as it starts and only
// the first gridDim.x tickets issued before the set closes do the work.
__device__ __forceinline__ int nm_G(const NmShared &sh) { return __builtin_amdgcn_readfirstlane(sh.grid); }
__device__ __forceinline__ int nm_B(const NmShared &sh) { return __builtin_amdgcn_readfirstlane(sh.rank); }

// align.rs:630-634

__device__ __forceinline__ double rms_of(double sd, double sd2, double cnt) {
    if (cnt < 16.0) return 1e9;
    double vv = (sd2 - sd * sd / cnt) / cnt;
    return sqrt(vv > 0.0 ? vv : 0.0);
}

// Orders LDS accesses between the lanes of wave 0, which runs the search's
// control alone: the lanes execute in step and a wave's LDS operations complete
// in issue order, so only the compiler has to be kept from moving accesses
// across this point.
__device__ __forceinline__ void wsync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ void zero_waves(NmShared &sh) {
    for (int i = threadIdx.x; i < 3 * NM_MAXP * NM_WAVES; i += NM_THREADS) (&sh.red[0][0])[i] = 0.0;
}

// The rows of a pass's tiles (64, 32, 16, 8 or 4; solo: 64): the fewest
// tiles for the busiest block (each tile costs its staging, its reference
// loads and its sums over the waves, while more rows per thread overlap),
// and among those the smaller tiles, which spread the level over more blocks
// (below 16 rows a tile leaves waves idle, and the block's SIMDs then run one
// wave each).
__device__ __forceinline__ int nm_tile_rows(int tw, int th, bool solo, int G) {
    if (solo) return 64;
    const int ntx = (tw + NM_TILE - 1) / NM_TILE;
    int best = 64, fewest = 0x7fffffff;
    for (int TH = 64; TH >= 4; TH >>= 1) {
        const int T = (ntx * ((th + TH - 1) / TH) + G - 1) / G;
        if (T <= fewest) {
            fewest = T;
            best = TH;
        }
    }
    return best;
}

// Whether a grid level's tiles are held (no block has more than NM_HMAXT),
// and whether their reference pixels are kept with them: when every tile of
// the busiest block fits with a window of a few pixels' bound.
__device__ __forceinline__ bool nm_holdable(int tw, int th, int G, bool *refh) {
    const int ntx = (tw + NM_TILE - 1) / NM_TILE, TH = nm_tile_rows(tw, th, false, G);
    const int T = (ntx * ((th + TH - 1) / TH) + G - 1) / G;
    const int win = (NM_TILE + 2 * NM_HMARGIN + 7) * (TH + 2 * NM_HMARGIN + 7);
    *refh = T * (NM_TILE * TH + win) <= NM_HB;
    return T <= NM_HMAXT;
}

// The pass's tiles, the sums into the waves' accumulators.  MODE 1: a
// resident solo level.  MODE 0: each tile's source window (the pass's bound)
// staged per pass, or, on a held level, staged by the level's first pass
// (hold) with the bound widened by NM_HMARGIN on each side for the moves to
// come, at the block's next free place in hb (together with the tile's
// reference pixels when refh), and read from there by every later pass; a
// tile that no longer fits there is staged by every pass instead.  When the
// block's windows all fit but not every tile's reference pixels beside them,
// the reference pixels go after the last window, tile by tile while they fit,
// into the rest of hb and the window buffer after it (free on a level whose
// windows are all held); the others are read from memory by every pass.
template <int MODE>
__device__ __forceinline__ void nm_tiles(const NmConst &K, NmShared &sh, float *boxs, float *hb, int lvl, int np,
                                         bool solo, bool held, bool hold, bool refh) {
    const int tw = K.lw[lvl], th = K.lh[lvl];
    const float *rf = K.ref[lvl], *tg = K.tgt[lvl];
    const double xmax = (double)(tw - 1), ymax = (double)(th - 1);
    const int G = nm_G(sh);
    // tiles column by column, so that the last (narrowest) column is what a
    // block takes as its second tile
    // (a resident level is one tile: its pixels strided over the block)
    const int npx = tw * th, ntx = (tw + NM_TILE - 1) / NM_TILE, TH = nm_tile_rows(tw, th, solo, G);
    const int R = MODE == 1 ? (npx + NM_THREADS - 1) / NM_THREADS : max(1, TH / 16);
    const int nty = (th + TH - 1) / TH, ntiles = MODE == 1 ? 1 : ntx * nty;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const bool act = MODE == 1 || ty < TH;  // the wave has a row of the tile
    // (the accumulators were cleared, and the pass's maps and bound written, by
    // wave 0 before the barrier that starts the pass)
    const int B = nm_B(sh);
    // the hold pass without refh: the block's windows' total size, and
    // whether they all fit (then reference pixels may follow them)
    int wsum = 0;
    bool allw = false;
    if (MODE == 0 && hold && !refh && sh.box[4]) {
        for (int t = B; t < ntiles; t += G) {
            const int X0 = (t / nty) * NM_TILE, Y0 = (t % nty) * TH;
            const int X1 = min(X0 + NM_TILE - 1, tw - 1), Y1 = min(Y0 + TH - 1, th - 1);
            const int bx0 = max(-1, X0 + sh.box[0] - NM_HMARGIN), by0 = max(-1, Y0 + sh.box[2] - NM_HMARGIN);
            const int bw = min(tw + 1, X1 + sh.box[1] + NM_HMARGIN) - bx0 + 1;
            const int bh = min(th + 1, Y1 + sh.box[3] + NM_HMARGIN) - by0 + 1;
            wsum += bw > 0 && bh > 0 ? bw * bh : NM_HB + 1;
        }
        allw = wsum <= NM_HB;
    }
    for (int t = solo ? 0 : B; t < ntiles; t += solo ? 1 : G) {
        const int X0 = (t / nty) * NM_TILE, Y0 = (t % nty) * TH;
        const int X1 = min(X0 + NM_TILE - 1, tw - 1), Y1 = min(Y0 + TH - 1, th - 1);
        int bx0 = 0, by0 = 0, bw = tw, bh = th, staged = 1;
        const float *win = boxs;  // the staged target samples (window origin at bx0, by0)
        float rv[4];
        if constexpr (MODE == 1) {
            bx0 = -1;
            by0 = -1;
            bw = tw + 3;
            bh = th + 3;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int i = (int)threadIdx.x + r * NM_THREADS;
                rv[r] = (r < R && i < npx) ? boxs[bw * bh + i] : 0.0f;
            }
        } else {
            const int q = held ? (t - B) / G : 0;  // the tile's place among the block's held ones
            int *hw = sh.hwin[q];
            // a held tile after the level's first pass: its window (and its
            // reference pixels, refh) where that pass put them in hb
            const bool kept = held && !hold && __builtin_amdgcn_readfirstlane(hw[5]);
            if (!kept) {
                __syncthreads();  // the bound is written; the previous tile's window is no longer read
                // the window may reach one sample before and two past the level
                // on each axis (a support's reach), where it holds the edge samples
                int m = hold ? NM_HMARGIN : 0, off = -1, roff = -1;
                bx0 = max(-1, X0 + sh.box[0] - m);
                by0 = max(-1, Y0 + sh.box[2] - m);
                bw = min(tw + 1, X1 + sh.box[1] + m) - bx0 + 1;
                bh = min(th + 1, Y1 + sh.box[3] + m) - by0 + 1;
                if (hold) {
                    // the level's first pass holds the tile at the block's next
                    // free place in hb when it fits; otherwise the tile is staged
                    // by every pass, with that pass's bound
                    const int rsz = NM_TILE * TH, need = (refh ? rsz : 0) + bw * bh, at = sh.hoff;
                    if (sh.box[4] && bw > 0 && bh > 0 && at + need <= NM_HB) {
                        off = refh ? at + rsz : at;
                        if (refh)
                            roff = at;
                        else if (allw && wsum + (q + 1) * rsz <= NM_HB + NM_BOX)
                            roff = wsum + q * rsz;
                    } else {
                        m = 0;
                        bx0 = max(-1, X0 + sh.box[0]);
                        by0 = max(-1, Y0 + sh.box[2]);
                        bw = min(tw + 1, X1 + sh.box[1]) - bx0 + 1;
                        bh = min(th + 1, Y1 + sh.box[3]) - by0 + 1;
                    }
                }
                float *wb = off >= 0 ? hb + off : boxs;
                staged = sh.box[4] && bw > 0 && bh > 0 && (off >= 0 || bw * bh <= NM_BOX);
                // (issuing all of a thread's window loads before storing any measured
                // no faster here and spilled registers)
                if (staged)
                    for (int i = threadIdx.x; i < bw * bh; i += NM_THREADS) {
                        const int r = i / bw, c = i - r * bw;
                        wb[i] = tg[(size_t)min(max(by0 + r, 0), th - 1) * tw + min(max(bx0 + c, 0), tw - 1)];
                    }
                win = wb;
                __syncthreads();  // every thread has read sh.hoff
                if (hold && threadIdx.x == 0) {
                    hw[0] = bx0;
                    hw[1] = by0;
                    hw[2] = bw;
                    hw[3] = bh;
                    hw[4] = staged;
                    hw[5] = off >= 0;
                    hw[6] = off;
                    hw[7] = roff;
                    if (off >= 0)
                        sh.hoff = off + bw * bh;
                    else
                        sh.restage = 1;
                }
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int x = X0 + tx, y = Y0 + ty + 16 * r;
                    rv[r] = (r < R && act && x < tw && y < th) ? rf[(size_t)y * tw + x] : 0.0f;
                    // kept for the later passes (each thread reads back its own)
                    if (roff >= 0 && r < R && act) hb[roff + (ty + 16 * r) * NM_TILE + tx] = rv[r];
                }
            } else {
                bx0 = hw[0];
                by0 = hw[1];
                bw = hw[2];
                bh = hw[3];
                staged = hw[4];
                const int off = __builtin_amdgcn_readfirstlane(hw[6]), roff = __builtin_amdgcn_readfirstlane(hw[7]);
                win = hb + off;
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int x = X0 + tx, y = Y0 + ty + 16 * r;
                    rv[r] = r < R && act ? (roff >= 0 ? hb[roff + (ty + 16 * r) * NM_TILE + tx]
                                                      : (x < tw && y < th ? rf[(size_t)y * tw + x] : 0.0f))
                                         : 0.0f;
                }
            }
        }
        // the window is the block's (scalar registers)
        bx0 = __builtin_amdgcn_readfirstlane(bx0);
        by0 = __builtin_amdgcn_readfirstlane(by0);
        bw = __builtin_amdgcn_readfirstlane(bw);
        bh = __builtin_amdgcn_readfirstlane(bh);
        staged = __builtin_amdgcn_readfirstlane(staged);
        // per point: the thread's 4 pixels, summed over the wave, added to the
        // wave's accumulators (one lane per wave writes them: no race)
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
        for (int p = 0; p < (act ? np : 0); p++) {
            Inv9 iv;
            for (int k = 0; k < 9; k++) iv.m[k] = sh.inv[p][k];
            const int aff = iv.m[6] == 0.0 && iv.m[7] == 0.0 && iv.m[8] == 1.0;
            double s0 = 0.0, s1 = 0.0, s2 = 0.0;
#pragma unroll 1
            for (int r = 0; r < R; r++) {
                int x = X0 + tx, y = Y0 + ty + 16 * r;
                if constexpr (MODE == 1) {
                    const int i = (int)threadIdx.x + r * NM_THREADS;
                    if (i >= npx) continue;
                    y = i / tw;
                    x = i - y * tw;
                }
                if (x >= tw || y >= th) continue;
                double sx, sy;
                src_point(iv, aff, x, y, sx, sy);
                if (!(sx >= 0.0 && sx <= xmax && sy >= 0.0 && sy <= ymax)) continue;
                double fx = floor(sx), fy = floor(sy);
                int x0 = (int)fx, y0 = (int)fy;
                float wx[4], wy[4];
                spline4f((float)(sx - fx), wx);
                spline4f((float)(sy - fy), wy);
                // the 4 x 4 support (clamped to the level) in the staged window
                // (whose edge samples stand for the clamped ones), or global reads
                float a = 0.0f;
                if (staged && x0 - 1 >= bx0 && x0 + 2 < bx0 + bw && y0 - 1 >= by0 && y0 + 2 < by0 + bh) {
                    const float *sup = win + (y0 - 1 - by0) * bw + (x0 - 1 - bx0);
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        float rr = 0.0f;
#pragma unroll
                        for (int q = 0; q < 4; q++) rr += wx[q] * sup[j * bw + q];
                        a += wy[j] * rr;
                    }
                } else {
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        const float *row = tg + (size_t)min(max(y0 + j - 1, 0), th - 1) * tw;
                        float rr = 0.0f;
#pragma unroll
                        for (int q = 0; q < 4; q++) rr += wx[q] * row[min(max(x0 + q - 1, 0), tw - 1)];
                        a += wy[j] * rr;
                    }
                }
                const double d = (double)(rv[r] - a);
                s0 += d;
                s1 += d * d;
                s2 += 1.0;
            }
            s0 = wave_total(s0);
            s1 = wave_total(s1);
            s2 = wave_total(s2);
            if (lane == 0) {
                sh.red[3 * p][wv] += s0;
                sh.red[3 * p + 1][wv] += s1;
                sh.red[3 * p + 2][wv] += s2;
            }
        }
    }
}

// The search's control runs in wave 0 alone, between two block barriers per
// pass: after the pixels, wave 0 publishes the pass's sums and waits for the
// other participants' (nm_publish), every wave folds them (nm_fold), and wave 0
// takes the costs (nm_costs), steers
// the simplex (nm_steer) and prepares the next pass's maps and window bound
// (nm_prep); the other waves wait at the barrier that starts the next pass.
// Within wave 0 the vertices, coordinates and points are spread over lanes.

// wave 0: the maps of the pass's points sh.PT[0..np) (the inverse homographies
// the warps sample through), and, when the pass stages windows, its bound: the
// displacement (source - destination) of an affine map is affine in the pixel,
// so over the level it is bounded by its values at the 4 corners (a lane per
// point and corner; min and max are exact in any order).  A projective point
// (or a wild one) turns the staging off for the pass.
__device__ void nm_prep(const NmConst &K, NmShared &sh, int lvl, int np, bool bound) {
    const int lane = threadIdx.x;
    const int tw = K.lw[lvl], th = K.lh[lvl];
    if (lane < np) {
        double v[8];
        for (int k = 0; k < 8; k++) v[k] = sh.cur[k];
        for (int k = 0; k < K.nfree; k++) v[K.free_idx[k]] = sh.PT[lane][k];
        sim_inverse_dev<false>(v, tw, th, sh.inv[lane]);
    }
    if (!bound) return;
    wsync();
    double mnx = 1e300, mxx = -1e300, mny = 1e300, mxy = -1e300;
    int ok = 1;
    if (lane < 4 * np) {
        const int p = lane >> 2, c = lane & 3;
        Inv9 iv;
        for (int k = 0; k < 9; k++) iv.m[k] = sh.inv[p][k];
        if (!(iv.m[6] == 0.0 && iv.m[7] == 0.0 && iv.m[8] == 1.0)) ok = 0;
        const int cx = (c & 1) ? tw - 1 : 0, cy = (c & 2) ? th - 1 : 0;
        double sx, sy;
        src_point(iv, 1, cx, cy, sx, sy);
        mnx = fmin(mnx, sx - cx);
        mxx = fmax(mxx, sx - cx);
        mny = fmin(mny, sy - cy);
        mxy = fmax(mxy, sy - cy);
    }
    auto mn = [](double a, double b) { return fmin(a, b); };
    auto mx = [](double a, double b) { return fmax(a, b); };
    mnx = wave_reduce(mnx, mn);
    mxx = wave_reduce(mxx, mx);
    mny = wave_reduce(mny, mn);
    mxy = wave_reduce(mxy, mx);
    ok = __all(ok) && mnx <= mxx && mny <= mxy && mnx > -256.0 && mxx < 256.0 && mny > -256.0 && mxy < 256.0;
    if (lane == 0) {
        sh.box[0] = ok ? (int)floor(mnx) - 1 : 0;  // window = tile + [lo, hi] on each axis
        sh.box[1] = ok ? (int)floor(mxx) + 2 : 0;
        sh.box[2] = ok ? (int)floor(mny) - 1 : 0;
        sh.box[3] = ok ? (int)floor(mxy) + 2 : 0;
        sh.box[4] = ok;
    }
}

// The combine of a pass's sums, after its pixels.  Wave 0 sums each value over
// the block's waves (in wave order); on a solo level that is the pass's sum.
// On the grid, wave 0 publishes the block's partial sums (one lane per value,
// one store instruction) as agent-scope atomic stores, written through to the
// level every XCD reads coherently, stored value-major (value k of block b at
// k * G + b, so that one value of 64 consecutive blocks is 512 contiguous
// bytes); their completion is waited for explicitly (s_waitcnt vmcnt(0): on
// gfx9 a store leaves the vector memory count only once the write is
// acknowledged at its scope) before the block's arrival flag is raised to the
// number of passes it has completed; the workgroup release fence keeps the
// compiler from moving the stores past the flag, but emits no wait of its own
// on this target.  Wave 0 reads all flags in one round trip per poll (a lane
// per block, up to NM_MAXG / 64 each).  Then (nm_fold) wave w loads value w
// (and w + 16, ...) of every block with agent-scope atomic loads, issued after
// the barrier that follows the wait, lane l holding blocks l, l + 64, ..., and
// sums it over them in block order and then over the lanes by a fixed tree:
// an order that depends on lane indices only, so every block holds the
// identical costs with no broadcast.  A block a pass ahead writes the other
// slot; none can be two ahead before every block has arrived at the pass
// between.  No XCD's L2 is written back or invalidated for the exchange: an
// agent-scope release / acquire does both to the whole L2 on every pass, which
// here (the side work's dirty lines in it, the level's samples lost from it)
// cost more than a small level's pass.  A flag per block (not a shared
// counter) keeps the arrival a plain store.

// wave 0: the block's sums of the pass (solo: into sh.sums); on the grid their
// publication and the wait for every participant's.  False when the search
// stops (another block's failure or a wait's time-out).
__device__ bool nm_publish(const NmConst &K, NmShared &sh, int lvl, int np, bool solo, double *part, unsigned *ctr,
                           unsigned *err, unsigned gen) {
    const int lane = threadIdx.x;
    const int nv = 3 * np;
    const bool tm = K.phases != 0;
    double t = 0.0;
    if (lane < nv) {
        for (int q = 0; q < NM_WAVES; q++) t += sh.red[lane][q];
        for (int q = 0; q < NM_WAVES; q++) sh.red[lane][q] = 0.0;  // cleared for the next pass
    }
    if (lane == 0) sh.lev_passes[lvl]++;
    if (solo) {
        if (lane < nv) sh.sums[lane] = t;
        return true;
    }
    const uint64_t tq2 = tm ? wall_clock64() : 0;
    const int G = nm_G(sh), B = nm_B(sh);
    if (lane < nv)
        __hip_atomic_store(&part[(size_t)(gen & 1) * G * 3 * NM_MAXP + (size_t)lane * G + B], t, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_s_waitcnt(0);
    unsigned *arr = ctr + NM_FL_ARR;
    const unsigned target = gen + 1;
    if (lane == 0) __hip_atomic_store(&arr[B], target, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint64_t tp = 0;
    if (tm) {
        tp = wall_clock64();
        if (lane == 0) sh.lph[lvl][4] += tp - tq2;
    }
    // bounded, and left when any block has raised the error flag (looked
    // at, with the clock, every 64 polls; the clock starts at the first look)
    uint64_t t_in = 0;
    unsigned spins = 0;
    bool stop = false;
    for (;;) {
        unsigned v[NM_MAXG / 64];
#pragma unroll
        for (int j = 0; j < NM_MAXG / 64; j++) {
            const int b = lane + 64 * j;
            v[j] = b < G ? __hip_atomic_load(&arr[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : target;
        }
        int all = 1;
#pragma unroll
        for (int j = 0; j < NM_MAXG / 64; j++) all &= v[j] >= target;
        if (__all(all)) break;
        __builtin_amdgcn_s_sleep(1);
        if ((++spins & 63u) == 0) {
            // another block's failure, or this wait's time-out, ends the search
            if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                stop = true;
                break;
            }
            const uint64_t now = wall_clock64();
            if (spins == 64u) t_in = now;
            if (now - t_in > K.wait_ticks) {
                if (lane == 0) atomicOr(err, 1u);
                stop = true;
                break;
            }
        }
    }
    if (lane == 0) {
        if (stop) sh.stop = 1;
        if (tm) {
            sh.tmark = wall_clock64();
            sh.lph[lvl][5] += sh.tmark - tp;
        }
    }
    return !stop;
}

// every wave, once every participant's partials are in place: value k = w,
// w + 16, ... of the pass summed over the blocks into sh.sums[k]
__device__ void nm_fold(NmShared &sh, int np, const double *part, unsigned gen) {
    const int nv = 3 * np, G = nm_G(sh);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const double *pp = part + (size_t)(gen & 1) * G * 3 * NM_MAXP;
    for (int k = wv; k < nv; k += NM_WAVES) {
        // every load issued before the first is used: unconditional, from an
        // address clamped into the value's run (a block past the grid is
        // loaded and left out of the sum)
        double v[NM_MAXG / 64];
#pragma unroll
        for (int u = 0; u < NM_MAXG / 64; u++)
            v[u] = __hip_atomic_load(&pp[(size_t)k * G + min(lane + 64 * u, G - 1)], __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_AGENT);
        double a = 0.0;
#pragma unroll
        for (int u = 0; u < NM_MAXG / 64; u++)
            if (lane + 64 * u < G) a += v[u];
        a = wave_total(a);
        if (lane == 0) sh.sums[k] = a;
    }
}

// wave 0: the costs (align.rs:630-634) of the pass's np points into sh.PC;
// on the grid the digest of the sums this block folded (FNV-1a over their
// bits): every participant must end a search with the same one
__device__ void nm_costs(const NmConst &K, NmShared &sh, int lvl, int np, bool solo) {
    const int lane = threadIdx.x;
    if (!solo && lane == 0) {
        unsigned long long d = sh.dig;
        for (int k = 0; k < 3 * np; k++)
            d = (d ^ (unsigned long long)__double_as_longlong(sh.sums[k])) * 1099511628211ull;
        sh.dig = d;
        if (K.phases) sh.lph[lvl][6] += wall_clock64() - sh.tmark;
    }
    wsync();
    if (lane < np) sh.PC[lane] = rms_of(sh.sums[3 * lane], sh.sums[3 * lane + 1], sh.sums[3 * lane + 2]);
    wsync();
}

// A counted cost into the trace (thread 0 of rank 0, when tracing), in the
// order the reference evaluates them.
__device__ __forceinline__ void nm_trace(const NmConst &K, NmShared &sh, double v) {
    if (!K.trace || nm_B(sh) != 0) return;
    if (sh.ntrace < K.trace_cap) K.trace[sh.ntrace] = v;
    sh.ntrace++;
}

__device__ __forceinline__ double clampd_dev(double v, double lo, double hi) {
    if (v < lo) v = lo;
    if (v > hi) v = hi;
    return v;
}

#define CLAMPK(v, k) clampd_dev((v), sh.lo[k], sh.hi[k])
enum { NM_INIT, NM_ITER, NM_EXPAND, NM_CONTRACT, NM_SHRINK };

// wave 0: one step of a level's bounded Nelder-Mead (align.rs:796-883) once a
// pass's costs are in sh.PC: the branch the pass was for, and when the
// iteration is over the next one's set-up (the stable sort, the convergence
// test, the centroid and the trial points).  Returns the next pass's point
// count, or 0 when the level is done.  The
// iterates and the counted evaluations are the reference's: on the grid levels
// where a pass costs mostly synchronisation (spec), an iteration evaluates the
// reflection, the expansion and the contraction in one pass (the latter two
// are fixed by the centroid and the worst vertex before the reflection's cost
// is known) and then takes the reference's branch, and a speculative cost the
// branch does not use is not counted; elsewhere they are evaluated one by one
// as the reference does.  A shrink evaluates its n new vertices in one pass.
__device__ int nm_steer(const NmConst &K, NmShared &sh, int lvl, int &state, int &it, bool spec) {
    const int n = K.nfree, lane = threadIdx.x;
    const bool l0 = lane == 0, tm = K.phases != 0;
    const uint64_t t_c = tm ? wall_clock64() : 0;
    int next = -1;
    if (state == NM_INIT) {
        if (lane <= n) sh.FV[lane] = sh.PC[lane];
        if (l0) {
            for (int i = 0; i <= n; i++) nm_trace(K, sh, sh.PC[i]);
            sh.lev_evals[lvl] += n + 1;
        }
    } else if (state == NM_ITER) {
        const double fr = sh.PC[0], fe = sh.PC[1], fc = sh.PC[2];
        const int br = fr < sh.FV[0] ? 0 : (fr < sh.FV[n - 1] ? 1 : 2);
        const int bc = spec && br == 2 ? (fc < sh.FV[n] ? 2 : 3) : br;
        wsync();  // every lane has read the branch's inputs
        const double *src = nullptr;
        double fv = 0.0;
        if (br == 1) {  // reflect
            src = sh.XR;
            fv = fr;
        } else if (!spec) {  // the expansion or the contraction is the next pass
            if (lane < n) sh.PT[0][lane] = br == 0 ? sh.XE[lane] : sh.XC[lane];
        } else if (br == 0) {  // expand (align.rs:838-849)
            src = fe < fr ? sh.XE : sh.XR;
            fv = fe < fr ? fe : fr;
        } else if (bc == 2) {  // contract (align.rs:853-862)
            src = sh.XC;
            fv = fc;
        }
        if (src) {
            if (lane < n) sh.S[n][lane] = src[lane];
            if (l0) sh.FV[n] = fv;
        }
        if (l0) {
            sh.FR = fr;
            sh.lev_evals[lvl] += 1;
            nm_trace(K, sh, fr);
            if (spec && br != 1) {
                sh.lev_evals[lvl] += 1;
                nm_trace(K, sh, br == 0 ? fe : fc);
            }
        }
        if (!spec && br != 1) next = br == 0 ? NM_EXPAND : NM_CONTRACT;
        if (spec && bc == 3) next = NM_SHRINK;
    } else if (state == NM_EXPAND) {
        const double fe = sh.PC[0], fr = sh.FR;
        const double *src = fe < fr ? sh.XE : sh.XR;
        if (lane < n) sh.S[n][lane] = src[lane];
        if (l0) {
            sh.FV[n] = fe < fr ? fe : fr;
            sh.lev_evals[lvl] += 1;
            nm_trace(K, sh, fe);
        }
    } else if (state == NM_CONTRACT) {
        const double fc = sh.PC[0];
        const bool take = fc < sh.FV[n];
        wsync();
        if (take) {
            if (lane < n) sh.S[n][lane] = sh.XC[lane];
            if (l0) sh.FV[n] = fc;
        }
        if (l0) {
            sh.lev_evals[lvl] += 1;
            nm_trace(K, sh, fc);
        }
        if (!take) next = NM_SHRINK;
    } else {  // NM_SHRINK: the n new vertices' costs
        if (lane >= 1 && lane <= n) sh.FV[lane] = sh.PC[lane - 1];
        if (l0) {
            for (int i = 1; i <= n; i++) nm_trace(K, sh, sh.PC[i - 1]);
            sh.lev_evals[lvl] += n;
        }
    }
    wsync();
    if (next == NM_SHRINK && lane < n * n) {  // shrink towards the best (align.rs:863-870)
        const int i = 1 + lane / n, k = lane % n;
        const double v = CLAMPK(sh.S[0][k] + 0.5 * (sh.S[i][k] - sh.S[0][k]), k);
        sh.S[i][k] = v;
        sh.PT[i - 1][k] = v;
    }
    wsync();
    const bool stop = sh.stop != 0;
    if (next >= 0 && !stop) {
        state = next;
        if (tm && l0) {
            const uint64_t t_e = wall_clock64();
            sh.lph[lvl][7] += t_e - t_c;
            sh.ph[3] += t_e - t_c;
            sh.lph[lvl][3] += t_e - t_c;
        }
        return next == NM_SHRINK ? n : 1;
    }
    if (stop || it == 200) return 0;
    if (tm && l0) sh.lph[lvl][7] += wall_clock64() - t_c;  // the branch's share of the steering
    it++;
    // the iteration's set-up (every step keeps the reference's arithmetic and
    // order per coordinate):
    // 1. the stable sort of the vertices by value (align.rs:816-818): vertex
    //    i goes to its rank (values below it, then equal values before it)
    if (lane <= n) {
        const double f = sh.FV[lane];
        int r = 0;
        for (int j = 0; j <= n; j++) r += sh.FV[j] < f || (sh.FV[j] == f && j < lane);
        for (int k = 0; k < n; k++) sh.S2[r][k] = sh.S[lane][k];
        sh.FV2[r] = f;
    }
    wsync();
    for (int t = lane; t < (n + 1) * 8; t += 64) {
        const int i = t >> 3, k = t & 7;
        if (k < n) sh.S[i][k] = sh.S2[i][k];
        if (k == 0) sh.FV[i] = sh.FV2[i];
    }
    wsync();
    // 2. convergence (align.rs:888-894): the value spread and every vertex
    //    within tol of the best, a (vertex, coordinate) per lane
    int ok = 1;
    if (l0) ok = !(fabs(sh.FV[n] - sh.FV[0]) > 1e-4 * (1.0 + fabs(sh.FV[0])));
    for (int t = lane; t < (n + 1) * 8; t += 64) {
        const int i = t >> 3, k = t & 7;
        if (i >= 1 && k < n && !(fabs(sh.S[i][k] - sh.S[0][k]) <= sh.tolv[k])) ok = 0;
    }
    // 3. per coordinate: the centroid of the best n, the reflection, the
    //    expansion and the contraction (align.rs:823-862)
    if (lane < n) {
        const int k = lane;
        double c = 0.0;
        for (int i = 0; i < n; i++) c += sh.S[i][k];
        c /= (double)n;
        sh.C[k] = c;
        const double xr = CLAMPK(c + 1.0 * (c - sh.S[n][k]), k);
        const double xe = CLAMPK(c + 2.0 * (xr - c), k);
        const double xc = CLAMPK(c + 0.5 * (sh.S[n][k] - c), k);
        sh.XR[k] = sh.PT[0][k] = xr;
        sh.XE[k] = sh.PT[1][k] = xe;
        sh.XC[k] = sh.PT[2][k] = xc;
    }
    if (__all(ok)) return 0;
    state = NM_ITER;
    wsync();
    if (tm && l0) {
        const uint64_t t_e = wall_clock64();
        sh.ph[3] += t_e - t_c;
        sh.lph[lvl][3] += t_e - t_c;
    }
    return spec ? 3 : 1;
}

// One level's bounded Nelder-Mead (align.rs:796-883), collective over the
// block (solo) or the grid.  A pass (the costs of up to NM_MAXP points at the
// level, align.rs:581-635) walks the level in 64-wide destination tiles
// (nm_tiles); each tile's source window (the union over the points of the
// bounding box of the tile corners' source points, plus the 4 x 4 support) is
// staged in LDS when it fits, and every tap inside it is read from there, any
// other from global memory.  solo: block 0 alone over all tiles; otherwise
// every block over its grid-strided share, the blocks' partial sums combined by
// every block in the same order (nm_publish, nm_fold).  Batching the points does not
// change any cost: each point's sums are the ones a pass of its own would form.
// resident (solo levels small enough): the whole target level and the
// reference level are staged in boxs once for the level, so a pass reads no
// global memory.  held: a grid level whose tiles the block holds in hb
// (nm_tiles), staged by the level's first pass.  The taps and the samples they
// read are the same in every case (a window holds the level's own samples, and
// a tap outside it is read from memory); the order of the sums follows the
// level's tiling.
__device__ void nm_level(const NmConst &K, NmShared &sh, float *boxs, float *hb, int lvl, bool solo, double *part,
                         unsigned *ctr, unsigned *err, unsigned &gen) {
    const int n = K.nfree;
    const bool w0 = threadIdx.x < 64;
    const int lane = threadIdx.x;  // wave 0's lanes
    const bool spec = !solo && (size_t)K.lw[lvl] * K.lh[lvl] <= (size_t)K.spec_max;
    const double *step = K.step[lvl];
    // a solo level whose target and reference fit the window buffer together
    // is staged once and read from there by every pass: the target with one
    // edge sample before and two past it on each axis ((w + 3) x (h + 3), as
    // a tile's window holds them), then the reference
    const int lw = K.lw[lvl], lh = K.lh[lvl], npx = lw * lh, pw = lw + 3, np3 = pw * (lh + 3);
    const bool resident = solo && np3 + npx <= NM_BOX;
    if (resident) {
        const float *tg = K.tgt[lvl], *rf = K.ref[lvl];
        constexpr int SU = NM_BOX / NM_THREADS;
        float v[SU];
#pragma unroll
        for (int u = 0; u < SU; u++) {
            const int i = threadIdx.x + u * NM_THREADS;
            if (i < np3) {
                const int r = i / pw, c = i - r * pw;
                v[u] = tg[min(max(r - 1, 0), lh - 1) * lw + min(max(c - 1, 0), lw - 1)];
            } else if (i < np3 + npx) {
                v[u] = rf[i - np3];
            }
        }
#pragma unroll
        for (int u = 0; u < SU; u++) {
            const int i = threadIdx.x + u * NM_THREADS;
            if (i < np3 + npx) boxs[i] = v[u];
        }
    }
    bool refh = false;
    const bool held = !solo && nm_holdable(K.lw[lvl], K.lh[lvl], nm_G(sh), &refh);
    int state = NM_INIT, it = 0, nxt = n + 1;  // wave 0's
    if (w0) {
        if (lane < n) sh.tolv[lane] = K.tol[lvl][lane];  // read by the iterations' set-up
        if (lane == 0) {
            sh.hoff = 0;  // the level's held tiles start over
            sh.restage = 0;
        }
        // the initial simplex (align.rs:800-812): the start, and a step along
        // each free axis from it, clamped to the box; a vertex per lane
        if (lane <= n)
            for (int k = 0; k < n; k++) {
                double v = sh.cur[K.free_idx[k]];
                if (lane >= 1) {
                    if (k == lane - 1) v += step[k];
                    v = CLAMPK(v, k);
                }
                sh.S[lane][k] = v;
                sh.PT[lane][k] = v;
            }
        wsync();
    }
    for (bool first = true;; first = false) {
        if (w0) {
            // the next pass's maps and window bound (the level's first pass
            // stages the held tiles), or the end of the level
            const uint64_t t_e = K.phases ? wall_clock64() : 0;
            if (nxt) nm_prep(K, sh, lvl, nxt, !resident && (first || !held || sh.restage));
            if (lane == 0) {
                sh.ctl = nxt;
                if (K.phases) {
                    sh.tmark = wall_clock64();
                    sh.ph[0] += sh.tmark - t_e;
                    sh.lph[lvl][0] += sh.tmark - t_e;
                }
            }
        }
        __syncthreads();
        const int np = __builtin_amdgcn_readfirstlane(sh.ctl);
        if (np == 0) break;
        if (resident)
            nm_tiles<1>(K, sh, boxs, hb, lvl, np, solo, false, false, false);
        else
            nm_tiles<0>(K, sh, boxs, hb, lvl, np, solo, held, held && first, refh);
        __syncthreads();  // every wave's sums are in sh.red
        uint64_t tc = 0;
        if (w0) {
            if (K.phases && lane == 0) {
                tc = wall_clock64();
                sh.ph[1] += tc - sh.tmark;
                sh.lph[lvl][1] += tc - sh.tmark;
                sh.tmark = tc;
            }
            nm_publish(K, sh, lvl, np, solo, part, ctr, err, gen);
        }
        if (!solo) {
            __syncthreads();  // every participant's partials are in place, or the search stops
            if (!sh.stop) nm_fold(sh, np, part, gen);
            gen++;
            __syncthreads();
        }
        if (w0) {
            nxt = 0;
            if (!sh.stop) {
                nm_costs(K, sh, lvl, np, solo);
                if (K.phases && lane == 0) {
                    const uint64_t t = wall_clock64();
                    sh.ph[2] += t - tc;
                    sh.lph[lvl][2] += t - tc;
                }
                nxt = nm_steer(K, sh, lvl, state, it, spec);
            }
        }
    }
    if (w0 && lane < n) {
        int best = 0;
        for (int i = 1; i <= n; i++)
            if (sh.FV[i] < sh.FV[best]) best = i;
        sh.cur[K.free_idx[lane]] = sh.S[best][lane];
    }
    __syncthreads();
}
#undef CLAMPK

// state: [0..8) the initial Sim (the guess) in, the found Sim out; [8] the
// evaluations; [9 + l] the evaluations at level l; [9 + NM_MAXL + l] rank 0's
// wall-clock ticks at level l; then the phase ticks, the passes per level,
// the participants, the traced costs and the hand-off Sim.  The coarse levels
// (up to solo_max pixels) are searched by rank 0 alone; it then hands the Sim
// to the grid (state[NM_ST_HAND..] and the hand-off flag), and all
// participants search the fine levels together.
__global__ __launch_bounds__(NM_THREADS) void k_nm_search(const NmConst *__restrict__ Kp, double *__restrict__ state,
                                                          double *__restrict__ part, unsigned *__restrict__ ctr,
                                                          unsigned *__restrict__ err) {
    __shared__ NmShared sh;
    // a held level's tiles (hb), then a tile's source window (boxs), one
    // array: the reference pixels of a level whose windows are all held may
    // run on from hb into boxs
    __shared__ __attribute__((aligned(16))) float hbx[NM_HB + NM_BOX];
    float *const hb = hbx, *const boxs = hbx + NM_HB;
    const NmConst &K = *Kp;
    const int n = K.nfree;
    unsigned gen = 0;
    const bool t0 = threadIdx.x == 0;
    zero_waves(sh);  // the waves' accumulators (wave 0 clears them after each pass)
    if (t0) {
        sh.stop = 0;
        sh.rank = -1;
        sh.grid = 0;
        // an earlier search of a stream that failed leaves the flag raised:
        // every later one leaves at once (the host reports the failure)
        if (!__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
            // the participants: a ticket per block as it starts; ticket 0
            // closes the set once every block it may take has joined, or
            // when none has joined for join_ticks, and publishes its size
            const unsigned cap = K.max_part > 0 ? min(gridDim.x, (unsigned)K.max_part) : gridDim.x;
            const unsigned tk = __hip_atomic_fetch_add(&ctr[NM_FL_JOIN], 1u, __ATOMIC_RELAXED,
                                                       __HIP_MEMORY_SCOPE_AGENT);
            unsigned G = 0;
            if (tk == 0) {
                unsigned seen = 1;
                uint64_t t_last = wall_clock64();
                for (;;) {
                    const unsigned j = __hip_atomic_load(&ctr[NM_FL_JOIN], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if (j >= cap) {
                        seen = cap;
                        break;
                    }
                    if (j != seen) {
                        seen = j;
                        t_last = wall_clock64();
                    } else if (wall_clock64() - t_last > K.join_ticks) {
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
                G = seen;
                __hip_atomic_store(&ctr[NM_FL_CLOSED], G | 0x80000000u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            } else if (tk < cap) {
                const uint64_t t_in = wall_clock64();
                unsigned v = 0, spins = 0;
                while ((v = __hip_atomic_load(&ctr[NM_FL_CLOSED], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) == 0u) {
                    __builtin_amdgcn_s_sleep(1);
                    if ((++spins & 63u) == 0 && wall_clock64() - t_in > K.wait_ticks) {
                        atomicOr(err, 1u);
                        break;
                    }
                }
                G = v & 0x7fffffffu;
            }
            if (tk < G) {
                sh.rank = (int)tk;
                sh.grid = (int)G;
            }
        }
        if (sh.rank >= 0) {
            for (int k = 0; k < 8; k++) sh.cur[k] = state[k];
            for (int k = 0; k < n; k++) {  // the search box about the guess (align.rs:912-913)
                double iv = state[K.free_idx[k]];
                sh.lo[k] = iv - K.span[K.free_idx[k]];
                sh.hi[k] = iv + K.span[K.free_idx[k]];
            }
            for (int l = 0; l < NM_MAXL; l++) sh.lev_evals[l] = sh.lev_passes[l] = 0;
            for (int q = 0; q < 4; q++) sh.ph[q] = 0;
            for (int l = 0; l < NM_MAXL; l++)
                for (int q = 0; q < NM_NPH; q++) sh.lph[l][q] = 0;
            sh.dig = 14695981039346656037ull;
            sh.ntrace = 0;
            sh.tmark = wall_clock64();
            if (sh.rank == 0)
                for (int l = 0; l < NM_MAXL; l++) state[9 + NM_MAXL + l] = 0.0;
        }
    }
    __syncthreads();
    if (sh.rank < 0) return;  // not a participant (late, beyond the cap, or a failed stream)
    const bool r0 = nm_B(sh) == 0;
    // the levels, coarse to fine; those of at most solo_max pixels (the
    // coarsest) by rank 0 alone, which then hands the Sim to the grid
    bool handed = false;
    for (int lvl = K.nlev - 1; n > 0 && lvl >= K.finest && !sh.stop; lvl--) {
        const bool solo = (size_t)K.lw[lvl] * K.lh[lvl] <= (size_t)K.solo_max;
        if (!solo && !handed) {
            handed = true;
            unsigned *hflag = ctr + NM_FL_HAND;
            if (t0) {
                if (r0) {
                    for (int k = 0; k < 8; k++)
                        __hip_atomic_store(&state[NM_ST_HAND + k], sh.cur[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    __threadfence();
                    __hip_atomic_store(hflag, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                } else {
                    const uint64_t t_in = wall_clock64();
                    unsigned spins = 0;
                    while (__hip_atomic_load(hflag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u) {
                        if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
                        __builtin_amdgcn_s_sleep(1);
                        if ((++spins & 63u) == 0 && wall_clock64() - t_in > K.wait_ticks) {
                            atomicOr(err, 1u);
                            break;
                        }
                    }
                    __threadfence();
                    for (int k = 0; k < 8; k++)
                        sh.cur[k] = __hip_atomic_load(&state[NM_ST_HAND + k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) sh.stop = 1;
                }
            }
            __syncthreads();
            if (sh.stop) break;
        }
        if (solo && !r0) continue;
        uint64_t t = wall_clock64();
        nm_level(K, sh, boxs, hb, lvl, solo, part, ctr, err, gen);
        if (t0 && r0) state[9 + NM_MAXL + lvl] = (double)(wall_clock64() - t);
    }
    if (t0 && K.digest) K.digest[nm_B(sh)] = sh.dig;
    if (r0 && t0) {
        double tot = 0.0;
        for (int l = 0; l < NM_MAXL; l++) {
            state[9 + l] = (double)sh.lev_evals[l];
            tot += (double)sh.lev_evals[l];
        }
        for (int k = 0; k < 8; k++) state[k] = sh.cur[k];
        state[8] = tot;
        for (int q = 0; q < 4; q++) state[9 + 2 * NM_MAXL + q] = (double)sh.ph[q];
        for (int l = 0; l < NM_MAXL; l++)
            for (int q = 0; q < NM_NPH; q++) state[NM_ST_LPH + NM_NPH * l + q] = (double)sh.lph[l][q];
        for (int l = 0; l < NM_MAXL; l++) state[13 + 2 * NM_MAXL + l] = (double)sh.lev_passes[l];
        state[NM_ST_PART] = (double)nm_G(sh);
        state[NM_ST_TRACE] = (double)sh.ntrace;
    }
}

// clears the arrival flags, the join count and the hand-off flag before a
// search (a kernel, so the sequence can be captured in a graph), and the error
// flag unless keep_err (a stream's later frames: one failed search fails the
// rest at once, and the host reports it after the last frame)
__global__ void k_nm_reset(unsigned *flags, int keep_err) {
    for (int i = threadIdx.x; i < NM_FL_WORDS; i += blockDim.x)
        if (!(keep_err && i == NM_FL_ERR)) flags[i] = 0;
}

// align.rs:416-474 in f64: the interpolation weights of kernel `ik` at fraction t
__device__ double keys_k(double d) {
    if (d < 1.0) return (1.5 * d - 2.5) * d * d + 1.0;
    if (d < 2.0) return ((-0.5 * d + 2.5) * d - 4.0) * d + 2.0;
    return 0.0;
}
__device__ double spline36_k(double d) {
    if (d < 1.0) return ((13.0 / 11.0 * d - 453.0 / 209.0) * d - 3.0 / 209.0) * d + 1.0;
    if (d < 2.0) {
        double u = d - 1.0;
        return ((-6.0 / 11.0 * u + 270.0 / 209.0) * u - 156.0 / 209.0) * u;
    }
    if (d < 3.0) {
        double u = d - 2.0;
        return ((1.0 / 11.0 * u - 45.0 / 209.0) * u + 26.0 / 209.0) * u;
    }
    return 0.0;
}
__device__ double lanczos3_k(double d) {
    if (d < 1e-9) return 1.0;
    if (d < 3.0) {
        double a = 3.14159265358979323846 * d, b = 3.14159265358979323846 * d / 3.0;
        return sin(a) / a * (sin(b) / b);
    }
    return 0.0;
}
// N: the kernel's tap count (2: ik 0-1, 4: ik 2-3, 6: ik 4-5), so that each
// warp's instantiation holds only the kernels of its width (the 6-tap ones
// bring sines and divisions whose registers the 2-tap warps would pay for)
template <int N>
__device__ __forceinline__ void interp_w(int ik, double t, float *wf) {
    double w[N];
    if constexpr (N == 2) {
        if (ik == 0) {
            w[0] = t < 0.5 ? 1.0 : 0.0;
            w[1] = t < 0.5 ? 0.0 : 1.0;
        } else {
            w[0] = 1.0 - t;
            w[1] = t;
        }
    } else if constexpr (N == 4) {
        if (ik == 3) {
            w[0] = ((-1.0 / 3.0 * t + 0.8) * t - 0.46666667) * t;
            w[1] = ((t - 1.8) * t - 0.2) * t + 1.0;
            w[2] = ((1.2 - t) * t + 0.8) * t;
            w[3] = ((1.0 / 3.0 * t - 0.2) * t - 0.13333334) * t;
        } else {
            double sum = 0.0;
            for (int i = 0; i < N; i++) {
                w[i] = keys_k(fabs(t - ((double)i + 1.0 - (double)(N / 2))));
                sum += w[i];
            }
            for (int i = 0; i < N; i++) w[i] /= sum;
        }
    } else {
        double sum = 0.0;
        for (int i = 0; i < N; i++) {
            double d = fabs(t - ((double)i + 1.0 - (double)(N / 2)));
            w[i] = ik == 4 ? spline36_k(d) : lanczos3_k(d);
            sum += w[i];
        }
        for (int i = 0; i < N; i++) w[i] /= sum;
    }
    for (int i = 0; i < N; i++) wf[i] = (float)w[i];
}

// One warped sample (align.rs:489-538) of channel `c` of a u16 frame, or of an f32 plane when c < 0.
template <int N>
__device__ __forceinline__ float warp_sample(const uint16_t *__restrict__ rgb, const float *__restrict__ plane, int c,
                                             int w, int h, int x0, int y0, const float *wx, const float *wy) {
    constexpr int start = 1 - N / 2;
    float acc = 0.0f;
#pragma unroll
    for (int j = 0; j < N; j++) {
        int yy = min(max(y0 + start + j, 0), h - 1);
        float r = 0.0f;
#pragma unroll
        for (int i = 0; i < N; i++) {
            int xx = min(max(x0 + start + i, 0), w - 1);
            size_t o = (size_t)yy * w + xx;
            float s = c < 0 ? plane[o] : (float)rgb[3 * o + c] * INV16;
            r += wx[i] * s;
        }
        acc += wy[j] * r;
    }
    return acc;
}

struct Gains {
    float g[3];
    int on[3];
};

// A frame's map and gains where the device derives them itself (the streaming
// path, no host round trip): the inverse of the searched Sim and the
// brightness gains from the device sums.
struct FrameDev {
    Inv9 inv;
    int affine, is_id;
    Gains g;
    float gains[3];
    double sim[8], evals;  // the searched Sim and its evaluation count, for the record
    unsigned err;          // the search's error flag as it ended
};

// align.rs:241-286: the Sim the search left in state[0..8), inverted (device
// sin/cos); the Sim, its evaluation count and the search's error flag are kept
// with the map for the frame's record, which may run after the next search
__global__ void k_frame_inv(const double *__restrict__ state, const unsigned *__restrict__ err, int w, int h,
                            FrameDev *__restrict__ fp) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s[8];
    for (int k = 0; k < 8; k++) s[k] = fp->sim[k] = state[k];
    fp->evals = state[8];
    fp->err = *err;
    sim_inverse_dev<true>(s, w, h, fp->inv.m);
    fp->affine = fp->inv.m[6] == 0.0 && fp->inv.m[7] == 0.0 && fp->inv.m[8] == 1.0;
    fp->is_id = s[0] == 0.0 && s[1] == 0.0 && s[2] == 1.0 && s[3] == 0.0 && s[4] == 1.0 && s[5] == 0.0 &&
                s[6] == 0.0 && s[7] == 0.0;
}

// brightness.rs:93-98 and :119-133 on the device: the gains from the sums the
// brightness kernel left (frame means in [0..3), frame 0's in [3..6), count
// in [6]); sums null: unit gains
__global__ void k_frame_gains(const double *__restrict__ sums, int enabled, FrameDev *__restrict__ fp) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    float g[3] = {1.0f, 1.0f, 1.0f};
    if (sums && enabled && sums[6] >= 64.0)
        for (int c = 0; c < 3; c++) {
            float v = sums[c] > 1e-9 ? (float)(sums[3 + c] / sums[c]) : 1.0f;
            g[c] = v < 0.25f ? 0.25f : (v > 4.0f ? 4.0f : v);
        }
    bool unity = true;
    for (int c = 0; c < 3; c++)
        if (!(fabsf(g[c] - 1.0f) < 1e-6f)) unity = false;
    for (int c = 0; c < 3; c++) {
        fp->g.g[c] = g[c];
        fp->g.on[c] = enabled && !unity && fabsf(g[c] - 1.0f) > 1e-6f;
        fp->gains[c] = g[c];
    }
}

// The final resample of a frame (align.rs:554-559) with the as-shot fallback
// outside the frame (stack.rs:669-678), the brightness gains (brightness.rs:
// 124-133), and optionally the warped registration luma (stack.rs:527) or,
// FY, the luma of the resampled frame itself (the focus measure's input,
// depth.rs:160-165, with the focus slice's own expression on the same values).
// The warp of one pixel: the three channels (and the luma when LUMA) summed in
// one pass over the taps, each sum in warp_sample's order, every load
// independent of whether the point is in the frame (clamped indices are always
// in range), so a thread's loads are all in flight together.
template <int N, bool LUMA, bool RGB, bool FY = false>
__device__ __forceinline__ void warp_px(const uint16_t *__restrict__ rgb, const float *__restrict__ yin, int w, int h,
                                        size_t o, bool valid, int x0, int y0, const float *wx, const float *wy,
                                        const Gains &g, float *__restrict__ planes, float *__restrict__ yout,
                                        float *__restrict__ fy = nullptr) {
    constexpr int start = 1 - N / 2;
    const size_t n = (size_t)w * h;
    float raw[3];
#pragma unroll
    for (int c = 0; c < 3; c++) raw[c] = RGB ? (float)rgb[3 * o + c] * INV16 : 0.0f;
    float acc[3] = {0.0f, 0.0f, 0.0f}, ay = 0.0f;
#pragma unroll
    for (int j = 0; j < N; j++) {
        const size_t row = (size_t)min(max(y0 + start + j, 0), h - 1) * w;
        float r[3] = {0.0f, 0.0f, 0.0f}, ry = 0.0f;
#pragma unroll
        for (int i = 0; i < N; i++) {
            const size_t q = row + min(max(x0 + start + i, 0), w - 1);
            if (RGB)
#pragma unroll
                for (int c = 0; c < 3; c++) r[c] += wx[i] * ((float)rgb[3 * q + c] * INV16);
            if (LUMA) ry += wx[i] * yin[q];
        }
#pragma unroll
        for (int c = 0; c < 3; c++) acc[c] += wy[j] * r[c];
        if (LUMA) ay += wy[j] * ry;
    }
    if (RGB) {
        float v[3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            v[c] = valid ? acc[c] : raw[c];
            if (g.on[c]) v[c] *= g.g[c];
            planes[c * n + o] = v[c];
        }
        if (FY) fy[o] = 0.299f * v[0] + 0.587f * v[1] + 0.114f * v[2];
    }
    if (LUMA) yout[o] = ay;
}

template <int N>
__global__ __launch_bounds__(256) void k_warp(const FrameRef src, const float *__restrict__ yin, int w, int h, Inv9 inv,
                                              int affine, int ik, Gains g, float *__restrict__ planes,
                                              float *__restrict__ yout, const FrameDev *__restrict__ fp,
                                              float *__restrict__ flum) {
    int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= w || y >= h) return;
    const uint16_t *__restrict__ rgb = src.get();
    if (fp) {
        inv = fp->inv;
        affine = fp->affine;
        g = fp->g;
    }
    size_t o = (size_t)y * w + x;
    double sx, sy;
    src_point(inv, affine, x, y, sx, sy);
    bool valid = sx >= 0.0 && sx <= (double)(w - 1) && sy >= 0.0 && sy <= (double)(h - 1);
    double fx = floor(sx), fy = floor(sy);
    int x0 = (int)fmax(fmin(fx, 1e9), -1e9), y0 = (int)fmax(fmin(fy, 1e9), -1e9);
    float wx[6], wy[6];
    interp_w<N>(ik, sx - fx, wx);
    interp_w<N>(ik, sy - fy, wy);
    if (yout && planes)
        warp_px<N, true, true>(rgb, yin, w, h, o, valid, x0, y0, wx, wy, g, planes, yout);
    else if (yout)
        warp_px<N, true, false>(rgb, yin, w, h, o, valid, x0, y0, wx, wy, g, planes, yout);
    else if (flum)
        warp_px<N, false, true, true>(rgb, yin, w, h, o, valid, x0, y0, wx, wy, g, planes, yout, flum);
    else
        warp_px<N, false, true>(rgb, yin, w, h, o, valid, x0, y0, wx, wy, g, planes, yout);
}

// The warped registration luma at one pixel: k_warp's luma path (warp_px's
// LUMA sum, the same taps in the same order), for a consumer that needs it
// only as an input sample.
template <int N>
__device__ __forceinline__ float warp_luma_at(const float *__restrict__ yin, int w, int h, const Inv9 &inv, int affine,
                                              int ik, int x, int y) {
    double sx, sy;
    src_point(inv, affine, x, y, sx, sy);
    const double fx = floor(sx), fy = floor(sy);
    const int x0 = (int)fmax(fmin(fx, 1e9), -1e9), y0 = (int)fmax(fmin(fy, 1e9), -1e9);
    float wx[6], wy[6];
    interp_w<N>(ik, sx - fx, wx);
    interp_w<N>(ik, sy - fy, wy);
    constexpr int start = 1 - N / 2;
    float ay = 0.0f;
#pragma unroll
    for (int j = 0; j < N; j++) {
        const size_t row = (size_t)min(max(y0 + start + j, 0), h - 1) * w;
        float ry = 0.0f;
#pragma unroll
        for (int i = 0; i < N; i++) ry += wx[i] * yin[row + min(max(x0 + start + i, 0), w - 1)];
        ay += wy[j] * ry;
    }
    return ay;
}

// Level 1 of the registration pyramid of the warped luma straight from the
// unwarped luma (align.rs:554-559 then :683-734): each tile's input window is
// the warped luma computed in LDS, so the full-resolution warped plane is
// never stored.  Only for pyramids whose level 0 nothing else reads (a search
// that stops short of full resolution).
template <int N>
__global__ __launch_bounds__(256) void k_warp_burt(const float *__restrict__ yin, int w, int h, int ik,
                                                   const FrameDev *__restrict__ fp, float *__restrict__ dst, int ow,
                                                   int oh) {
    __shared__ BurtLds<BT_TY> S;
    const Inv9 inv = fp->inv;
    const int affine = fp->affine;
    const int ox0 = blockIdx.x * BT_TX, oy0 = blockIdx.y * BT_TY;
    const bool inner = ox0 >= 1 && oy0 >= 1 && 2 * ox0 + 2 * BT_TX <= w - 1 && 2 * oy0 + 2 * BT_TY <= h - 1 &&
                       ox0 + BT_TX <= ow && oy0 + BT_TY <= oh;
    const auto warped = [&](int gx, int gy) { return warp_luma_at<N>(yin, w, h, inv, affine, ik, gx, gy); };
    if (inner)
        burt_body<BT_TY, false>(warped, w, h, dst, ow, oh, ox0, oy0, S);
    else
        burt_body<BT_TY, true>(warped, w, h, dst, ow, oh, ox0, oy0, S);
}

// The frame as decoded (io.rs:89), times the gains
__global__ void k_ingest(const uint16_t *__restrict__ rgb, size_t n, Gains g, float *__restrict__ planes) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        float v = (float)rgb[3 * i + c] * INV16;
        if (g.on[c]) v *= g.g[c];
        planes[c * n + i] = v;
    }
}

// brightness.rs:42-53: frame 0 on the step-4 grid
__global__ void k_grid(const uint16_t *__restrict__ rgb, int w, int h, int gw, int gh, float *__restrict__ grid) {
    int gx = blockIdx.x * blockDim.x + threadIdx.x, gy = blockIdx.y * blockDim.y + threadIdx.y;
    if (gx >= gw || gy >= gh) return;
    size_t o = (size_t)(gy * 4) * w + gx * 4;
    for (int c = 0; c < 3; c++) grid[(size_t)c * gw * gh + (size_t)gy * gw + gx] = (float)rgb[3 * o + c] * INV16;
}

// brightness.rs:67-90: sums over the step-4 grid of the warped frame (and of
// frame 0's grid) where the source point lies in the frame.  mode 0: the frame
// is warped by inv (the full warp's value, fallback included, is recomputed at
// each grid pixel); mode 1: the frame as decoded, every grid pixel counts.
template <int N>
__global__ __launch_bounds__(256) void k_bright_sums(const FrameRef src, int w, int h, Inv9 inv,
                                                     int affine, int ik, int mode, const float *__restrict__ grid,
                                                     int gw, int gh, double *__restrict__ part,
                                                     unsigned *__restrict__ ticket, double *__restrict__ out,
                                                     const FrameDev *__restrict__ fp) {
    if (fp) {
        inv = fp->inv;
        affine = fp->affine;
        mode = fp->is_id;
    }
    const uint16_t *__restrict__ rgb = src.get();
    double v[7] = {0, 0, 0, 0, 0, 0, 0};
    size_t ng = (size_t)gw * gh;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < ng; i += (size_t)gridDim.x * blockDim.x) {
        int gy = (int)(i / gw), gx = (int)(i - (size_t)gy * gw);
        int x = gx * 4, y = gy * 4;
        size_t o = (size_t)y * w + x;
        float val[3];
        if (mode == 1) {
            for (int c = 0; c < 3; c++) val[c] = (float)rgb[3 * o + c] * INV16;
        } else {
            // the membership test uses the general map (brightness.rs:74)
            double xf = (double)x, yf = (double)y;
            double d = inv.m[6] * xf + inv.m[7] * yf + inv.m[8];
            double mx = (inv.m[0] * xf + inv.m[1] * yf + inv.m[2]) / d;
            double my = (inv.m[3] * xf + inv.m[4] * yf + inv.m[5]) / d;
            if (!(mx >= 0.0 && mx <= (double)(w - 1) && my >= 0.0 && my <= (double)(h - 1))) continue;
            double sx, sy;
            src_point(inv, affine, x, y, sx, sy);
            bool valid = sx >= 0.0 && sx <= (double)(w - 1) && sy >= 0.0 && sy <= (double)(h - 1);
            double fx = floor(sx), fy = floor(sy);
            int x0 = (int)fmax(fmin(fx, 1e9), -1e9), y0 = (int)fmax(fmin(fy, 1e9), -1e9);
            float wx[6], wy[6];
            interp_w<N>(ik, sx - fx, wx);
            interp_w<N>(ik, sy - fy, wy);
            for (int c = 0; c < 3; c++)
                val[c] = valid ? warp_sample<N>(rgb, nullptr, c, w, h, x0, y0, wx, wy) : (float)rgb[3 * o + c] * INV16;
        }
        for (int c = 0; c < 3; c++) {
            v[c] += (double)val[c];
            if (grid) v[3 + c] += (double)grid[(size_t)c * ng + i];
        }
        v[6] += 1.0;
    }
    grid_sum<7>(v, part, ticket, out);
}

static int taps_n(uint32_t ik) { return ik <= 1 ? 2 : (ik <= 3 ? 4 : 6); }

// luma_only: the warped registration luma alone (no frame read, no planes);
// fy (with planes, without yout): the resampled frame's luma beside its planes
static void warp_launch(lsh_frames *f, const FrameRef &rgb, const Inv9 &inv, int affine, const Gains &g,
                        float *yout, const FrameDev *fp = nullptr, const float *yin_ = nullptr,
                        float *planes_out = nullptr, bool luma_only = false, float *fy = nullptr) {
    dim3 gr(cdiv(f->w, 64), cdiv(f->h, 4)), b(64, 4);
    const float *yin = yin_ ? yin_ : f->yt.p;
    float *planes = luma_only ? nullptr : (planes_out ? planes_out : f->planes.p);
    int ik = (int)f->interp, w = (int)f->w, h = (int)f->h;
    // u16 frame read, 3 f32 planes written; the luma: luma read, warped luma written
    if (yout || !planes) fy = nullptr;
    const double by = (double)f->w * f->h * ((planes ? 18.0 : 0.0) + (yout ? 8.0 : 0.0) + (fy ? 4.0 : 0.0));
    const char *name = planes ? "warp" : "warp_luma";
    switch (taps_n(f->interp)) {
    case 2: launch_b(name, by, k_warp<2>, gr, b, 0, rgb, yin, w, h, inv, affine, ik, g, planes, yout, fp, fy); break;
    case 4: launch_b(name, by, k_warp<4>, gr, b, 0, rgb, yin, w, h, inv, affine, ik, g, planes, yout, fp, fy); break;
    default: launch_b(name, by, k_warp<6>, gr, b, 0, rgb, yin, w, h, inv, affine, ik, g, planes, yout, fp, fy); break;
    }
}

// host7 null: the sums stay in red_out for a device consumer (no round trip)
static void sums_launch(lsh_frames *f, const FrameRef &rgb, const Inv9 &inv, int affine, int mode, double *host7,
                        const FrameDev *fp = nullptr) {
    size_t ng = f->gw * f->gh;
    unsigned grid = std::max(1u, std::min<unsigned>(cdiv(ng, 256), 1024));
    f->red.ensure((size_t)grid * 7);
    const float *gp = f->have_grid ? f->grid.p : nullptr;
    int ik = (int)f->interp, w = (int)f->w, h = (int)f->h, gw = (int)f->gw, gh = (int)f->gh;
    // per step-4 grid point: the kernel's N x N taps of the three u16 channels, frame 0's three f32 samples
    const double nt = (double)taps_n(f->interp), by = (double)ng * (6.0 * nt * nt + 12.0);
    switch (taps_n(f->interp)) {
    case 2: launch_b("bright_sums", by, k_bright_sums<2>, dim3(grid), dim3(256), 0, rgb, w, h, inv, affine, ik, mode, gp, gw, gh, f->red.p, f->ticket.p, f->red_out.p, fp); break;
    case 4: launch_b("bright_sums", by, k_bright_sums<4>, dim3(grid), dim3(256), 0, rgb, w, h, inv, affine, ik, mode, gp, gw, gh, f->red.p, f->ticket.p, f->red_out.p, fp); break;
    default: launch_b("bright_sums", by, k_bright_sums<6>, dim3(grid), dim3(256), 0, rgb, w, h, inv, affine, ik, mode, gp, gw, gh, f->red.p, f->ticket.p, f->red_out.p, fp); break;
    }
    if (!host7) return;
    double *pin = (double *)pinned_scratch(64);
    LSH_CHECK(hipMemcpyAsync(pin, f->red_out.p, 7 * sizeof(double), hipMemcpyDeviceToHost, stream()));
    sync();
    std::memcpy(host7, pin, 7 * sizeof(double));
}

// brightness.rs:93-98
static void ratio(const double *ref, const double *frm, uint64_t n, float *g) {
    for (int c = 0; c < 3; c++) {
        if (n < 64) {
            g[c] = 1.0f;
            continue;
        }
        float v = frm[c] > 1e-9 ? (float)(ref[c] / frm[c]) : 1.0f;
        g[c] = rclampf(v, 0.25f, 4.0f);
    }
}

// brightness.rs:119-133: which channels apply multiplies
static Gains gains_of(const float *g, bool enabled) {
    Gains G{};
    bool unity = true;
    for (int c = 0; c < 3; c++) {
        G.g[c] = g[c];
        if (!(std::fabs(g[c] - 1.0f) < 1e-6f)) unity = false;
    }
    for (int c = 0; c < 3; c++) G.on[c] = enabled && !unity && std::fabs(g[c] - 1.0f) > 1e-6f;
    return G;
}

static Inv9 inv_for(const Sim &s, size_t w, size_t h, int &affine) {
    double m[9];
    Inv9 v;
    sim_matrix(s, w, h, m);
    inverse9(m, v.m);
    affine = v.m[6] == 0.0 && v.m[7] == 0.0 && v.m[8] == 1.0;
    return v;
}

// the registration pyramid above level 0 (align.rs:741-754)
static void burt_level(const float *src, size_t w, size_t h, float *dst, size_t ow, size_t oh) {
    const dim3 grid(cdiv(ow, BT_TX), cdiv(oh, BT_TY));
    launch_b("burt_reduce", 4.0 * (w * h + ow * oh), k_burt<BT_TY>, grid, dim3(256), 0, src, (int)w, (int)h, dst, (int)ow,
             (int)oh);
}

static void build_reg_pyramid(lsh_frames *f, float *base) {
    for (size_t l = 1; l < f->lw.size(); l++)
        burt_level(base + f->loff[l - 1], f->lw[l - 1], f->lh[l - 1], base + f->loff[l], f->lw[l], f->lh[l]);
}

// align.rs:220: the search box's half-width around the guess, per parameter
static const double SPAN[8] = {0.10, 0.10, 0.10, 0.087266462599716474, 0.05, 0.05, 0.05, 0.05};

// multiscale_align (align.rs:896-942) through the cooperative search kernel:
// the target pyramid in yt, the reference in yr, the guess in, the Sim out.
// The search's constants (levels, steps, tolerances, free parameters) on the
// device, and its buffers; the search box follows the guess in the kernel.
static void search_const(lsh_frames *f, const uint8_t *free8, size_t coarsen) {
    size_t n = f->lw.size();  // levels 0..n-1, level 0 the full-resolution plane
    LSH_REQUIRE(n <= NM_MAXL, "registration pyramid deeper than the search supports");
    NmConst K;
    std::memset(&K, 0, sizeof(K));
    K.nlev = (int)n;
    K.finest = (int)std::min(coarsen, n - 1);
    K.solo_max = 4096;
    if (const char *e = std::getenv("LSH_SEARCH_SOLO_MAX")) K.solo_max = std::atoi(e);
    // speculation pays where a pass is dominated by the grid's synchronisation
    K.spec_max = 1000000;
    if (const char *e = std::getenv("LSH_SEARCH_SPEC_MAX")) K.spec_max = std::atoi(e);
    for (size_t l = 0; l < n; l++) {
        K.ref[l] = f->yr.p + f->loff[l];
        K.tgt[l] = f->yt.p + f->loff[l];
        K.lw[l] = (int)f->lw[l];
        K.lh[l] = (int)f->lh[l];
    }
    for (int k = 0; k < 8; k++) {
        K.span[k] = SPAN[k];
        if (free8[k]) K.free_idx[K.nfree++] = k;
    }
    {
        // the wait bounds in ticks of the device wall clock: a wait for the
        // other participants ends with an error after 4 s (a pass takes well
        // under a millisecond), and the set of participants closes 20 ms after
        // the last block joined (any other kernel's blocks have long left the
        // CUs by then, so a normal launch always gets its whole grid)
        int dev = 0, khz = 0;
        LSH_CHECK(hipGetDevice(&dev));
        LSH_CHECK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
        LSH_REQUIRE(khz > 0, "the device reports no wall clock");
        K.wait_ticks = (uint64_t)khz * 4000u;
        K.join_ticks = (uint64_t)khz * 20u;
    }
    if (const char *e = std::getenv("LSH_SEARCH_MAX_BLOCKS")) K.max_part = std::max(0, std::atoi(e));
    {
        const char *e = std::getenv("LSH_SEARCH_PHASES");
        K.phases = prof().on || (e && std::atoi(e) != 0);
    }
    double W = (double)f->w, H = (double)f->h, r = std::max(W, H) / 2.0;
    const double gain[8] = {W, H, r, r, H / 2.0, H / 2.0, W / 4.0, H / 4.0};  // align.rs:766-770
    for (size_t l = 0; l < n; l++) {
        double px = (double)((uint64_t)1 << l);
        for (int k = 0; k < K.nfree; k++) {
            K.step[l][k] = 1.0 * px / gain[K.free_idx[k]];  // align.rs:784-790
            K.tol[l][k] = 0.1 * px / gain[K.free_idx[k]];
        }
    }
    if (!f->nm_const.p) {
        f->nm_const.alloc(sizeof(NmConst));
        f->nm_state.alloc(NM_STATE);
        f->nm_flags.alloc(NM_FL_WORDS);
        int dev = 0, cus = 0, per_cu = 0;
        LSH_CHECK(hipGetDevice(&dev));
        LSH_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        LSH_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k_nm_search, NM_THREADS, 0));
        LSH_REQUIRE(per_cu > 0, "the search kernel cannot be resident");
        int want = 1;  // blocks per CU: more hide gather latency, fewer arrive faster
        if (const char *e = std::getenv("LSH_SEARCH_BLOCKS_PER_CU")) want = std::max(1, std::atoi(e));
        f->nm_bpc = (unsigned)std::min(per_cu, want);
        f->nm_grid = f->nm_bpc * (unsigned)cus;
        f->nm_grid = std::min(f->nm_grid, (unsigned)NM_MAXG);  // an arrival flag each
        f->nm_grid -= f->nm_grid % NM_GRID_STEP;
        LSH_REQUIRE(f->nm_grid >= NM_GRID_STEP, "the search needs at least 16 resident blocks");
        f->nm_part.alloc((size_t)2 * 3 * NM_MAXP * f->nm_grid + 2 * NM_MAXP);
        f->nm_digest.alloc(f->nm_grid);
    }
    K.digest = (unsigned long long *)f->nm_digest.p;
    if (f->trace_cap) {
        f->nm_trace.ensure(f->trace_cap);
        K.trace = f->nm_trace.p;
        K.trace_cap = (int)f->trace_cap;
    }
    // upload only what changed (a stack's frames share one set)
    if (f->nm_host.size() == sizeof(NmConst) && std::memcmp(f->nm_host.data(), &K, sizeof(NmConst)) == 0) return;
    f->nm_host.assign((const uint8_t *)&K, (const uint8_t *)&K + sizeof(NmConst));
    uint8_t *pin = (uint8_t *)pinned_scratch(sizeof(NmConst));
    std::memcpy(pin, &K, sizeof(NmConst));
    LSH_CHECK(hipMemcpyAsync(f->nm_const.p, pin, sizeof(NmConst), hipMemcpyHostToDevice, stream()));
    sync();  // the pinned block is reused
}

// One search from the guess in state[0..8) (left there by the caller or by the
// previous search), its result written back there.  plain: an ordinary launch
// of the same grid (inside a captured graph), which every block of fits at
// once on an otherwise idle device; the kernel's bounded waits end it with an
// error rather than a hang if that ever fails.
static void search_launch(lsh_frames *f, bool plain = false, hipEvent_t started = nullptr) {
    unsigned *ctr = f->nm_flags.p, *err = f->nm_flags.p + NM_FL_ERR;
    // a stream's searches (plain) keep a raised error flag: the frames after a
    // failed one leave at once
    launch_b("search_reset", 4.0 * NM_FL_WORDS, k_nm_reset, dim3(1), dim3(256), 0, f->nm_flags.p, plain ? 1 : 0);
    if (started) LSH_CHECK(hipEventRecord(started, stream()));
    const NmConst *kp = (const NmConst *)f->nm_const.p;
    double *state = f->nm_state.p, *part = f->nm_part.p;
    unsigned grid = f->nm_grid;
    if (const unsigned c = stream_cus()) {
        // a stream restricted to c CUs holds only that many blocks at once
        grid = std::min(grid, f->nm_bpc * c);
        grid -= grid % NM_GRID_STEP;
        LSH_REQUIRE(grid >= NM_GRID_STEP, "the search needs at least 16 resident blocks");
    }
    if (plain) {
        launch("nm_search", k_nm_search, dim3(grid), dim3(NM_THREADS), 0, kp, state, part, ctr, err);
        return;
    }
    void *args[] = {(void *)&kp, (void *)&state, (void *)&part, (void *)&ctr, (void *)&err};
    launch_coop("nm_search", 0.0, (const void *)k_nm_search, dim3(grid), dim3(NM_THREADS), args, 0);
}

// The search's record (its state words, as the kernel left them) into the
// handle's statistics: evaluations, ticks and passes per level, participants
static void search_state_in(lsh_frames *f, const double *out) {
    f->search_part = (unsigned)out[NM_ST_PART];
    f->trace_n = (size_t)out[NM_ST_TRACE];
    f->evals = (uint64_t)out[8];
    for (size_t l = 0; l < NM_MAXL; l++) {
        f->lev_evals[l] = (uint64_t)out[9 + l];
        f->lev_ticks[l] = (uint64_t)out[9 + NM_MAXL + l];
    }
    for (int q = 0; q < 4; q++) f->phase_ticks[q] = (uint64_t)out[9 + 2 * NM_MAXL + q];
    for (int l = 0; l < NM_MAXL; l++)
        for (int q = 0; q < NM_NPH; q++) f->lev_phase_ticks[l][q] = (uint64_t)out[NM_ST_LPH + NM_NPH * l + q];
    for (size_t l = 0; l < NM_MAXL; l++) f->lev_passes[l] = (uint64_t)out[13 + 2 * NM_MAXL + l];
}

static Sim multiscale_align(lsh_frames *f, const Sim &init, const uint8_t *free8, size_t coarsen) {
    const size_t n = f->lw.size();
    search_const(f, free8, coarsen);
    uint8_t *pin = (uint8_t *)pinned_scratch((24 + NM_STATE) * sizeof(double));
    double *st = (double *)pin;
    std::memcpy(st, init.v, 8 * sizeof(double));
    LSH_CHECK(hipMemcpyAsync(f->nm_state.p, st, 8 * sizeof(double), hipMemcpyHostToDevice, stream()));
    search_launch(f);
    unsigned *err = f->nm_flags.p + NM_FL_ERR;
    double *out = st + 16;
    LSH_CHECK(hipMemcpyAsync(out, f->nm_state.p, NM_STATE * sizeof(double), hipMemcpyDeviceToHost, stream()));
    unsigned *flags = (unsigned *)(out + NM_STATE);
    LSH_CHECK(hipMemcpyAsync(flags, err, sizeof(unsigned), hipMemcpyDeviceToHost, stream()));
    sync();
    LSH_REQUIRE(flags[0] == 0, "search: a wait for the other blocks timed out");
    search_state_in(f, out);
    // algorithmic bytes: each counted cost evaluation (the reference's own,
    // align.rs:581-635) reads the level's reference and target once; the
    // speculative costs a branch leaves unused are this kernel's overhead
    double bytes = 0.0;
    for (size_t l = 0; l < n; l++) bytes += out[9 + l] * 8.0 * (double)(f->lw[l] * f->lh[l]);
    prof_set_last_bytes(bytes);
    Sim s;
    std::memcpy(s.v, out, 8 * sizeof(double));
    return s;
}

// the per-frame record of a stream: Sim, gains, evaluations at the frame
// index, a sticky copy of the search's error flag, then the index advances
__global__ void k_frame_record(const FrameDev *__restrict__ fp, unsigned *__restrict__ cursor,
                               double *__restrict__ sims, float *__restrict__ gains, double *__restrict__ evals) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const unsigned i = cursor[0];
    for (int k = 0; k < 8; k++) sims[(size_t)i * 8 + k] = fp->sim[k];
    for (int c = 0; c < 3; c++) gains[(size_t)i * 3 + c] = fp->gains[c];
    evals[i] = fp->evals;
    cursor[1] |= fp->err;
    cursor[0] = i + 1;
}

__global__ void k_cursor_set(unsigned *cursor, unsigned v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        cursor[0] = v;
        cursor[1] = 0;
    }
}

__global__ void k_state_set(double *state, Sim s) {
    if (threadIdx.x < 8 && blockIdx.x == 0) state[threadIdx.x] = s.v[threadIdx.x];
}

void frames_stream_setup(lsh_frames *f, const uint8_t *free8, uint32_t coarsen, size_t nframes) {
    LSH_REQUIRE(f->have_ref && f->processed > 0, "stream: the first frame goes through lsh_frames_process");
    search_const(f, free8, coarsen);
    f->coarsen = coarsen;
    f->fdev.ensure(2 * sizeof(FrameDev));  // one per parity of the frame
    f->cursor.ensure(2);
    f->rec_sims.ensure(nframes * 8);
    f->rec_gains.ensure(nframes * 3);
    f->rec_evals.ensure(nframes);
    f->red.ensure((size_t)std::max(1u, std::min<unsigned>(cdiv(f->gw * f->gh, 256), 1024)) * 7);
    launch_b("search_reset", 4.0 * NM_FL_WORDS, k_nm_reset, dim3(1), dim3(256), 0, f->nm_flags.p, 0);  // a fresh error flag
    launch("stream_cursor", k_cursor_set, dim3(1), dim3(64), 0, f->cursor.p, (unsigned)f->processed);
    launch("stream_state", k_state_set, dim3(1), dim3(64), 0, f->nm_state.p, Sim());  // frame 0 sits at the identity
}

void frames_stream_slots(lsh_frames *f) {
    const size_t tot = f->yt.n;
    for (auto &b : f->yts) b.ensure(tot);
    f->planes_b.ensure(3 * f->w * f->h);
}

void frames_target(lsh_frames *f, const FrameRef &src, int slot) {
    const size_t n = f->w * f->h;
    float *dst = f->yts[slot].p;
    launch_b("luma", 10.0 * n, k_luma_u16, dim3(cdiv(n, 1024)), dim3(256), 0, src, n, dst);
    build_reg_pyramid(f, dst);
}

void frames_chain(lsh_frames *f, int slot, bool plain_search, hipEvent_t started) {
    const size_t w = f->w, h = f->h;
    const float *tgt = f->yts[slot].p;
    // the levels the search reads (the finest searched and coarser) into its target buffer
    const size_t fin = f->loff[std::min<size_t>(f->coarsen, f->lw.size() - 1)];
    copy_mem(f->yt.p + fin, LSH_DEVICE, tgt + fin, LSH_DEVICE, (f->yt.n - fin) * sizeof(float));
    FrameDev *fp = (FrameDev *)f->fdev.p + slot;
    search_launch(f, plain_search, started);
    launch("frame_inv", k_frame_inv, dim3(1), dim3(64), 0, (const double *)f->nm_state.p,
           (const unsigned *)(f->nm_flags.p + NM_FL_ERR), (int)w, (int)h, fp);
    // the warped registration luma from the target's full-resolution luma in
    // its slot, and its pyramid; when the search stops short of level 0, the
    // warp is folded into the first REDUCE and level 0 is never stored
    if (f->coarsen >= 1 && f->lw.size() >= 2) {
        const dim3 grid(cdiv(f->lw[1], BT_TX), cdiv(f->lh[1], BT_TY));
        const int wi = (int)w, hi = (int)h, ik = (int)f->interp, ow = (int)f->lw[1], oh = (int)f->lh[1];
        float *l1 = f->yr.p + f->loff[1];
        // the luma taps of the warped samples in the window (about 1.2 per level-0 pixel), level 1 written
        const double by = 4.0 * (double)taps_n(f->interp) * taps_n(f->interp) * 1.2 * w * h + 4.0 * ow * oh;
        switch (taps_n(f->interp)) {
        case 2: launch_b("warp_burt", by, k_warp_burt<2>, grid, dim3(256), 0, tgt, wi, hi, ik, (const FrameDev *)fp, l1, ow, oh); break;
        case 4: launch_b("warp_burt", by, k_warp_burt<4>, grid, dim3(256), 0, tgt, wi, hi, ik, (const FrameDev *)fp, l1, ow, oh); break;
        default: launch_b("warp_burt", by, k_warp_burt<6>, grid, dim3(256), 0, tgt, wi, hi, ik, (const FrameDev *)fp, l1, ow, oh); break;
        }
        for (size_t l = 2; l < f->lw.size(); l++)
            burt_level(f->yr.p + f->loff[l - 1], f->lw[l - 1], f->lh[l - 1], f->yr.p + f->loff[l], f->lw[l], f->lh[l]);
    } else {
        const Inv9 unused{};
        const Gains ug{};
        warp_launch(f, FrameRef(), unused, 1, ug, f->yr.p, fp, tgt, nullptr, true);
        build_reg_pyramid(f, f->yr.p);
    }
    f->processed++;
}

void frames_side_warp(lsh_frames *f, const FrameRef &src, int slot, float *planes_out, bool brightness, float *fy_out) {
    FrameDev *fp = (FrameDev *)f->fdev.p + slot;
    const Inv9 unused{};
    const Gains ug{};
    if (brightness && f->have_grid) {
        sums_launch(f, src, unused, 1, 0, nullptr, fp);
        launch("frame_gains", k_frame_gains, dim3(1), dim3(64), 0, (const double *)f->red_out.p, 1, fp);
    } else {
        launch("frame_gains", k_frame_gains, dim3(1), dim3(64), 0, (const double *)nullptr, 0, fp);
    }
    warp_launch(f, src, unused, 1, ug, nullptr, fp, nullptr, planes_out, false, fy_out);
    launch("frame_record", k_frame_record, dim3(1), dim3(64), 0, (const FrameDev *)fp, f->cursor.p, f->rec_sims.p,
           f->rec_gains.p, f->rec_evals.p);
}

void frames_stream_results(lsh_frames *f, size_t n, double *sims, float *gains, uint64_t *evals) {
    std::vector<unsigned> cur(2);
    copy_mem(cur.data(), LSH_HOST, f->cursor.p, LSH_DEVICE, 2 * sizeof(unsigned));
    LSH_REQUIRE(cur[1] == 0, "search: a cost broadcast timed out");
    if (sims) copy_mem(sims, LSH_HOST, f->rec_sims.p, LSH_DEVICE, n * 8 * sizeof(double));
    if (gains) copy_mem(gains, LSH_HOST, f->rec_gains.p, LSH_DEVICE, n * 3 * sizeof(float));
    if (evals) {
        std::vector<double> e(n);
        copy_mem(e.data(), LSH_HOST, f->rec_evals.p, LSH_DEVICE, n * sizeof(double));
        for (size_t i = 0; i < n; i++) evals[i] = (uint64_t)e[i];
    }
    // the last search's statistics (lsh_frames_search_stats)
    std::vector<double> st(NM_STATE);
    copy_mem(st.data(), LSH_HOST, f->nm_state.p, LSH_DEVICE, NM_STATE * sizeof(double));
    search_state_in(f, st.data());
}

static const uint16_t *frame_source(lsh_frames *f, const uint16_t *rgb, lsh_mem mem) {
    if (mem == LSH_DEVICE) return rgb;
    f->staging.ensure(3 * f->w * f->h);
    copy_mem(f->staging.p, LSH_DEVICE, rgb, LSH_HOST, 3 * f->w * f->h * sizeof(uint16_t));
    return f->staging.p;
}

static void slice_out_(lsh_frames *f, const lsh_depth_params *measure, float *slice_out, lsh_mem slice_mem) {
    if (!measure) return;
    LSH_REQUIRE(slice_out, "frames: a focus slice is asked for without an output");
    size_t k = (size_t)1 << measure->scale;
    size_t dw = (f->w + k - 1) / k, dh = (f->h + k - 1) / k;
    if (slice_mem == LSH_DEVICE) {
        focus_slice(f, f->planes.p, *measure, slice_out);
    } else {
        DBuf<float> tmp(dw * dh);
        focus_slice(f, f->planes.p, *measure, tmp.p);
        copy_mem(slice_out, LSH_HOST, tmp.p, LSH_DEVICE, dw * dh * sizeof(float));
    }
}

}  // namespace lsh

using namespace lsh;

extern "C" {

char *lsh_frames_new(uint32_t w, uint32_t h, uint32_t interp, lsh_frames **out) {
    LSH_API_BEGIN
    LSH_REQUIRE(out, "lsh_frames_new: null output");
    *out = nullptr;
    LSH_REQUIRE(w > 0 && h > 0, "lsh_frames_new: empty frame");
    LSH_REQUIRE(interp <= LSH_LANCZOS3, "lsh_frames_new: unknown interpolation kernel");
    auto f = new lsh_frames();
    try {
        f->w = w;
        f->h = h;
        f->interp = interp;
        size_t n = (size_t)w * h;
        f->planes.alloc(3 * n);
        // registration levels: the plane itself, then halved while h > 64 && w > 8
        size_t cw = w, ch = h, off = 0;
        f->lw.push_back(cw);
        f->lh.push_back(ch);
        f->loff.push_back(0);
        off = cw * ch;
        while (ch > 64 && cw > 8) {
            cw = (cw + 1) / 2;
            ch = (ch + 1) / 2;
            f->lw.push_back(cw);
            f->lh.push_back(ch);
            f->loff.push_back(off);
            off += cw * ch;
        }
        f->yt.alloc(off);
        f->yr.alloc(off);
        f->gw = (w + 3) / 4;
        f->gh = (h + 3) / 4;
        f->grid.alloc(3 * f->gw * f->gh);
        f->red.alloc(1024 * 7);
        f->red_out.alloc(8);
        f->ticket.alloc(1);
        f->ticket.zero();
    } catch (...) {
        delete f;
        throw;
    }
    *out = f;
    LSH_API_END
}

char *lsh_frames_process(lsh_frames *f, const uint16_t *rgb, lsh_mem rgb_mem, const double *guess,
                         const uint8_t *free8, uint32_t coarsen, const double *known_sim, const float *known_gains,
                         uint32_t brightness, const lsh_depth_params *measure, double *sim_out, float *gains_out,
                         float *slice_out, lsh_mem slice_mem) {
    LSH_API_BEGIN
    LSH_REQUIRE(f && rgb && free8, "lsh_frames_process: null argument");
    LSH_REQUIRE((known_sim == nullptr) == (known_gains == nullptr), "lsh_frames_process: known needs both sim and gains");
    const size_t w = f->w, h = f->h, n = w * h;
    const uint16_t *src = frame_source(f, rgb, rgb_mem);
    bool first = !f->have_ref && !known_sim;
    bool searching = !known_sim && !first;
    Sim sim;
    float g[3] = {1.0f, 1.0f, 1.0f};
    f->evals = 0;
    if (known_sim) {
        std::memcpy(sim.v, known_sim, sizeof(sim.v));
        std::memcpy(g, known_gains, sizeof(g));
    } else if (searching) {
        Sim init;
        if (guess) std::memcpy(init.v, guess, sizeof(init.v));
        launch_b("luma", 10.0 * n, k_luma_u16, dim3(cdiv(n, 1024)), dim3(256), 0, src, n, f->yt.p);
        build_reg_pyramid(f, f->yt.p);
        sim = multiscale_align(f, init, free8, coarsen);
    }
    int affine = 1;
    Inv9 inv = inv_for(sim, w, h, affine);
    // brightness gains against frame 0's grid (stack.rs:681-690)
    if (!known_sim && brightness) {
        if (first) {
            dim3 gg(cdiv(f->gw, 64), cdiv(f->gh, 4));
            launch_b("bright_grid", 18.0 * f->gw * f->gh, k_grid, gg, dim3(64, 4), 0, src, (int)w, (int)h, (int)f->gw, (int)f->gh, f->grid.p);
            f->have_grid = true;
        } else if (f->have_grid) {
            double s7[7];
            sums_launch(f, src, inv, affine, sim.is_id() ? 1 : 0, s7);
            ratio(s7 + 3, s7, (uint64_t)s7[6], g);
        }
    }
    Gains G = gains_of(g, brightness || known_sim);
    // the next frame's reference: this frame's luma, warped (stack.rs:527)
    if (!known_sim) {
        if (!searching) {
            launch_b("luma", 10.0 * n, k_luma_u16, dim3(cdiv(n, 1024)), dim3(256), 0, src, n, f->yt.p);
        }
        if (sim.is_id()) {
            std::swap(f->yt, f->yr);
            launch_b("ingest", 18.0 * n, k_ingest, dim3(cdiv(n, 256)), dim3(256), 0, src, n, G, f->planes.p);
        } else {
            warp_launch(f, src, inv, affine, G, f->yr.p);
        }
        build_reg_pyramid(f, f->yr.p);
        f->have_ref = true;
    } else if (sim.is_id()) {
        launch_b("ingest", 18.0 * n, k_ingest, dim3(cdiv(n, 256)), dim3(256), 0, src, n, G, f->planes.p);
    } else {
        warp_launch(f, src, inv, affine, G, nullptr);
    }
    slice_out_(f, measure, slice_out, slice_mem);
    f->processed++;
    if (sim_out) std::memcpy(sim_out, sim.v, sizeof(sim.v));
    if (gains_out) std::memcpy(gains_out, g, sizeof(g));
    LSH_API_END
}

char *lsh_frames_ingest(lsh_frames *f, const uint16_t *rgb, lsh_mem rgb_mem, uint32_t brightness,
                        const lsh_depth_params *measure, float *gains_out, float *slice_out, lsh_mem slice_mem) {
    LSH_API_BEGIN
    LSH_REQUIRE(f && rgb, "lsh_frames_ingest: null argument");
    const size_t n = f->w * f->h;
    const uint16_t *src = frame_source(f, rgb, rgb_mem);
    float g[3] = {1.0f, 1.0f, 1.0f};
    if (brightness) {
        bool had = f->have_grid;
        f->have_grid = false;  // means only: no reference grid in the sums
        double s7[7];
        Inv9 id{};
        sums_launch(f, src, id, 1, 1, s7);
        f->have_grid = had;
        double m[3];
        for (int c = 0; c < 3; c++) m[c] = s7[6] == 0.0 ? 0.0 : s7[c] / s7[6];  // brightness.rs:107-110
        if (!f->have_means) {
            std::memcpy(f->ref_means, m, sizeof(m));
            f->have_means = true;
        } else {
            ratio(f->ref_means, m, UINT64_MAX, g);  // brightness.rs:114-117
        }
    }
    Gains G = gains_of(g, brightness != 0);
    launch_b("ingest", 18.0 * n, k_ingest, dim3(cdiv(n, 256)), dim3(256), 0, src, n, G, f->planes.p);
    slice_out_(f, measure, slice_out, slice_mem);
    f->processed++;
    if (gains_out) std::memcpy(gains_out, g, sizeof(g));
    LSH_API_END
}

char *lsh_frames_planes(lsh_frames *f, const float **planes_dev) {
    LSH_API_BEGIN
    LSH_REQUIRE(f && planes_dev, "lsh_frames_planes: null argument");
    *planes_dev = f->planes.p;
    LSH_API_END
}

char *lsh_frames_download(lsh_frames *f, float *img3, lsh_mem mem) {
    LSH_API_BEGIN
    LSH_REQUIRE(f && img3, "lsh_frames_download: null argument");
    copy_mem(img3, mem, f->planes.p, LSH_DEVICE, 3 * f->w * f->h * sizeof(float));
    sync();
    LSH_API_END
}

void lsh_frames_free(lsh_frames *f) {
    if (!f) return;
    try {
        sync();
    } catch (...) {
    }
    delete f;
}

uint64_t lsh_frames_last_evals(lsh_frames *f) { return f ? f->evals : 0; }

char *lsh_frames_search_stats(lsh_frames *f, uint64_t *evals, uint64_t *ticks, uint32_t cap) {
    LSH_API_BEGIN
    LSH_REQUIRE(f, "lsh_frames_search_stats: null handle");
    for (uint32_t l = 0; l < cap && l < NM_MAXL; l++) {
        if (evals) evals[l] = f->lev_evals[l];
        if (ticks) ticks[l] = f->lev_ticks[l];
    }
    // past the levels: block 0's ticks per phase (matrix, pixels, combine,
    // steering), and the collective passes per level
    for (uint32_t q = 0; q < 4 && NM_MAXL + q < cap; q++)
        if (ticks) ticks[NM_MAXL + q] = f->phase_ticks[q];
    for (uint32_t l = 0; l < NM_MAXL && NM_MAXL + l < cap; l++)
        if (evals) evals[NM_MAXL + l] = f->lev_passes[l];
    // then per level the phases' ticks
    for (uint32_t i = 0; i < NM_NPH * NM_MAXL && NM_MAXL + 4 + i < cap; i++)
        if (ticks) ticks[NM_MAXL + 4 + i] = f->lev_phase_ticks[i / NM_NPH][i % NM_NPH];
    LSH_API_END
}

char *lsh_frames_search_trace(lsh_frames *f, uint32_t cap) {
    LSH_API_BEGIN
    LSH_REQUIRE(f, "lsh_frames_search_trace: null handle");
    f->trace_cap = cap;
    f->trace_n = 0;
    if (!cap) f->nm_trace.release();
    LSH_API_END
}

char *lsh_frames_search_costs(lsh_frames *f, double *costs, uint32_t cap, uint32_t *n) {
    LSH_API_BEGIN
    LSH_REQUIRE(f && n, "lsh_frames_search_costs: null argument");
    LSH_REQUIRE(f->trace_cap, "lsh_frames_search_costs: tracing is off (lsh_frames_search_trace)");
    LSH_REQUIRE(f->trace_n <= f->trace_cap, "lsh_frames_search_costs: the search counted more costs than traced");
    *n = (uint32_t)f->trace_n;
    const size_t m = std::min<size_t>(cap, f->trace_n);
    if (costs && m) copy_mem(costs, LSH_HOST, f->nm_trace.p, LSH_DEVICE, m * sizeof(double));
    LSH_API_END
}

char *lsh_frames_search_blocks(lsh_frames *f, uint64_t *digests, uint32_t cap, uint32_t *n) {
    LSH_API_BEGIN
    LSH_REQUIRE(f && n, "lsh_frames_search_blocks: null argument");
    *n = f->search_part;
    const size_t m = std::min<size_t>(cap, f->search_part);
    if (digests && m) copy_mem(digests, LSH_HOST, f->nm_digest.p, LSH_DEVICE, m * sizeof(uint64_t));
    LSH_API_END
}

}  // extern "C"

// ---------------------------------------------------------- the output side
// What happens to the result after the path: the crop to the area every frame
// covers (align.rs:344-403 common_area, Img3::crop), the restretch back to
// the frame's size (stack.rs:951-974 through prep.rs:46-71 resize), and the
// 16-bit quantisation of the writers (io.rs:283-288).
namespace lsh {

// prep.rs:63-71 resize_plane: the warp of align.rs:489-559 (edge-clamped taps)
// onto the ow x oh grid with the Sim that scales the source over the output
// about its center, pixel centers onto pixel centers; np planes
template <int N>
__global__ __launch_bounds__(256) void k_resize(const float *__restrict__ src, int pw, int ph, Inv9 inv, int affine,
                                                int ik, int np, float *__restrict__ dst, int ow, int oh) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= ow || y >= oh) return;
    double sx, sy;
    src_point(inv, affine, x, y, sx, sy);
    const double fx = floor(sx), fy = floor(sy);
    const int x0 = (int)fmax(fmin(fx, 1e9), -1e9), y0 = (int)fmax(fmin(fy, 1e9), -1e9);
    float wx[6], wy[6];
    interp_w<N>(ik, sx - fx, wx);
    interp_w<N>(ik, sy - fy, wy);
    const size_t pn = (size_t)pw * ph, on = (size_t)ow * oh, o = (size_t)y * ow + x;
    for (int c = 0; c < np; c++) dst[c * on + o] = warp_sample<N>(nullptr, src + c * pn, -1, pw, ph, x0, y0, wx, wy);
}

// Img3::crop of np planes: the rect (x, y, w, h) of each
__global__ void k_crop(const float *__restrict__ src, int w, int h, int np, int rx, int ry, int rw, int rh,
                       float *__restrict__ dst) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= rw || y >= rh) return;
    const size_t n = (size_t)w * h, rn = (size_t)rw * rh;
    for (int c = 0; c < np; c++) dst[c * rn + (size_t)y * rw + x] = src[c * n + (size_t)(ry + y) * w + rx + x];
}

// io.rs:283-288: (v.clamp(0, 1) * 65535 + 0.5) as u16, interleaved RGB
__global__ void k_to_u16(const float *__restrict__ img3, size_t n, uint16_t *__restrict__ rgb) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int c = 0; c < 3; c++) {
        const float v = rclampf(img3[c * n + i], 0.0f, 1.0f) * 65535.0f + 0.5f;
        rgb[3 * i + c] = (uint16_t)fminf(v, 65535.0f);
    }
}

// align.rs:178-192 Interp::taps / margin
static double interp_margin(uint32_t ik) {
    static const int taps[6] = {1, 2, 4, 4, 6, 6};
    return (double)(taps[ik] / 2);
}

// align.rs:344-403 common_area, on the host: per row the columns every
// (non-identity) frame's source point stays `margin` inside the frame for
// (each bound of an affine or projective row is linear in the column), then
// the largest rectangle over consecutive rows; the whole frame when nothing
// is cut or nothing is common.
static void common_area_host(const double *sims, size_t n, size_t w, size_t h, uint32_t ik, uint32_t rect[4]) {
    rect[0] = 0;
    rect[1] = 0;
    rect[2] = (uint32_t)w;
    rect[3] = (uint32_t)h;
    if (w == 0 || h == 0) return;
    std::vector<std::array<double, 9>> invs;
    for (size_t i = 0; i < n; i++) {
        Sim s;
        std::copy(sims + 8 * i, sims + 8 * i + 8, s.v);
        if (s.is_id()) continue;
        double m[9];
        std::array<double, 9> iv;
        sim_matrix(s, w, h, m);
        inverse9(m, iv.data());
        invs.push_back(iv);
    }
    if (invs.empty()) return;
    const double mg = interp_margin(ik), xmax = (double)(w - 1) - mg, ymax = (double)(h - 1) - mg;
    std::vector<std::pair<int64_t, int64_t>> rows(h);
    for (size_t y = 0; y < h; y++) {
        double lo = 0.0, hi = (double)(w - 1);
        const double yf = (double)y;
        for (const auto &v : invs) {
            const double g = v[6], d = v[7] * yf + v[8];
            if (d <= 0.0 || g * (double)(w - 1) + d <= 0.0) {
                lo = 1.0;
                hi = 0.0;
                break;
            }
            const double cf[2][4] = {{v[0], v[1] * yf + v[2], mg, xmax}, {v[3], v[4] * yf + v[5], mg, ymax}};
            for (const auto &r : cf) {
                const double a = r[0], k = r[1], smin = r[2], smax = r[3];
                const double pq[2][2] = {{a - smin * g, k - smin * d}, {smax * g - a, smax * d - k}};
                for (const auto &e : pq) {
                    const double p = e[0], q = e[1];
                    if (std::fabs(p) < 1e-12) {
                        if (q < 0.0) {
                            lo = 1.0;
                            hi = 0.0;
                        }
                    } else if (p > 0.0) {
                        lo = std::max(lo, -q / p);
                    } else {
                        hi = std::min(hi, -q / p);
                    }
                }
            }
        }
        rows[y] = lo <= hi ? std::make_pair((int64_t)std::ceil(lo), (int64_t)std::floor(hi)) : std::make_pair((int64_t)1, (int64_t)0);
    }
    size_t bx = 0, by = 0, bw = 0, bh = 0;
    for (size_t y0 = 0; y0 < h; y0++) {
        int64_t lo = rows[y0].first, hi = rows[y0].second;
        for (size_t y1 = y0; y1 < h; y1++) {
            lo = std::max(lo, rows[y1].first);
            hi = std::min(hi, rows[y1].second);
            if (lo > hi) break;
            const size_t rw = (size_t)(hi - lo + 1), rh = y1 - y0 + 1;
            if (rw * rh > bw * bh) {
                bx = (size_t)lo;
                by = y0;
                bw = rw;
                bh = rh;
            }
        }
    }
    if (bw * bh == 0) return;
    rect[0] = (uint32_t)bx;
    rect[1] = (uint32_t)by;
    rect[2] = (uint32_t)bw;
    rect[3] = (uint32_t)bh;
}

}  // namespace lsh

extern "C" {

char *lsh_gauss_levels(const float *y, lsh_mem y_mem, uint32_t w, uint32_t h, float *packed, lsh_mem out_mem,
                       uint32_t *nlevels) {
    LSH_API_BEGIN
    LSH_REQUIRE(y && nlevels, "lsh_gauss_levels: null argument");
    LSH_REQUIRE(w > 0 && h > 0, "registration pyramid: empty plane");
    // align.rs:741-754: half size (rounded up) while h > 64 and w > 8
    std::vector<size_t> lw{w}, lh{h}, off{0};
    size_t tot = 0, cw = w, ch = h;
    while (ch > 64 && cw > 8) {
        cw = (cw + 1) / 2;
        ch = (ch + 1) / 2;
        lw.push_back(cw);
        lh.push_back(ch);
        off.push_back(tot);
        tot += cw * ch;
    }
    *nlevels = (uint32_t)(lw.size() - 1);
    if (!packed || tot == 0) return nullptr;
    DBuf<float> ys, outs;
    const float *prev = y;
    if (y_mem == LSH_HOST) {
        ys.alloc((size_t)w * h);
        copy_mem(ys.p, LSH_DEVICE, y, LSH_HOST, (size_t)w * h * sizeof(float));
        prev = ys.p;
    }
    float *out = packed;
    if (out_mem == LSH_HOST) {
        outs.alloc(tot);
        out = outs.p;
    }
    for (size_t l = 1; l < lw.size(); l++) {
        burt_level(prev, lw[l - 1], lh[l - 1], out + off[l], lw[l], lh[l]);
        prev = out + off[l];
    }
    if (out_mem == LSH_HOST) copy_mem(packed, LSH_HOST, outs.p, LSH_DEVICE, tot * sizeof(float));
    sync();
    LSH_API_END
}

char *lsh_common_area(const double *sims, uint32_t n, uint32_t w, uint32_t h, uint32_t interp, uint32_t *rect4) {
    LSH_API_BEGIN
    LSH_REQUIRE(rect4 && (sims || n == 0), "lsh_common_area: null argument");
    LSH_REQUIRE(interp <= 5, "common area: unknown interpolation");
    common_area_host(sims, n, w, h, interp, rect4);
    LSH_API_END
}

char *lsh_crop(const float *src, lsh_mem src_mem, uint32_t planes, uint32_t w, uint32_t h, const uint32_t *rect4,
               float *dst, lsh_mem dst_mem) {
    LSH_API_BEGIN
    LSH_REQUIRE(src && dst && rect4, "lsh_crop: null argument");
    const uint32_t rx = rect4[0], ry = rect4[1], rw = rect4[2], rh = rect4[3];
    LSH_REQUIRE(rw > 0 && rh > 0 && (uint64_t)rx + rw <= w && (uint64_t)ry + rh <= h, "crop: the rectangle leaves the image");
    const size_t n = (size_t)w * h, rn = (size_t)rw * rh;
    DBuf<float> s, d;
    const float *sp = src;
    if (src_mem == LSH_HOST) {
        s.alloc(planes * n);
        copy_mem(s.p, LSH_DEVICE, src, LSH_HOST, planes * n * sizeof(float));
        sp = s.p;
    }
    float *dp = dst;
    if (dst_mem == LSH_HOST) {
        d.alloc(planes * rn);
        dp = d.p;
    }
    launch_b("crop", 8.0 * planes * rn, k_crop, dim3(cdiv(rw, 64), cdiv(rh, 4)), dim3(64, 4), 0, sp, (int)w, (int)h,
             (int)planes, (int)rx, (int)ry, (int)rw, (int)rh, dp);
    if (dst_mem == LSH_HOST) copy_mem(dst, LSH_HOST, d.p, LSH_DEVICE, planes * rn * sizeof(float));
    sync();
    LSH_API_END
}

char *lsh_resize(const float *src, lsh_mem src_mem, uint32_t planes, uint32_t pw, uint32_t ph, uint32_t w, uint32_t h,
                 uint32_t interp, float *dst, lsh_mem dst_mem) {
    LSH_API_BEGIN
    LSH_REQUIRE(src && dst, "lsh_resize: null argument");
    LSH_REQUIRE(pw > 0 && ph > 0 && w > 0 && h > 0, "resize: empty image");
    LSH_REQUIRE(interp <= 5, "resize: unknown interpolation");
    const size_t pn = (size_t)pw * ph, on = (size_t)w * h;
    if (pw == w && ph == h) {  // prep.rs:47-49, 64-66: the same size is a copy
        copy_mem(dst, dst_mem, src, src_mem, planes * pn * sizeof(float));
        sync();
        return nullptr;
    }
    // prep.rs:67-69: scale kx, ky and the offset that puts output pixel o on source (o + 1/2) / k - 1/2
    const double sw = pw, sh = ph, kx = (double)w / sw, ky = (double)h / sh;
    Sim s;
    s.v[0] = (kx - 1.0) * (sw + 1.0) / (2.0 * sw);
    s.v[1] = (ky - 1.0) * (sh + 1.0) / (2.0 * sh);
    s.v[2] = kx;
    s.v[4] = ky / kx;
    double m[9];
    Inv9 inv;
    sim_matrix(s, pw, ph, m);
    inverse9(m, inv.m);
    const int affine = inv.m[6] == 0.0 && inv.m[7] == 0.0 && inv.m[8] == 1.0;
    DBuf<float> sb, db;
    const float *sp = src;
    if (src_mem == LSH_HOST) {
        sb.alloc(planes * pn);
        copy_mem(sb.p, LSH_DEVICE, src, LSH_HOST, planes * pn * sizeof(float));
        sp = sb.p;
    }
    float *dp = dst;
    if (dst_mem == LSH_HOST) {
        db.alloc(planes * on);
        dp = db.p;
    }
    const dim3 g(cdiv(w, 64), cdiv(h, 4)), b(64, 4);
    const double by = 4.0 * planes * ((double)on + (double)pn);
    const int ik = (int)interp;
    switch (taps_n(interp)) {
    case 2: launch_b("resize", by, k_resize<2>, g, b, 0, sp, (int)pw, (int)ph, inv, affine, ik, (int)planes, dp, (int)w, (int)h); break;
    case 4: launch_b("resize", by, k_resize<4>, g, b, 0, sp, (int)pw, (int)ph, inv, affine, ik, (int)planes, dp, (int)w, (int)h); break;
    default: launch_b("resize", by, k_resize<6>, g, b, 0, sp, (int)pw, (int)ph, inv, affine, ik, (int)planes, dp, (int)w, (int)h); break;
    }
    if (dst_mem == LSH_HOST) copy_mem(dst, LSH_HOST, db.p, LSH_DEVICE, planes * on * sizeof(float));
    sync();
    LSH_API_END
}

char *lsh_to_u16(const float *img3, lsh_mem img_mem, uint32_t w, uint32_t h, uint16_t *rgb, lsh_mem rgb_mem) {
    LSH_API_BEGIN
    LSH_REQUIRE(img3 && rgb, "lsh_to_u16: null argument");
    const size_t n = (size_t)w * h;
    DBuf<float> sb;
    DBuf<uint16_t> db;
    const float *sp = img3;
    if (img_mem == LSH_HOST) {
        sb.alloc(3 * n);
        copy_mem(sb.p, LSH_DEVICE, img3, LSH_HOST, 3 * n * sizeof(float));
        sp = sb.p;
    }
    uint16_t *dp = rgb;
    if (rgb_mem == LSH_HOST) {
        db.alloc(3 * n);
        dp = db.p;
    }
    launch_b("to_u16", 18.0 * n, k_to_u16, dim3(cdiv(n, 256)), dim3(256), 0, sp, n, dp);
    if (rgb_mem == LSH_HOST) copy_mem(rgb, LSH_HOST, db.p, LSH_DEVICE, 3 * n * sizeof(uint16_t));
    sync();
    LSH_API_END
}

}  // extern "C"
