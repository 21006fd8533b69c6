// wls.hip -- the confidence-weighted edge-aware WLS solve (depth.rs:653-1087)
// on gfx950, in three precisions (wls.hpp): the reference's FP32 MG-PCG, FP32
// MG-PCG with FP64 residual refinement and an exact FP64 coarsest solve, and an
// all-FP64 oracle.
//
// The operator is the 5-point stencil of (W + lambda L) held as per-cell
// coefficients (data weight, edge to the right, edge below) on a 2x2
// aggregation hierarchy.  Every kernel is a streaming pass over cell planes;
// dot products sum in f64 through per-block partials folded in a fixed order.
#include <algorithm>
#include <chrono>
#include <map>
#include <mutex>
#include <string>

#include "api.hpp"
#include "common.hpp"
#include "wls.hpp"

namespace lsh {

#define MG_OMEGA 0.8f  // depth.rs:770
#define MG_COARSE 32   // depth.rs:775
#define MG_MIN 16      // depth.rs:777

static dim3 g2(int w, int h) { return dim3(cdiv(w, 64), cdiv(h, 4)); }
static const dim3 B2(64, 4);
static unsigned g1(size_t n) { return cdiv(n, 256); }
static unsigned gred(size_t n) { return std::max(1u, std::min<unsigned>(cdiv(n, 256), 1024)); }

// depth.rs:653-672, and the lambda-scaled coefficients of depth.rs:980-981
__global__ void k_edges(const float *__restrict__ g, int w, int h, float inv, float lambda, float *__restrict__ ex,
                        float *__restrict__ ey, float *__restrict__ ax, float *__restrict__ ay) {
    int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= w || y >= h) return;
    size_t o = (size_t)y * w + x;
    float e1 = x + 1 < w ? expf(fabsf(g[o] - g[o + 1]) * inv) : 0.0f;
    float e2 = y + 1 < h ? expf(fabsf(g[o] - g[o + w]) * inv) : 0.0f;
    ex[o] = e1;
    ey[o] = e2;
    ax[o] = lambda * e1;
    ay[o] = lambda * e2;
}

// depth.rs:852-898: 2x2 Galerkin aggregation.  what: 1 data weights, 2 edges.
__global__ void k_coarsen(const float *__restrict__ wd, const float *__restrict__ ax, const float *__restrict__ ay,
                          int w, int h, float *__restrict__ cwd, float *__restrict__ cax, float *__restrict__ cay,
                          int cw, int ch, int what) {
    int xx = blockIdx.x * blockDim.x + threadIdx.x, yy = blockIdx.y * blockDim.y + threadIdx.y;
    if (xx >= cw || yy >= ch) return;
    size_t co = (size_t)yy * cw + xx;
    int ye = min(2 * yy + 2, h), xe = min(2 * xx + 2, w);
    if (what & 1) {
        float o = 0.0f;
        for (int y = 2 * yy; y < ye; y++) {
            float s = 0.0f;
            for (int x = 2 * xx; x < xe; x++) s += wd[(size_t)y * w + x];
            o += s;
        }
        cwd[co] = o;
    }
    if (what & 2) {
        float o = 0.0f;
        if (xx + 1 < cw)
            for (int y = 2 * yy; y < ye; y++) o += ax[(size_t)y * w + 2 * xx + 1];
        cax[co] = o;
        float s = 0.0f;
        if (yy + 1 < ch) {
            int y = 2 * yy + 1;
            for (int x = 2 * xx; x < xe; x++) s += ay[(size_t)y * w + x];
        }
        cay[co] = s;
    }
}

// depth.rs:812-834 (f32, as the reference); the FP64 hierarchy computes its own in f64
template <typename T>
__global__ void k_dinv(const T *__restrict__ wd, const T *__restrict__ ax, const T *__restrict__ ay, int w, int h,
                       T *__restrict__ dinv) {
    int i = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= w || y >= h) return;
    size_t k = (size_t)y * w + i;
    T v = wd[k];
    if (i > 0) v += ax[k - 1];
    if (i + 1 < w) v += ax[k];
    if (y > 0) v += ay[k - w];
    if (y + 1 < h) v += ay[k];
    dinv[k] = (T)1 / (v > (T)1e-12 ? v : (T)1e-12);
}

// depth.rs:994 and :1168
__global__ void k_data_w(const float *__restrict__ conf, const float *__restrict__ d, const float *__restrict__ u,
                         float tau, size_t n, float *__restrict__ wd) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float c = conf[i];
    if (u) c = c * fminf(tau / fabsf(d[i] - u[i]), 1.0f);
    wd[i] = fmaxf(c, 0.0f) + 1e-4f;
}

template <typename T>
__device__ __forceinline__ T apply_at(const T *__restrict__ wd, const T *__restrict__ ax, const T *__restrict__ ay,
                                      const T *__restrict__ x, int w, int h, int i, int y, size_t k) {
    T xk = x[k];
    T v = wd[k] * xk;
    if (i > 0) v += ax[k - 1] * (xk - x[k - 1]);
    if (i + 1 < w) v += ax[k] * (xk - x[k + 1]);
    if (y > 0) v += ay[k - w] * (xk - x[k - w]);
    if (y + 1 < h) v += ay[k] * (xk - x[k + w]);
    return v;
}

// depth.rs:786-810 apply; mode 0: A x; 1: b - A x; 2: x + w D^-1 (b - A x) (depth.rs:837-847)
template <typename T>
__global__ __launch_bounds__(256) void k_stencil(const T *__restrict__ wd, const T *__restrict__ ax,
                                                 const T *__restrict__ ay, const T *__restrict__ dinv,
                                                 const T *__restrict__ x, const T *__restrict__ b, int w, int h,
                                                 int mode, T *__restrict__ out, const int *__restrict__ skip) {
    int i = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= w || y >= h || (skip && *skip)) return;
    size_t k = (size_t)y * w + i;
    T v = apply_at(wd, ax, ay, x, w, h, i, y, k);
    if (mode == 0)
        out[k] = v;
    else if (mode == 1)
        out[k] = b[k] - v;
    else
        out[k] = x[k] + (T)MG_OMEGA * dinv[k] * (b[k] - v);
}

template <typename T>
__global__ void k_jacobi0(const T *__restrict__ dinv, const T *__restrict__ b, size_t n, T *__restrict__ x,
                          const int *__restrict__ skip) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && !(skip && *skip)) x[i] = (T)MG_OMEGA * dinv[i] * b[i];
}

// The column and row of a grid-stride loop's index over a w-wide grid,
// advanced by the stride with no division per step.
struct GridWalk {
    int i, y, di, dy, w;
    __device__ GridWalk(size_t k0, size_t stride, int w_) : w(w_) {
        y = (int)(k0 / (size_t)w_);
        i = (int)(k0 - (size_t)y * w_);
        dy = (int)(stride / (size_t)w_);
        di = (int)(stride - (size_t)dy * w_);
    }
    __device__ __forceinline__ void next() {
        i += di;
        y += dy;
        if (i >= w) {
            i -= w;
            y++;
        }
    }
};

// A level's descent step in one pass (depth.rs:837-847 from zero, :924-930 and
// :911-921): per coarse cell, the 2 x 2 fine cells under it get x = w D^-1 b
// (k_jacobi0), their residuals b - A x (k_stencil mode 1, the neighbours' x
// formed the same way where they are read), and the coarse cell gets their sum
// as the restriction adds them (row by row); the fine residual is never stored.
template <typename T>
__device__ __forceinline__ T jac0_at(const T *__restrict__ dinv, const T *__restrict__ b, size_t k) {
    return (T)MG_OMEGA * dinv[k] * b[k];
}
template <typename T>
__global__ __launch_bounds__(256) void k_mg_down(const T *__restrict__ wd, const T *__restrict__ ax,
                                                 const T *__restrict__ ay, const T *__restrict__ dinv,
                                                 const T *__restrict__ b, int w, int h, T *__restrict__ x,
                                                 T *__restrict__ bc, int cw, int ch, const int *__restrict__ skip) {
    const int xx = blockIdx.x * blockDim.x + threadIdx.x, yy = blockIdx.y * blockDim.y + threadIdx.y;
    if (xx >= cw || yy >= ch || (skip && *skip)) return;
    const int ye = min(2 * yy + 2, h), xe = min(2 * xx + 2, w);
    T o = 0;
    for (int y = 2 * yy; y < ye; y++) {
        T sr = 0;
        for (int i = 2 * xx; i < xe; i++) {
            const size_t k = (size_t)y * w + i;
            const T xk = jac0_at(dinv, b, k);
            x[k] = xk;
            T v = wd[k] * xk;
            if (i > 0) v += ax[k - 1] * (xk - jac0_at(dinv, b, k - 1));
            if (i + 1 < w) v += ax[k] * (xk - jac0_at(dinv, b, k + 1));
            if (y > 0) v += ay[k - w] * (xk - jac0_at(dinv, b, k - w));
            if (y + 1 < h) v += ay[k] * (xk - jac0_at(dinv, b, k + w));
            sr += b[k] - v;
        }
        o += sr;
    }
    bc[(size_t)yy * cw + xx] = o;
}

// A level's ascent step in one pass (depth.rs:901-908 then :837-847): the
// post-smoothing sweep applied to x + P xc (the prolongation's sum, formed where each
// value is read), into out; the prolonged x itself is never stored.
template <typename T>
__global__ __launch_bounds__(256) void k_mg_up(const T *__restrict__ wd, const T *__restrict__ ax,
                                               const T *__restrict__ ay, const T *__restrict__ dinv,
                                               const T *__restrict__ b, const T *__restrict__ x,
                                               const T *__restrict__ xc, int w, int h, int cw, T *__restrict__ out,
                                               const int *__restrict__ skip) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= w || y >= h || (skip && *skip)) return;
    const size_t k = (size_t)y * w + i;
    auto px = [&](int ii, int yy2) -> T {
        return x[(size_t)yy2 * w + ii] + xc[(size_t)(yy2 / 2) * cw + ii / 2];
    };
    const T xk = px(i, y);
    T v = wd[k] * xk;
    if (i > 0) v += ax[k - 1] * (xk - px(i - 1, y));
    if (i + 1 < w) v += ax[k] * (xk - px(i + 1, y));
    if (y > 0) v += ay[k - w] * (xk - px(i, y - 1));
    if (y + 1 < h) v += ay[k] * (xk - px(i, y + 1));
    out[k] = xk + (T)MG_OMEGA * dinv[k] * (b[k] - v);
}

// depth.rs:677-702: the Thomas solve of every row (wd null: unit weights).
// A block owns 16 rows and walks them in 64-column tiles staged through LDS
// (coalesced loads and stores); one thread per row runs the recurrence over
// the tile, in place, with the reference's arithmetic in the reference's order.
#define FGS_R 16
#define FGS_T 64
__global__ __launch_bounds__(256) void k_solve_rows(float *__restrict__ u, const float *__restrict__ f,
                                                    const float *__restrict__ wd, const float *__restrict__ a, int w,
                                                    int h, float lam, float *__restrict__ cp, float *__restrict__ dp) {
    __shared__ float TA[FGS_R][FGS_T + 1], TF[FGS_R][FGS_T + 1], TW[FGS_R][FGS_T + 1];
    const int y0 = blockIdx.x * FGS_R, r = threadIdx.x;
    const bool live = r < FGS_R && y0 + r < h;
    float prev_a = 0.0f, cprev = 0.0f, dprev = 0.0f;
    for (int x0 = 0; x0 < w; x0 += FGS_T) {
        __syncthreads();
        for (int i = threadIdx.x; i < FGS_R * FGS_T; i += 256) {
            const int rr = i / FGS_T, c = i % FGS_T, y = y0 + rr, x = x0 + c;
            if (y < h && x < w) {
                const size_t o = (size_t)y * w + x;
                TA[rr][c] = a[o];
                TF[rr][c] = f[o];
                TW[rr][c] = wd ? wd[o] : 1.0f;
            }
        }
        __syncthreads();
        if (live) {
            const int cn = min(FGS_T, w - x0);
            for (int c = 0; c < cn; c++) {
                const float ai = x0 + c + 1 < w ? TA[r][c] : 0.0f;
                const float wdi = TW[r][c];
                const float diag = wdi + lam * (prev_a + ai);
                const float lower = -lam * prev_a;
                const float upper = -lam * ai;
                float m = diag - lower * cprev;
                if (fabsf(m) < 1e-12f) m = 1e-12f;
                cprev = upper / m;
                dprev = (wdi * TF[r][c] - lower * dprev) / m;
                TA[r][c] = cprev;
                TF[r][c] = dprev;
                prev_a = ai;
            }
        }
        __syncthreads();
        for (int i = threadIdx.x; i < FGS_R * FGS_T; i += 256) {
            const int rr = i / FGS_T, c = i % FGS_T, y = y0 + rr, x = x0 + c;
            if (y < h && x < w) {
                const size_t o = (size_t)y * w + x;
                cp[o] = TA[rr][c];
                dp[o] = TF[rr][c];
            }
        }
    }
    float nx = 0.0f;
    for (int x0 = (w - 1) / FGS_T * FGS_T; x0 >= 0; x0 -= FGS_T) {
        __syncthreads();
        for (int i = threadIdx.x; i < FGS_R * FGS_T; i += 256) {
            const int rr = i / FGS_T, c = i % FGS_T, y = y0 + rr, x = x0 + c;
            if (y < h && x < w) {
                const size_t o = (size_t)y * w + x;
                TA[rr][c] = cp[o];
                TF[rr][c] = dp[o];
            }
        }
        __syncthreads();
        if (live) {
            for (int c = min(FGS_T, w - x0) - 1; c >= 0; c--) {
                nx = x0 + c == w - 1 ? TF[r][c] : TF[r][c] - TA[r][c] * nx;
                TA[r][c] = nx;
            }
        }
        __syncthreads();
        for (int i = threadIdx.x; i < FGS_R * FGS_T; i += 256) {
            const int rr = i / FGS_T, c = i % FGS_T, y = y0 + rr, x = x0 + c;
            if (y < h && x < w) u[(size_t)y * w + x] = TA[rr][c];
        }
    }
}

// depth.rs:707-746: the same down every column, one column per thread
// (coalesced across threads); the inputs of 16 rows are loaded ahead of the
// 16 recurrence steps that use them
#define FGS_C 16
// ---- the FGS sweeps, partitioned
// The row and column solves of depth.rs:674-746 (tridiagonal: lower -lam a_{i-1},
// diagonal wd_i + lam (a_{i-1} + a_i), upper -lam a_i, right-hand side wd_i f_i)
// by the partition method, so a row or column is solved by many lanes instead
// of one: its cells are cut into segments [L, R] whose last cell R is an
// interface unknown z.  Each segment's interior [L, R-1] is eliminated by its
// own Thomas sweep with three right-hand sides -- the data y, and the
// responses g, h to its left and right interface values -- so
// x_i = y_i + z_left g_i + z_right h_i there; the equation of each interface
// cell then couples z_{s-1}, z_s, z_{s+1} only, a tridiagonal system of one
// unknown per segment, and a last backward sweep per segment writes x.  The
// arithmetic per cell is the reference's (its pivot guard included); the
// order of elimination differs, so the sweep's result matches the sequential
// one to rounding (the guess PCG starts from, not the solution).
#define FGS_SEG 64  // segments per row: one wave, one lane each

struct Seg {
    int L, R;  // the segment's interior is [L, R-1], its interface cell R
};
__device__ __forceinline__ Seg seg_of(int s, int nseg, int n) {
    return Seg{(int)((long long)s * n / nseg), (int)((long long)(s + 1) * n / nseg) - 1};
}

// one row per block of one wave; the row's a, wd, f staged in LDS (3 * w
// floats, dynamic), where the forward sweep leaves c', y', g' and the last
// sweep the solution
__global__ __launch_bounds__(64) void k_fgs_rows_seg(float *__restrict__ u, const float *__restrict__ f,
                                                     const float *__restrict__ wd, const float *__restrict__ a, int w,
                                                     float lam) {
    extern __shared__ float lds[];
    float *A = lds, *W = lds + w, *F = lds + 2 * w;
    const int lane = threadIdx.x;
    const size_t row = (size_t)blockIdx.x * w;
    for (int i = lane; i < w; i += 64) {
        A[i] = a[row + i];
        W[i] = wd ? wd[row + i] : 1.0f;
        F[i] = f[row + i];
    }
    __syncthreads();
    const Seg sg = seg_of(lane, FGS_SEG, w);
    // forward over the interior, c' y' g' in place
    float prev_a = sg.L > 0 ? A[sg.L - 1] : 0.0f;
    float cpv = 0.0f, ypv = 0.0f, gpv = 0.0f, den_last = 1.0f, up_last = 0.0f;
    for (int i = sg.L; i < sg.R; i++) {
        const float ai = A[i];  // i + 1 <= R < w: an interior cell always has a right neighbour
        const float wdi = W[i];
        const float diag = wdi + lam * (prev_a + ai);
        const float lower = -lam * prev_a;
        const float upper = -lam * ai;
        float m = diag - lower * cpv;
        if (fabsf(m) < 1e-12f) m = 1e-12f;
        cpv = upper / m;
        ypv = (wdi * F[i] - lower * ypv) / m;
        gpv = i == sg.L ? -lower / m : (0.0f - lower * gpv) / m;
        A[i] = cpv;
        W[i] = ypv;
        F[i] = gpv;
        prev_a = ai;
        den_last = m;
        up_last = upper;
    }
    // the responses at the interior's two ends: backward sweeps of y, g, h
    float ye = 0.0f, ge = 1.0f, he = 0.0f, ys = 0.0f, gs = 0.0f, hs = 1.0f;  // (an empty interior: x_{R-1} = z_{s-1})
    if (sg.R > sg.L) {
        float yb = W[sg.R - 1], gb = F[sg.R - 1], hb = -up_last / den_last;
        ye = yb;
        ge = gb;
        he = hb;
        for (int i = sg.R - 2; i >= sg.L; i--) {
            const float c = A[i];
            yb = W[i] - c * yb;
            gb = F[i] - c * gb;
            hb = -c * hb;
        }
        ys = yb;
        gs = gb;
        hs = hb;
    }
    // the interface cell's equation in z_{s-1}, z_s, z_{s+1}
    const int R = sg.R;
    const float ar = R + 1 < w ? A[R] : 0.0f, wr = W[R];
    const float lo = -lam * prev_a, up = -lam * ar, di = wr + lam * (prev_a + ar), rhs = wr * F[R];
    float ysn = __shfl_down(ys, 1, 64), gsn = __shfl_down(gs, 1, 64), hsn = __shfl_down(hs, 1, 64);
    if (lane == FGS_SEG - 1) ysn = gsn = hsn = 0.0f;
    float Ac = lo * ge, Bc = lo * he + di + up * gsn, Cc = up * hsn, Dc = rhs - lo * ye - up * ysn;
    // parallel cyclic reduction over the wave
    for (int d = 1; d < FGS_SEG; d <<= 1) {
        float Am = __shfl_up(Ac, d, 64), Bm = __shfl_up(Bc, d, 64), Cm = __shfl_up(Cc, d, 64), Dm = __shfl_up(Dc, d, 64);
        float Ap = __shfl_down(Ac, d, 64), Bp = __shfl_down(Bc, d, 64), Cp = __shfl_down(Cc, d, 64),
              Dp = __shfl_down(Dc, d, 64);
        if (lane < d) {
            Am = Cm = Dm = 0.0f;
            Bm = 1.0f;
        }
        if (lane + d >= FGS_SEG) {
            Ap = Cp = Dp = 0.0f;
            Bp = 1.0f;
        }
        const float al = -Ac / Bm, ga = -Cc / Bp;
        const float nA = al * Am, nC = ga * Cp, nB = Bc + al * Cm + ga * Ap, nD = Dc + al * Dm + ga * Dp;
        Ac = nA;
        Bc = nB;
        Cc = nC;
        Dc = nD;
    }
    const float z = Dc / Bc;
    float zl = __shfl_up(z, 1, 64);
    if (lane == 0) zl = 0.0f;
    // the segment's cells: x_i = (y'_i + z_{s-1} g'_i + z_s h'_i) - c'_i x_{i+1}
    if (sg.R > sg.L) {
        float x = W[sg.R - 1] + zl * F[sg.R - 1] + z * (-up_last / den_last);
        A[sg.R - 1] = x;
        for (int i = sg.R - 2; i >= sg.L; i--) {
            x = (W[i] + zl * F[i]) - A[i] * x;
            A[i] = x;
        }
    }
    A[R] = z;
    __syncthreads();
    for (int i = lane; i < w; i += 64) u[row + i] = A[i];
}

// 64 columns per block, FGS_CSEG segments down each column (one wave per
// segment, one lane per column): c', y', g' go through global scratch (cq, yq,
// gq), the interface system of each column through LDS, solved by one lane
#define FGS_CSEG 16
__global__ __launch_bounds__(64 * FGS_CSEG) void k_fgs_cols_seg(float *__restrict__ u, const float *__restrict__ f,
                                                              const float *__restrict__ wd, const float *__restrict__ a,
                                                              int w, int h, float lam, float *__restrict__ cq,
                                                              float *__restrict__ yq, float *__restrict__ gq) {
    __shared__ float S0[FGS_CSEG][64], S1[FGS_CSEG][64], S2[FGS_CSEG][64];
    __shared__ float EA[FGS_CSEG][64], EB[FGS_CSEG][64], EC[FGS_CSEG][64], ED[FGS_CSEG][64], Z[FGS_CSEG][64];
    const int tx = threadIdx.x & 63, s = threadIdx.x >> 6;
    const int x = blockIdx.x * 64 + tx;
    const bool live = x < w;
    const Seg sg = seg_of(s, FGS_CSEG, h);
    const int xc = live ? x : w - 1;
    auto at = [&](int y) { return (size_t)y * w + xc; };
    float prev_a = sg.L > 0 ? a[at(sg.L - 1)] : 0.0f;
    float cpv = 0.0f, ypv = 0.0f, gpv = 0.0f, den_last = 1.0f, up_last = 0.0f;
    for (int y = sg.L; y < sg.R; y++) {
        const size_t o = at(y);
        const float ai = a[o], wdi = wd ? wd[o] : 1.0f, fi = f[o];
        const float diag = wdi + lam * (prev_a + ai);
        const float lower = -lam * prev_a;
        const float upper = -lam * ai;
        float m = diag - lower * cpv;
        if (fabsf(m) < 1e-12f) m = 1e-12f;
        cpv = upper / m;
        ypv = (wdi * fi - lower * ypv) / m;
        gpv = y == sg.L ? -lower / m : (0.0f - lower * gpv) / m;
        if (live) {
            cq[o] = cpv;
            yq[o] = ypv;
            gq[o] = gpv;
        }
        prev_a = ai;
        den_last = m;
        up_last = upper;
    }
    float ye = 0.0f, ge = 1.0f, he = 0.0f, ys = 0.0f, gs = 0.0f, hs = 1.0f;
    if (sg.R > sg.L) {
        float yb = ypv, gb = gpv, hb = -up_last / den_last;
        ye = yb;
        ge = gb;
        he = hb;
        for (int y = sg.R - 2; y >= sg.L; y--) {
            const size_t o = at(y);
            const float c = cq[o];
            yb = yq[o] - c * yb;
            gb = gq[o] - c * gb;
            hb = -c * hb;
        }
        ys = yb;
        gs = gb;
        hs = hb;
    }
    S0[s][tx] = ys;
    S1[s][tx] = gs;
    S2[s][tx] = hs;
    __syncthreads();
    {
        const int R = sg.R;
        const size_t o = at(R);
        const float ar = R + 1 < h ? a[o] : 0.0f, wr = wd ? wd[o] : 1.0f;
        const float lo = -lam * prev_a, up = -lam * ar, di = wr + lam * (prev_a + ar), rhs = wr * f[o];
        const bool last = s == FGS_CSEG - 1;
        const float ysn = last ? 0.0f : S0[s + 1][tx], gsn = last ? 0.0f : S1[s + 1][tx], hsn = last ? 0.0f : S2[s + 1][tx];
        EA[s][tx] = lo * ge;
        EB[s][tx] = lo * he + di + up * gsn;
        EC[s][tx] = up * hsn;
        ED[s][tx] = rhs - lo * ye - up * ysn;
    }
    __syncthreads();
    if (s == 0) {  // the column's interface system, by Thomas
        float cpz = 0.0f, dpz = 0.0f;
        for (int k = 0; k < FGS_CSEG; k++) {
            float m = EB[k][tx] - EA[k][tx] * cpz;
            if (fabsf(m) < 1e-12f) m = 1e-12f;
            cpz = EC[k][tx] / m;
            dpz = (ED[k][tx] - EA[k][tx] * dpz) / m;
            EC[k][tx] = cpz;
            ED[k][tx] = dpz;
        }
        float zz = ED[FGS_CSEG - 1][tx];
        Z[FGS_CSEG - 1][tx] = zz;
        for (int k = FGS_CSEG - 2; k >= 0; k--) {
            zz = ED[k][tx] - EC[k][tx] * zz;
            Z[k][tx] = zz;
        }
    }
    __syncthreads();
    if (!live) return;
    const float z = Z[s][tx], zl = s > 0 ? Z[s - 1][tx] : 0.0f;
    if (sg.R > sg.L) {
        size_t o = at(sg.R - 1);
        float xv = ypv + zl * gpv + z * (-up_last / den_last);
        u[o] = xv;
        for (int y = sg.R - 2; y >= sg.L; y--) {
            o = at(y);
            xv = (yq[o] + zl * gq[o]) - cq[o] * xv;
            u[o] = xv;
        }
    }
    u[at(sg.R)] = z;
}

__global__ void k_solve_cols(float *__restrict__ u, const float *__restrict__ f, const float *__restrict__ wd,
                             const float *__restrict__ a, int w, int h, float lam, float *__restrict__ cp,
                             float *__restrict__ dp) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= w) return;
    float cprev = 0.0f, dprev = 0.0f, prev_a = 0.0f;
    for (int y0 = 0; y0 < h; y0 += FGS_C) {
        float A[FGS_C], F[FGS_C], WD[FGS_C];
#pragma unroll
        for (int k = 0; k < FGS_C; k++) {
            const int y = y0 + k;
            const size_t i = (size_t)min(y, h - 1) * w + x;
            A[k] = y + 1 < h ? a[i] : 0.0f;
            F[k] = f[i];
            WD[k] = wd ? wd[i] : 1.0f;
        }
#pragma unroll
        for (int k = 0; k < FGS_C; k++) {
            const int y = y0 + k;
            if (y >= h) continue;
            const size_t i = (size_t)y * w + x;
            const float ai = A[k], wdi = WD[k];
            const float diag = wdi + lam * (prev_a + ai);
            const float lower = -lam * prev_a;
            const float upper = -lam * ai;
            float m = diag - lower * cprev;
            if (fabsf(m) < 1e-12f) m = 1e-12f;
            cprev = upper / m;
            dprev = (wdi * F[k] - lower * dprev) / m;
            cp[i] = cprev;
            dp[i] = dprev;
            prev_a = ai;
        }
    }
    float nx = 0.0f;
    for (int y1 = h - 1; y1 >= 0; y1 -= FGS_C) {
        float C[FGS_C], D[FGS_C];
#pragma unroll
        for (int k = 0; k < FGS_C; k++) {
            const size_t i = (size_t)max(y1 - k, 0) * w + x;
            C[k] = cp[i];
            D[k] = dp[i];
        }
#pragma unroll
        for (int k = 0; k < FGS_C; k++) {
            const int y = y1 - k;
            if (y < 0) continue;
            nx = y == h - 1 ? D[k] : D[k] - C[k] * nx;
            u[(size_t)y * w + x] = nx;
        }
    }
}

__device__ __forceinline__ double wsum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// fixed-order grid reduction of one double per thread (per-block partials, last block folds)
template <class Fin>
__device__ void grid_sum1f(double v, double *__restrict__ part, unsigned *__restrict__ ticket, Fin fin);
__device__ void grid_sum1(double v, double *__restrict__ part, unsigned *__restrict__ ticket, double *__restrict__ out) {
    grid_sum1f(v, part, ticket, [=](double t) { *out = t; });
}

// Sums v over the grid: per-block partials, then the last block to finish folds
// them in block order (deterministic) and thread 0 of it calls fin(total).
template <class Fin>
__device__ void grid_sum1f(double v, double *__restrict__ part, unsigned *__restrict__ ticket, Fin fin) {
    __shared__ double sh[4];
    __shared__ bool last;
    int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    double s = wsum(v);
    if (lane == 0) sh[wv] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        part[blockIdx.x] = ((sh[0] + sh[1]) + sh[2]) + sh[3];
        __threadfence();
        last = atomicAdd(ticket, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    double acc = 0.0;
    for (unsigned b = threadIdx.x; b < gridDim.x; b += blockDim.x) acc += part[b];
    __syncthreads();
    s = wsum(acc);
    if (lane == 0) sh[wv] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        fin(((sh[0] + sh[1]) + sh[2]) + sh[3]);
        *ticket = 0;
    }
}

// sum a*b in f64 (depth.rs:1037)
template <typename T>
__global__ __launch_bounds__(256) void k_dot(const T *__restrict__ a, const T *__restrict__ b, size_t n,
                                             double *__restrict__ part, unsigned *__restrict__ ticket,
                                             double *__restrict__ out) {
    double s = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        s += (double)a[i] * (double)b[i];
    grid_sum1(s, part, ticket, out);
}

// One MG-PCG solve's scalars in device memory: the loop's kernels read and
// update them, so the loop runs with no host round trip per iteration (the
// host checks `done` between batches of iterations).  Each scalar is computed
// by thread 0 of the last block of a reduction, with the reference's
// arithmetic: alpha = (T)(rz / pap), beta = (T)(rz_new / rz), the relative
// residual sqrt(rr) / bnorm (kept in f32 by the FP32 solver, depth.rs:1051).
struct CgState {
    double rz, pap, rr, rel, rel0, alpha, beta, bnorm, tol;
    unsigned iters;
    int done;  // 0: running; 1: below the tolerance; 2: stopped on p.Ap <= 0
};
#define CG_WORDS 16  // doubles the state takes in device memory

template <typename T>
__host__ __device__ __forceinline__ double cg_rel(double rr, double bnorm) {
    return sizeof(T) == 4 ? (double)(float)(sqrt(rr) / bnorm) : sqrt(rr) / bnorm;
}
template <typename T>
__host__ __device__ __forceinline__ bool cg_below(double rel, double tol) {
    return sizeof(T) == 4 ? (float)rel < (float)tol : rel < tol;
}

// the loop's entry: rr = r.r (rz = r.z was summed into cg->rz), the residual,
// and whether it is already below the tolerance
__global__ void k_cg_start(CgState *cg, double bnorm, double tol, int f32) {
    if (threadIdx.x || blockIdx.x) return;
    cg->bnorm = bnorm;
    cg->tol = tol;
    cg->rel = f32 ? cg_rel<float>(cg->rr, bnorm) : cg_rel<double>(cg->rr, bnorm);
    cg->rel0 = cg->rel;
    cg->iters = 0;
    cg->done = (f32 ? cg_below<float>(cg->rel, tol) : cg_below<double>(cg->rel, tol)) ? 1 : 0;
}

// an iteration's first step (depth.rs:1058-1063): ap = A p, pap = p.ap, alpha
template <typename T>
__global__ __launch_bounds__(256) void k_spmv_cg(const T *__restrict__ wd, const T *__restrict__ ax,
                                                 const T *__restrict__ ay, const T *__restrict__ p, int w, int h,
                                                 T *__restrict__ ap, double *__restrict__ part,
                                                 unsigned *__restrict__ ticket, CgState *__restrict__ cg) {
    if (cg->done) return;
    double s = 0.0;
    const size_t n = (size_t)w * h, k0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x,
                 stride = (size_t)gridDim.x * blockDim.x;
    GridWalk g(k0, stride, w);
    for (size_t k = k0; k < n; k += stride, g.next()) {
        const int y = g.y, i = g.i;
        T v = apply_at(wd, ax, ay, p, w, h, i, y, k);
        ap[k] = v;
        s += (double)p[k] * (double)v;
    }
    grid_sum1f(s, part, ticket, [=](double pap) {
        cg->iters++;
        cg->pap = pap;
        if (pap <= 0.0)
            cg->done = 2;
        else
            cg->alpha = (double)(T)(cg->rz / pap);
    });
}

// u += alpha p; r -= alpha ap; rr = r.r; the residual and the stop test (depth.rs:1064-1070)
template <typename T>
__global__ __launch_bounds__(256) void k_update_cg(T *__restrict__ u, T *__restrict__ r, const T *__restrict__ p,
                                                   const T *__restrict__ ap, size_t n, double *__restrict__ part,
                                                   unsigned *__restrict__ ticket, CgState *__restrict__ cg) {
    if (cg->done) return;
    const T alpha = (T)cg->alpha;
    double s = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        u[i] += alpha * p[i];
        T ri = r[i] - alpha * ap[i];
        r[i] = ri;
        s += (double)ri * (double)ri;
    }
    grid_sum1f(s, part, ticket, [=](double rr) {
        cg->rr = rr;
        cg->rel = cg_rel<T>(rr, cg->bnorm);
        if (cg_below<T>(cg->rel, cg->tol)) cg->done = 1;
    });
}

// rz_new = r.z, beta = rz_new / rz, rz = rz_new (depth.rs:1071-1073); rz
// itself at the loop's entry (first)
template <typename T>
__global__ __launch_bounds__(256) void k_dot_cg(const T *__restrict__ a, const T *__restrict__ b, size_t n,
                                                double *__restrict__ part, unsigned *__restrict__ ticket,
                                                CgState *__restrict__ cg, int first) {
    if (!first && cg->done) return;
    double s = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        s += (double)a[i] * (double)b[i];
    grid_sum1f(s, part, ticket, [=](double t) {
        if (first == 2) {
            cg->rr = t;
        } else if (first == 1) {
            cg->rz = t;
        } else {
            cg->beta = (double)(T)(t / cg->rz);
            cg->rz = t;
        }
    });
}

// p = z + beta p (depth.rs:1074)
template <typename T>
__global__ void k_pupdate_cg(T *__restrict__ p, const T *__restrict__ z, size_t n, const CgState *__restrict__ cg) {
    if (cg->done) return;
    const T beta = (T)cg->beta;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = z[i] + beta * p[i];
}

// b = wd * d (depth.rs:1039), in T
template <typename T>
__global__ void k_rhs(const float *__restrict__ wd, const float *__restrict__ d, size_t n, T *__restrict__ b) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (T)wd[i] * (T)d[i];
}

// The true residual in FP64 of the FP32 iterate: r = b - A u with the f32
// coefficients and u promoted; scaled copy for the next inner solve later.
template <typename U>
__global__ __launch_bounds__(256) void k_resid64(const float *__restrict__ wd, const float *__restrict__ ax,
                                                 const float *__restrict__ ay, const U *__restrict__ u,
                                                 const double *__restrict__ b, int w, int h, double *__restrict__ r,
                                                 double *__restrict__ part, unsigned *__restrict__ ticket,
                                                 double *__restrict__ out) {
    double s = 0.0;
    const size_t n = (size_t)w * h, k0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x,
                 stride = (size_t)gridDim.x * blockDim.x;
    GridWalk g(k0, stride, w);
    for (size_t k = k0; k < n; k += stride, g.next()) {
        const int y = g.y, i = g.i;
        double uk = u[k];
        double v = (double)wd[k] * uk;
        if (i > 0) v += (double)ax[k - 1] * (uk - (double)u[k - 1]);
        if (i + 1 < w) v += (double)ax[k] * (uk - (double)u[k + 1]);
        if (y > 0) v += (double)ay[k - w] * (uk - (double)u[k - w]);
        if (y + 1 < h) v += (double)ay[k] * (uk - (double)u[k + w]);
        double rk = b[k] - v;
        r[k] = rk;
        s += rk * rk;
    }
    grid_sum1(s, part, ticket, out);
}

__global__ void k_scale_to_f32(const double *__restrict__ a, double s, size_t n, float *__restrict__ out) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (float)(a[i] * s);
}

// u += e / s (the refinement's correction)
__global__ void k_correct(double *__restrict__ x, const float *__restrict__ e, double inv_s, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] += (double)e[i] * inv_s;
}

template <typename A, typename B>
__global__ void k_convert(const A *__restrict__ a, size_t n, B *__restrict__ b) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (B)a[i];
}

// max |u - u64| per block (the host folds the blocks)
__global__ void k_absdiff_max(const float *__restrict__ u, const double *__restrict__ v, size_t n,
                              double *__restrict__ part) {
    __shared__ double sh[4];
    double m = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        m = fmax(m, fabs((double)u[i] - v[i]));
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_down(m, o, 64));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
}

// The exact solve of the coarsest grid (at most 16x16 cells): a banded
// Cholesky factor in f64 of the 5-point operator, half bandwidth p = cw in the
// row-major cell order, L(i,j) for i-p <= j <= i at Lb[i*(p+1) + i-j].  The
// factor is right-looking in LDS: per column, the pivot, the p entries below
// it, then the p(p+1)/2 trailing updates in parallel (one thread per pair).
#define CHOL_MAXN 256
#define CHOL_STRIDE 17
template <typename T>
__global__ __launch_bounds__(256) void k_chol_factor(const T *__restrict__ wd, const T *__restrict__ ax,
                                                     const T *__restrict__ ay, int cw, int ch,
                                                     double *__restrict__ Lb) {
    __shared__ double S[CHOL_MAXN * CHOL_STRIDE];
    const int n = cw * ch, p = cw, st = p + 1, tid = threadIdx.x;
    for (int idx = tid; idx < n * st; idx += blockDim.x) {
        int i = idx / st, o = idx - i * st, j = i - o;
        double v = 0.0;
        if (o == 0) {
            int yi = i / cw, xi = i - yi * cw;
            v = (double)wd[i];
            if (xi > 0) v += (double)ax[i - 1];
            if (xi + 1 < cw) v += (double)ax[i];
            if (yi > 0) v += (double)ay[i - cw];
            if (yi + 1 < ch) v += (double)ay[i];
        } else if (o == 1 && j >= 0 && (i % cw) > 0) {
            v = -(double)ax[j];
        } else if (o == p && j >= 0) {
            v = -(double)ay[j];
        }
        S[idx] = v;
    }
    __syncthreads();
    const int a = tid / 16 + 1, b = tid % 16 + 1;  // trailing pair (row j+a, column j+b), b <= a <= p
    for (int j = 0; j < n; j++) {
        if (tid == 0) {
            double d = S[j * st];
            S[j * st] = sqrt(d > 1e-300 ? d : 1e-300);
        }
        __syncthreads();
        if (tid < p && j + 1 + tid < n) {
            int i = j + 1 + tid;
            S[i * st + i - j] /= S[j * st];
        }
        __syncthreads();
        if (b <= a && a <= p && j + a < n) {
            int i = j + a, k = j + b;
            S[i * st + i - k] -= S[i * st + i - j] * S[k * st + k - j];
        }
        __syncthreads();
    }
    for (int idx = tid; idx < n * st; idx += blockDim.x) Lb[idx] = S[idx];
}

// x = (L L^T)^-1 b by column-oriented substitutions in LDS, S and y in the
// caller's LDS (the block's threads, any count: the column work takes the
// first p of them)
template <typename T>
__device__ __forceinline__ void chol_solve_body(const double *__restrict__ Lb, int cw, int ch, const T *__restrict__ b,
                                                T *__restrict__ x, double *S, double *y) {
    const int n = cw * ch, p = cw, st = p + 1, tid = threadIdx.x;
    for (int idx = tid; idx < n * st; idx += blockDim.x) S[idx] = Lb[idx];
    for (int i = tid; i < n; i += blockDim.x) y[i] = (double)b[i];
    __syncthreads();
    for (int j = 0; j < n; j++) {  // L y = b
        if (tid == 0) y[j] /= S[j * st];
        __syncthreads();
        if (tid < p && j + 1 + tid < n) {
            int i = j + 1 + tid;
            y[i] -= S[i * st + i - j] * y[j];
        }
        __syncthreads();
    }
    for (int j = n - 1; j >= 0; j--) {  // L^T x = y
        if (tid == 0) y[j] /= S[j * st];
        __syncthreads();
        if (tid < p && j - 1 - tid >= 0) {
            int i = j - 1 - tid;
            y[i] -= S[j * st + j - i] * y[j];
        }
        __syncthreads();
    }
    for (int i = tid; i < n; i += blockDim.x) x[i] = (T)y[i];
}

// chol_solve_body by the 64 lanes of one wave (lanes in step; a wave's LDS
// accesses complete in issue order, so the steps need only be kept in order by
// the compiler)
__device__ __forceinline__ void wave_order() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
template <typename T>
__device__ __forceinline__ void chol_solve_wave(const double *__restrict__ Lb, int cw, int ch, const T *__restrict__ b,
                                                T *__restrict__ x, double *S, double *y) {
    const int n = cw * ch, p = cw, st = p + 1, tid = threadIdx.x;
    for (int idx = tid; idx < n * st; idx += 64) S[idx] = Lb[idx];
    for (int i = tid; i < n; i += 64) y[i] = (double)b[i];
    wave_order();
    for (int j = 0; j < n; j++) {  // L y = b
        if (tid == 0) y[j] /= S[j * st];
        wave_order();
        if (tid < p && j + 1 + tid < n) {
            int i = j + 1 + tid;
            y[i] -= S[i * st + i - j] * y[j];
        }
        wave_order();
    }
    for (int j = n - 1; j >= 0; j--) {  // L^T x = y
        if (tid == 0) y[j] /= S[j * st];
        wave_order();
        if (tid < p && j - 1 - tid >= 0) {
            int i = j - 1 - tid;
            y[i] -= S[j * st + j - i] * y[j];
        }
        wave_order();
    }
    for (int i = tid; i < n; i += 64) x[i] = (T)y[i];
}

// the coarsest grid solved by the FP64 factor (one wavefront)
template <typename T>
__global__ __launch_bounds__(64) void k_chol_solve(const double *__restrict__ Lb, int cw, int ch,
                                                   const T *__restrict__ b, T *__restrict__ x,
                                                   const int *__restrict__ skip) {
    __shared__ double S[CHOL_MAXN * CHOL_STRIDE];
    __shared__ double y[CHOL_MAXN];
    if (skip && *skip) return;
    chol_solve_body<T>(Lb, cw, ch, b, x, S, y);
}

// The V-cycle's coarse tail (depth.rs:924-947 from level l0 down) in one block:
// on levels this small every step of the recursion is a few cells per thread,
// so a launch per step (two per level, and the coarsest grid's 32 sweeps) costs
// far more than the arithmetic.  The block walks the reference's steps in its
// order with a barrier between them, each cell computed with the operations the
// finer levels' kernels use (apply_at, the restriction's row sums, the
// prolongation's add), so the values are the ones those steps give.  A level's
// post-smoothing writes its r buffer, as k_mg_up does (the host then swaps x and
// r of the levels, as it does after those launches).
#define VT_MAXL 12
#define VT_THREADS 1024
#define VT_CELLS 16384  // the tail starts at the first level of at most this many cells
template <typename T>
struct VTail {
    const T *wd[VT_MAXL], *ax[VT_MAXL], *ay[VT_MAXL], *dinv[VT_MAXL];
    T *x[VT_MAXL], *b[VT_MAXL], *r[VT_MAXL];
    int w[VT_MAXL], h[VT_MAXL];
    int nl;     // levels in the tail: l0 .. l0 + nl - 1, the last the coarsest
    int exact;  // the coarsest grid by the FP64 factor (else MG_COARSE damped-Jacobi sweeps)
    const double *chol;
    const int *skip;  // non-null and set: the launch does nothing (a PCG that has stopped)
};

template <typename T>
__device__ __forceinline__ void vt_sweep(const VTail<T> &P, int l, const T *__restrict__ x, T *__restrict__ o) {
    // one damped-Jacobi sweep of level l from x into o (k_stencil mode 2)
    const int w = P.w[l], h = P.h[l], n = w * h;
    const T *bb = P.b[l], *di = P.dinv[l];
    for (int k = threadIdx.x; k < n; k += VT_THREADS) {
        const int y = k / w, i = k - y * w;
        const T v = apply_at(P.wd[l], P.ax[l], P.ay[l], x, w, h, i, y, (size_t)k);
        o[k] = x[k] + (T)MG_OMEGA * di[k] * (bb[k] - v);
    }
    __syncthreads();
}

// k_jacobi0: x = w D^-1 b on level l
template <typename T>
__device__ __forceinline__ void vt_first(const VTail<T> &P, int l) {
    const int n = P.w[l] * P.h[l];
    for (int k = threadIdx.x; k < n; k += VT_THREADS) P.x[l][k] = (T)MG_OMEGA * P.dinv[l][k] * P.b[l][k];
    __syncthreads();
}

template <typename T>
__global__ __launch_bounds__(VT_THREADS) void k_vtail(const VTail<T> P) {
    __shared__ double S[CHOL_MAXN * CHOL_STRIDE];
    __shared__ double Y[CHOL_MAXN];
    const int tid = threadIdx.x, L = P.nl;
    if (P.skip && *P.skip) return;
    for (int l = 0; l + 1 < L; l++) {
        vt_first(P, l);
        // the residual into r (k_stencil mode 1), restricted into the next level's b (depth.rs:911-921)
        const int w = P.w[l], h = P.h[l], n = w * h;
        T *rr = P.r[l];
        for (int k = tid; k < n; k += VT_THREADS) {
            const int y = k / w, i = k - y * w;
            rr[k] = P.b[l][k] - apply_at(P.wd[l], P.ax[l], P.ay[l], (const T *)P.x[l], w, h, i, y, (size_t)k);
        }
        __syncthreads();
        const int cw = P.w[l + 1], chh = P.h[l + 1];
        for (int k = tid; k < cw * chh; k += VT_THREADS) {
            const int yy = k / cw, xx = k - yy * cw;
            const int ye = min(2 * yy + 2, h), xe = min(2 * xx + 2, w);
            T o = 0;
            for (int y = 2 * yy; y < ye; y++) {
                T s = 0;
                for (int x = 2 * xx; x < xe; x++) s += rr[(size_t)y * w + x];
                o += s;
            }
            P.b[l + 1][k] = o;
        }
        __syncthreads();
    }
    // the coarsest grid; its x ends in P.x (the factor, or an even number of
    // sweeps after the first) or in P.r
    const T *xcoarse = P.x[L - 1];
    if (P.exact) {
        // by wave 0 alone: the substitutions are a chain of small steps, and
        // a barrier of one wave is only an ordering of its LDS accesses
        if (tid < 64) chol_solve_wave<T>(P.chol, P.w[L - 1], P.h[L - 1], (const T *)P.b[L - 1], P.x[L - 1], S, Y);
        __syncthreads();
    } else {
        vt_first(P, L - 1);
        T *a = P.x[L - 1], *c = P.r[L - 1];
        for (int k = 1; k < MG_COARSE; k++) {
            vt_sweep(P, L - 1, a, c);
            T *t = a;
            a = c;
            c = t;
        }
        xcoarse = a;
    }
    for (int l = L - 2; l >= 0; l--) {
        // the prolongation into x (depth.rs:901-908), then the post-smoothing sweep from x into r (where
        // the level's x then is)
        const int w = P.w[l], h = P.h[l], cw = P.w[l + 1];
        T *x = P.x[l];
        for (int k = tid; k < w * h; k += VT_THREADS) {
            const int y = k / w, i = k - y * w;
            x[k] += xcoarse[(size_t)(y / 2) * cw + i / 2];
        }
        __syncthreads();
        vt_sweep(P, l, (const T *)x, P.r[l]);
        xcoarse = P.r[l];
    }
}

// Where a solver's buffers come from: device allocations (base null, counting
// off), or consecutive 256-byte-aligned pieces of a caller's block -- counted
// only, with nothing placed, when base is null and counting is on.
struct Arena {
    char *base = nullptr;
    size_t cap = 0, used = 0;
    bool counting = false;
    template <typename T>
    void get(DBuf<T> &b, size_t n) {
        if (!base && !counting) {
            b.alloc(n);
            return;
        }
        used = (used + 255) & ~(size_t)255;
        const size_t at = used;
        used += n * sizeof(T);
        if (counting) return;
        LSH_REQUIRE(used <= cap, "wls: the solver's buffers exceed the scratch block");
        b.view((T *)(base + at), n);
    }
};

template <typename T>
struct Lev {
    int w = 0, h = 0;
    DBuf<T> wd, ax, ay, dinv, x, b, r;
    void alloc(int w_, int h_, Arena &a) {
        w = w_;
        h = h_;
        size_t n = (size_t)w * h;
        for (auto *bf : {&wd, &ax, &ay, &dinv, &x, &b, &r}) a.get(*bf, n);
    }
    void alloc(int w_, int h_) {
        Arena a;
        alloc(w_, h_, a);
    }
};

struct WlsSolver {
    int w = 0, h = 0;
    float lambda = 0.f;
    DBuf<float> ex, ey, cp, dp, f;
    std::vector<Lev<float>> L;
    std::vector<Lev<double>> L64;
    bool l64_ok = false, chol_ok = false;
    DBuf<double> chol;
    DBuf<float> r, z, p, ap, e;
    DBuf<double> b64, r64, z64, p64, ap64, x64;
    DBuf<double> part, out;
    DBuf<double> cgs;  // a PCG solve's scalars (CgState)
    DBuf<unsigned> ticket;
    bool noalloc = false;  // buffers in a caller's block: scalars come back through pageable host memory
};

void WlsDeleter::operator()(WlsSolver *s) const { delete s; }

static double read_out(WlsSolver &s) {
    if (s.noalloc) {
        double v = 0.0;
        LSH_CHECK(hipMemcpyAsync(&v, s.out.p, sizeof(double), hipMemcpyDeviceToHost, stream()));
        LSH_CHECK(hipStreamSynchronize(stream()));
        return v;
    }
    double *pin = (double *)pinned_scratch(64);
    LSH_CHECK(hipMemcpyAsync(pin, s.out.p, sizeof(double), hipMemcpyDeviceToHost, stream()));
    sync();
    return pin[0];
}

__global__ void k_fill_u32(unsigned *p, size_t n, unsigned v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// The solver's buffers (depth.rs:976-988's vectors and the multigrid hierarchy
// down to MG_MIN), from the arena; with64 adds the FP64 right-hand side and
// residual the true residual needs.
static void wls_layout(WlsSolver &s, int w, int h, Arena &a, bool with64) {
    s.w = w;
    s.h = h;
    size_t n = (size_t)w * h;
    for (auto *b : {&s.ex, &s.ey, &s.cp, &s.dp, &s.f, &s.r, &s.z, &s.p, &s.ap}) a.get(*b, n);
    if (with64)
        for (auto *b : {&s.b64, &s.r64}) a.get(*b, n);
    a.get(s.part, 1024);
    a.get(s.out, 1);
    a.get(s.cgs, CG_WORDS);
    a.get(s.ticket, 1);
    s.L.clear();
    s.L.emplace_back();
    s.L[0].alloc(w, h, a);
    while (std::max(s.L.back().w, s.L.back().h) > MG_MIN) {
        int cw = (s.L.back().w + 1) / 2, chh = (s.L.back().h + 1) / 2;
        s.L.emplace_back();
        s.L.back().alloc(cw, chh, a);
    }
}

// depth.rs:976-988: the guide's edges and the hierarchy's coarse operators
static void wls_init(WlsSolver &s, const float *guide, float lambda, float sigma_c) {
    const int w = s.w, h = s.h;
    const size_t n = (size_t)w * h;
    s.lambda = lambda;
    launch_b("wls_ticket", 4.0, k_fill_u32, dim3(1), dim3(64), 0, s.ticket.p, (size_t)1, 0u);
    float inv = -1.0f / fmaxf(sigma_c, 1e-6f);
    launch_b("wls_edges", 20.0 * n, k_edges, g2(w, h), B2, 0, guide, w, h, inv, lambda, s.ex.p, s.ey.p, s.L[0].ax.p,
             s.L[0].ay.p);
    for (size_t l = 1; l < s.L.size(); l++) {
        Lev<float> &fl = s.L[l - 1], &c = s.L[l];
        launch_b("wls_coarsen", 8.0 * ((double)fl.w * fl.h + (double)c.w * c.h), k_coarsen, g2(c.w, c.h), B2, 0,
                 (const float *)fl.wd.p, (const float *)fl.ax.p, (const float *)fl.ay.p, fl.w, fl.h, c.wd.p, c.ax.p,
                 c.ay.p, c.w, c.h, 2);
    }
}

WlsPtr wls_create(const float *guide, int w, int h, float lambda, float sigma_c) {
    WlsPtr sp(new WlsSolver());
    Arena a;
    wls_layout(*sp, w, h, a, false);
    wls_init(*sp, guide, lambda, sigma_c);
    return sp;
}

size_t wls_scratch_bytes(int w, int h) {
    WlsSolver s;
    Arena a;
    a.counting = true;
    wls_layout(s, w, h, a, true);
    return a.used;
}

WlsPtr wls_create_in(void *block, size_t cap, const float *guide, int w, int h, float lambda, float sigma_c) {
    WlsPtr sp(new WlsSolver());
    Arena a;
    a.base = (char *)block;
    a.cap = cap;
    sp->noalloc = true;
    wls_layout(*sp, w, h, a, true);
    wls_init(*sp, guide, lambda, sigma_c);
    return sp;
}

void wls_set_weights(WlsSolver *sp, const float *conf, const float *d, const float *u, float tau) {
    WlsSolver &s = *sp;
    size_t n = (size_t)s.w * s.h;
    launch_b("wls_weights", 8.0 * n, k_data_w, dim3(g1(n)), dim3(256), 0, conf, d, u, tau, n, s.L[0].wd.p);
    for (size_t l = 0; l < s.L.size(); l++) {
        Lev<float> &lv = s.L[l];
        if (l > 0) {
            Lev<float> &fl = s.L[l - 1];
            launch_b("wls_coarsen", 4.0 * ((double)fl.w * fl.h + (double)lv.w * lv.h), k_coarsen, g2(lv.w, lv.h), B2,
                     0, (const float *)fl.wd.p, (const float *)fl.ax.p, (const float *)fl.ay.p, fl.w, fl.h, lv.wd.p,
                     lv.ax.p, lv.ay.p, lv.w, lv.h, 1);
        }
        launch_b("wls_dinv", 16.0 * lv.w * lv.h, k_dinv<float>, g2(lv.w, lv.h), B2, 0, (const float *)lv.wd.p,
                 (const float *)lv.ax.p,
               (const float *)lv.ay.p, lv.w, lv.h, lv.dinv.p);
    }
    s.l64_ok = false;
    s.chol_ok = false;
}

// the coarsest grid's factor (from the f32 coefficients, promoted)
static void ensure_chol(WlsSolver &s) {
    if (s.chol_ok) return;
    Lev<float> &c = s.L.back();
    LSH_REQUIRE(c.w <= 16 && c.h <= 16, "coarsest grid larger than 16x16");
    s.chol.ensure((size_t)c.w * c.h * (c.w + 1));
    // reads the three coefficient planes, writes the banded factor (f64)
    launch_b("wls_coarse_factor", 12.0 * c.w * c.h + 8.0 * c.w * c.h * (c.w + 1), k_chol_factor<float>, dim3(1),
             dim3(256), 0, (const float *)c.wd.p,
           (const float *)c.ax.p, (const float *)c.ay.p, c.w, c.h, s.chol.p);
    s.chol_ok = true;
}

// the FP64 hierarchy: the f32 coefficients promoted, dinv in f64
static void ensure_l64(WlsSolver &s) {
    if (s.l64_ok) return;
    if (s.L64.size() != s.L.size()) {
        s.L64.clear();
        for (auto &lv : s.L) {
            s.L64.emplace_back();
            s.L64.back().alloc(lv.w, lv.h);
        }
    }
    for (size_t l = 0; l < s.L.size(); l++) {
        Lev<float> &a = s.L[l];
        Lev<double> &b = s.L64[l];
        size_t n = (size_t)a.w * a.h;
        launch_b("wls_promote", 12.0 * n, k_convert<float, double>, dim3(g1(n)), dim3(256), 0, (const float *)a.wd.p, n, b.wd.p);
        launch_b("wls_promote", 12.0 * n, k_convert<float, double>, dim3(g1(n)), dim3(256), 0, (const float *)a.ax.p, n, b.ax.p);
        launch_b("wls_promote", 12.0 * n, k_convert<float, double>, dim3(g1(n)), dim3(256), 0, (const float *)a.ay.p, n, b.ay.p);
        launch("wls_dinv", k_dinv<double>, g2(b.w, b.h), B2, 0, (const double *)b.wd.p, (const double *)b.ax.p,
               (const double *)b.ay.p, b.w, b.h, b.dinv.p);
    }
    s.l64_ok = true;
}

// a damped-Jacobi sweep of level t (from zero: x = w D^-1 b), b its right-hand
// side; skip: the device flag that turns the launch into a no-op (nullable)
template <typename T>
static void jacobi(Lev<T> &t, bool zero, const T *b, const int *skip) {
    size_t n = (size_t)t.w * t.h;
    double sz = sizeof(T);
    if (zero) {
        launch_b("mg_jacobi", 3.0 * sz * n, k_jacobi0<T>, dim3(g1(n)), dim3(256), 0, (const T *)t.dinv.p, b, n, t.x.p,
                 skip);
        return;
    }
    launch_b("mg_jacobi", 7.0 * sz * n, k_stencil<T>, g2(t.w, t.h), B2, 0, (const T *)t.wd.p, (const T *)t.ax.p,
             (const T *)t.ay.p, (const T *)t.dinv.p, (const T *)t.x.p, b, t.w, t.h, 2, t.r.p, skip);
    std::swap(t.x, t.r);
}

// depth.rs:924-947; exact: the coarsest grid solved by the FP64 factor instead
// of 32 sweeps.  b0: level 0's right-hand side (the residual the caller
// preconditions, read where it is); skip: as for jacobi.
template <typename T>
static void vcycle(WlsSolver &s, std::vector<Lev<T>> &H, size_t l, bool exact, const T *b0, const int *skip) {
    Lev<T> &t = H[l];
    const T *bl = l == 0 ? b0 : (const T *)t.b.p;
    if ((size_t)t.w * t.h <= VT_CELLS && H.size() - l <= VT_MAXL) {
        // the coarse tail in one block (k_vtail), then the pointer swaps its
        // sweeps stand for
        VTail<T> P;
        std::memset(&P, 0, sizeof(P));
        P.nl = (int)(H.size() - l);
        P.exact = exact ? 1 : 0;
        P.chol = s.chol.p;
        P.skip = skip;
        double bytes = 0.0;
        for (int q = 0; q < P.nl; q++) {
            Lev<T> &v = H[l + q];
            P.wd[q] = v.wd.p;
            P.ax[q] = v.ax.p;
            P.ay[q] = v.ay.p;
            P.dinv[q] = v.dinv.p;
            P.x[q] = v.x.p;
            P.b[q] = q == 0 ? (T *)bl : v.b.p;  // (level l's b is only read)
            P.r[q] = v.r.p;
            P.w[q] = v.w;
            P.h[q] = v.h;
            // the bytes the per-level kernels count for these steps
            const double n = (double)v.w * v.h, sz = sizeof(T);
            if (q + 1 < P.nl)
                bytes += sz * (19.0 * n + 2.0 * (double)H[l + q + 1].w * H[l + q + 1].h);
            else
                bytes += exact ? 2.0 * sz * n : sz * (3.0 + 7.0 * (MG_COARSE - 1)) * n;
        }
        launch_b("mg_tail", bytes, k_vtail<T>, dim3(1), dim3(VT_THREADS), 0, P);
        for (int q = 0; q + 1 < P.nl; q++) std::swap(H[l + q].x, H[l + q].r);
        if (!exact)
            for (int k = 1; k < MG_COARSE; k++) std::swap(H.back().x, H.back().r);
        return;
    }
    if (l + 1 == H.size()) {
        if (exact)
            launch("mg_coarse_solve", k_chol_solve<T>, dim3(1), dim3(64), 0, (const double *)s.chol.p, t.w, t.h, bl,
                   t.x.p, skip);
        else
            for (int k = 0; k < MG_COARSE; k++) jacobi(t, k == 0, bl, skip);
        return;
    }
    // the pre-smoothing from zero, the residual and its restriction in one
    // pass; the coarse correction; its prolongation and the post-smoothing in one
    const double n = (double)t.w * t.h, sz = sizeof(T);
    Lev<T> &c = H[l + 1];
    const double nc = (double)c.w * c.h;
    launch_b("mg_down", sz * (6.0 * n + nc), k_mg_down<T>, g2(c.w, c.h), B2, 0, (const T *)t.wd.p, (const T *)t.ax.p,
             (const T *)t.ay.p, (const T *)t.dinv.p, bl, t.w, t.h, t.x.p, c.b.p, c.w, c.h, skip);
    vcycle(s, H, l + 1, exact, (const T *)nullptr, skip);
    launch_b("mg_up", sz * (7.0 * n + nc), k_mg_up<T>, g2(t.w, t.h), B2, 0, (const T *)t.wd.p, (const T *)t.ax.p,
             (const T *)t.ay.p, (const T *)t.dinv.p, bl, (const T *)t.x.p, (const T *)c.x.p, t.w, t.h, c.w, t.r.p,
             skip);
    std::swap(t.x, t.r);
}

template <typename T>
static double dot(WlsSolver &s, const T *a, const T *b, size_t n) {
    launch_b("cg_dot", 2.0 * sizeof(T) * n, k_dot<T>, dim3(gred(n)), dim3(256), 0, a, b, n, s.part.p, s.ticket.p, s.out.p);
    return read_out(s);
}

// z = M^-1 r: one V-cycle on r where it is; z is the hierarchy's level-0 x
// (valid until the next V-cycle)
template <typename T>
static const T *precond(WlsSolver &s, std::vector<Lev<T>> &H, const T *r, bool exact, const int *skip) {
    vcycle(s, H, 0, exact, r, skip);
    return H[0].x.p;
}

static CgState read_cg(WlsSolver &s) {
    CgState c;
    if (s.noalloc) {
        LSH_CHECK(hipMemcpyAsync(&c, s.cgs.p, sizeof(CgState), hipMemcpyDeviceToHost, stream()));
        LSH_CHECK(hipStreamSynchronize(stream()));
        return c;
    }
    void *pin = pinned_scratch(sizeof(CgState));
    LSH_CHECK(hipMemcpyAsync(pin, s.cgs.p, sizeof(CgState), hipMemcpyDeviceToHost, stream()));
    sync();
    std::memcpy(&c, pin, sizeof(CgState));
    return c;
}

// MG-PCG (depth.rs:1036-1077) on H from the iterate u for the right-hand side b:
// the reference's sequence of operations; rel is the recurrence's residual over
// bnorm.  The loop runs on the device: its scalars live in CgState, every
// kernel of an iteration does nothing once the solve has stopped, and the host
// queues iterations in batches and reads the state between them (the first
// batch a few iterations, the next sized by the convergence rate so far), so
// an iteration costs no host round trip; the iterations queued past the stop
// cost their launches only.
template <typename T>
static uint32_t pcg(WlsSolver &s, std::vector<Lev<T>> &H, T *u, const T *b, double bnorm, double tol, uint32_t max_iters,
                    bool exact, double &rel_out, T *r, T * /*z: the hierarchy's level-0 x instead*/, T *p, T *ap) {
    const int w = H[0].w, h = H[0].h;
    const size_t n = (size_t)w * h;
    Lev<T> &t0 = H[0];
    const double sz = sizeof(T);
    CgState *cg = (CgState *)s.cgs.p;
    const int *done = &cg->done;
    launch_b("cg_residual", 6.0 * sz * n, k_stencil<T>, g2(w, h), B2, 0, (const T *)t0.wd.p, (const T *)t0.ax.p,
             (const T *)t0.ay.p, (const T *)t0.dinv.p, (const T *)u, b, w, h, 1, r, (const int *)nullptr);
    const T *z = precond(s, H, (const T *)r, exact, (const int *)nullptr);
    LSH_CHECK(hipMemcpyAsync(p, z, n * sizeof(T), hipMemcpyDeviceToDevice, stream()));
    launch_b("cg_dot", 2.0 * sz * n, k_dot_cg<T>, dim3(gred(n)), dim3(256), 0, (const T *)r, z, n, s.part.p, s.ticket.p,
             cg, 1);
    launch_b("cg_dot", 2.0 * sz * n, k_dot_cg<T>, dim3(gred(n)), dim3(256), 0, (const T *)r, (const T *)r, n, s.part.p,
             s.ticket.p, cg, 2);
    launch_b("cg_start", 2.0 * sizeof(CgState), k_cg_start, dim3(1), dim3(64), 0, cg, bnorm, tol, sizeof(T) == 4 ? 1 : 0);
    auto iteration = [&]() {
        launch_b("cg_spmv", 5.0 * sz * n, k_spmv_cg<T>, dim3(gred(n)), dim3(256), 0, (const T *)t0.wd.p,
                 (const T *)t0.ax.p, (const T *)t0.ay.p, (const T *)p, w, h, ap, s.part.p, s.ticket.p, cg);
        launch_b("cg_update", 6.0 * sz * n, k_update_cg<T>, dim3(gred(n)), dim3(256), 0, u, r, (const T *)p,
                 (const T *)ap, n, s.part.p, s.ticket.p, cg);
        const T *zz = precond(s, H, (const T *)r, exact, done);
        launch_b("cg_dot", 2.0 * sz * n, k_dot_cg<T>, dim3(gred(n)), dim3(256), 0, (const T *)r, zz, n, s.part.p,
                 s.ticket.p, cg, 0);
        launch_b("cg_pupdate", 3.0 * sz * n, k_pupdate_cg<T>, dim3(g1(n)), dim3(256), 0, p, zz, n,
                 (const CgState *)cg);
    };
    uint32_t queued = 0, batch = std::min<uint32_t>(max_iters, 4);
    CgState st = read_cg(s);  // the entry's state (and the initial residual for the rate)
    while (!st.done && queued < max_iters) {
        for (uint32_t k = 0; k < batch; k++) iteration();
        queued += batch;
        st = read_cg(s);
        if (st.done || queued >= max_iters) break;
        // the iterations still needed at the rate so far (at least one)
        uint32_t more = 1;
        if (st.iters > 0 && st.rel > 0.0 && st.rel0 > st.rel) {
            const double rate = std::pow(st.rel / st.rel0, 1.0 / st.iters);
            const double need = std::log(tol / st.rel) / std::log(rate);
            if (need > 1.0 && need < 1e6) more = (uint32_t)std::ceil(need);
        }
        batch = std::min(more, max_iters - queued);
    }
    rel_out = st.rel;
    return st.iters;
}

// depth.rs:1019-1034: three rounds of row and column sweeps with the lambda_t
// schedule.  Each sweep reads the previous one's output directly (rows: d or
// f -> u, columns: u -> f), which is what the reference's copies between them
// amount to; the last result moves into u once.
static void fgs_guess(WlsSolver &s, const float *d, float *u) {
    const int w = s.w, h = s.h, T = 3;
    const size_t n = (size_t)w * h;
    const float *wd = s.L[0].wd.p;
    const float *src = d;
    // partitioned sweeps where every segment has an interior (rows of at least
    // 2 * FGS_SEG cells, columns of 2 * FGS_CSEG) and a row fits in LDS; else
    // the sequential ones
    const bool rows_seg = w >= 2 * FGS_SEG && (size_t)3 * w * sizeof(float) <= 64 * 1024;
    const bool cols_seg = h >= 2 * FGS_CSEG;
    for (int t = 1; t <= T; t++) {
        float lam_t = 1.5f * s.lambda * std::pow(4.0f, (float)(T - t)) / (std::pow(4.0f, (float)T) - 1.0f);
        if (rows_seg)
            launch_b("fgs_rows", (t == 1 ? 16.0 : 12.0) * n, k_fgs_rows_seg, dim3(h), dim3(64),
                     (size_t)3 * w * sizeof(float), u, src, t == 1 ? wd : nullptr, (const float *)s.ex.p, w, lam_t);
        else
            launch_b("fgs_rows", 16.0 * n, k_solve_rows, dim3(cdiv(h, FGS_R)), dim3(256), 0, u, src,
                     t == 1 ? wd : nullptr, (const float *)s.ex.p, w, h, lam_t, s.cp.p, s.dp.p);
        if (cols_seg)
            launch_b("fgs_cols", 12.0 * n, k_fgs_cols_seg, dim3(cdiv(w, 64)), dim3(64 * FGS_CSEG), 0, s.f.p,
                     (const float *)u, nullptr, (const float *)s.ey.p, w, h, lam_t, s.cp.p, s.dp.p, s.z.p);
        else
            launch_b("fgs_cols", 16.0 * n, k_solve_cols, dim3(cdiv(w, 64)), dim3(64), 0, s.f.p, (const float *)u,
                     nullptr, (const float *)s.ey.p, w, h, lam_t, s.cp.p, s.dp.p);
        src = s.f.p;
    }
    LSH_CHECK(hipMemcpyAsync(u, s.f.p, n * sizeof(float), hipMemcpyDeviceToDevice, stream()));
}

WlsResult wls_solve(WlsSolver *sp, const float *d, uint32_t max_iters, double tol, int mode, const float *warm,
                    float *u, double *u64) {
    WlsSolver &s = *sp;
    const int w = s.w, h = s.h;
    const size_t n = (size_t)w * h;
    const float *wd = s.L[0].wd.p;
    WlsResult res;
    if (w == 1 || h == 1) {  // depth.rs:1009-1017: the exact 1-D solve
        if (h == 1)
            launch("fgs_rows", k_solve_rows, dim3(1), dim3(256), 0, u, d, wd, (const float *)s.ex.p, w, h, s.lambda, s.cp.p, s.dp.p);
        else
            launch("fgs_cols", k_solve_cols, dim3(1), dim3(64), 0, u, d, wd, (const float *)s.ey.p, w, h, s.lambda, s.cp.p, s.dp.p);
        if (u64) launch_b("wls_promote", 12.0 * n, k_convert<float, double>, dim3(g1(n)), dim3(256), 0, (const float *)u, n, u64);
        return res;
    }
    if (warm) {
        if (warm != u) LSH_CHECK(hipMemcpyAsync(u, warm, n * sizeof(float), hipMemcpyDeviceToDevice, stream()));
    } else {
        fgs_guess(s, d, u);
    }
    if (mode == WLS_REFERENCE) {
        float *b = s.f.p;
        launch_b("wls_rhs", 12.0 * n, k_rhs<float>, dim3(g1(n)), dim3(256), 0, wd, d, n, b);
        double bnorm = std::sqrt(dot<float>(s, b, b, n));
        if (bnorm < 1e-30) bnorm = 1e-30;
        res.iters = pcg<float>(s, s.L, u, b, bnorm, tol, max_iters, false, res.rel, s.r.p, s.z.p, s.p.p, s.ap.p);
        return res;
    }
    for (auto *bf : {&s.b64, &s.r64}) bf->ensure(n);
    launch_b("wls_rhs", 16.0 * n, k_rhs<double>, dim3(g1(n)), dim3(256), 0, wd, d, n, s.b64.p);
    double bnorm = std::sqrt(dot<double>(s, s.b64.p, s.b64.p, n));
    if (bnorm < 1e-30) bnorm = 1e-30;
    ensure_chol(s);
    if (mode == WLS_REFINED) {
        // the iterate x (FP64) starts at the guess; each step: r = b - A x in
        // FP64, the FP32 MG-PCG solves A e = r / |r|, x += |r| e in FP64
        s.e.ensure(n);
        s.x64.ensure(n);
        double *x = u64 ? u64 : s.x64.p;
        launch_b("wls_promote", 12.0 * n, k_convert<float, double>, dim3(g1(n)), dim3(256), 0, (const float *)u, n, x);
        const Lev<float> &t0 = s.L[0];
        for (uint32_t outer = 0; outer < 64; outer++) {
            launch_b("wls_resid64", 36.0 * n, k_resid64<double>, dim3(gred(n)), dim3(256), 0, (const float *)t0.wd.p,
                     (const float *)t0.ax.p, (const float *)t0.ay.p, (const double *)x, (const double *)s.b64.p, w, h,
                     s.r64.p, s.part.p, s.ticket.p, s.out.p);
            double rr = read_out(s);
            res.rel = std::sqrt(rr) / bnorm;
            if (res.rel < tol || res.iters >= max_iters || rr == 0.0) break;
            res.outer++;
            double scale = 1.0 / std::sqrt(rr);
            launch_b("wls_demote", 12.0 * n, k_scale_to_f32, dim3(g1(n)), dim3(256), 0, (const double *)s.r64.p, scale, n,
                     s.f.p);
            s.e.zero();
            // the inner FP32 solve is asked for what is still missing, with a
            // margin, but never below what FP32 vectors hold
            double inner_tol = std::min(0.5, std::max(1e-5, 0.5 * tol / res.rel));
            double inner_rel = 0.0;
            uint32_t it = pcg<float>(s, s.L, s.e.p, s.f.p, 1.0, inner_tol, max_iters - res.iters, true, inner_rel, s.r.p,
                                     s.z.p, s.p.p, s.ap.p);
            res.iters += it;
            launch_b("wls_correct", 20.0 * n, k_correct, dim3(g1(n)), dim3(256), 0, x, (const float *)s.e.p, 1.0 / scale, n);
            if (it == 0) break;
        }
        launch("wls_demote", k_convert<double, float>, dim3(g1(n)), dim3(256), 0, (const double *)x, n, u);
        return res;
    }
    // FP64 oracle
    ensure_l64(s);
    for (auto *bf : {&s.z64, &s.p64, &s.ap64, &s.x64}) bf->ensure(n);
    double *x = u64 ? u64 : s.x64.p;
    launch_b("wls_promote", 12.0 * n, k_convert<float, double>, dim3(g1(n)), dim3(256), 0, (const float *)u, n, x);
    res.iters = pcg<double>(s, s.L64, x, (const double *)s.b64.p, bnorm, tol, max_iters, true, res.rel, s.r64.p, s.z64.p,
                            s.p64.p, s.ap64.p);
    launch("wls_demote", k_convert<double, float>, dim3(g1(n)), dim3(256), 0, (const double *)x, n, u);
    return res;
}

double wls_true_rel(WlsSolver *sp, const float *d, const float *u) {
    WlsSolver &s = *sp;
    const int w = s.w, h = s.h;
    const size_t n = (size_t)w * h;
    for (auto *bf : {&s.b64, &s.r64}) bf->ensure(n);
    const Lev<float> &t0 = s.L[0];
    launch_b("wls_rhs", 16.0 * n, k_rhs<double>, dim3(g1(n)), dim3(256), 0, (const float *)t0.wd.p, d, n, s.b64.p);
    double bnorm = std::sqrt(dot<double>(s, s.b64.p, s.b64.p, n));
    if (bnorm < 1e-30) bnorm = 1e-30;
    launch_b("wls_resid64", 28.0 * n, k_resid64<float>, dim3(gred(n)), dim3(256), 0, (const float *)t0.wd.p,
             (const float *)t0.ax.p, (const float *)t0.ay.p, u, (const double *)s.b64.p, w, h, s.r64.p, s.part.p,
             s.ticket.p, s.out.p);
    return std::sqrt(read_out(s)) / bnorm;
}

double wls_max_abs_diff(WlsSolver *sp, const float *u, const double *u64) {
    WlsSolver &s = *sp;
    size_t n = (size_t)s.w * s.h;
    unsigned g = gred(n);
    launch("wls_absdiff", k_absdiff_max, dim3(g), dim3(256), 0, u, u64, n, s.part.p);
    std::vector<double> h(g);
    copy_mem(h.data(), LSH_HOST, s.part.p, LSH_DEVICE, g * sizeof(double));
    double m = 0.0;
    for (double v : h) m = std::max(m, v);
    return m;
}

}  // namespace lsh

extern "C" {

char *lsh_wls_solve(const float *guide, const float *conf, const float *d, const float *warm, lsh_mem in_mem,
                    uint32_t w, uint32_t h, float lambda, float sigma_c, uint32_t mode, double tol,
                    uint32_t max_iters, uint32_t want_err, float *u, double *u64, lsh_mem out_mem,
                    lsh_wls_report *rep) {
    using namespace lsh;
    LSH_API_BEGIN
    LSH_REQUIRE(guide && conf && d && u, "lsh_wls_solve: null argument");
    LSH_REQUIRE(w > 0 && h > 0, "wls: empty grid");
    LSH_REQUIRE(mode <= LSH_WLS_FP64, "wls: unknown mode");
    LSH_REQUIRE(lambda > 0.0f, "wls: lambda must be positive");
    LSH_REQUIRE(!u64 || mode != LSH_WLS_REFERENCE, "wls: u64 is an output of the refined and FP64 modes");
    const size_t n = (size_t)w * h;
    DBuf<float> in_d;
    const float *g = guide, *c = conf, *dd = d, *wm = warm;
    if (in_mem == LSH_HOST) {
        in_d.alloc((warm ? 4 : 3) * n);
        copy_mem(in_d.p, LSH_DEVICE, guide, LSH_HOST, n * sizeof(float));
        copy_mem(in_d.p + n, LSH_DEVICE, conf, LSH_HOST, n * sizeof(float));
        copy_mem(in_d.p + 2 * n, LSH_DEVICE, d, LSH_HOST, n * sizeof(float));
        if (warm) copy_mem(in_d.p + 3 * n, LSH_DEVICE, warm, LSH_HOST, n * sizeof(float));
        g = in_d.p;
        c = in_d.p + n;
        dd = in_d.p + 2 * n;
        wm = warm ? in_d.p + 3 * n : nullptr;
    }
    DBuf<float> u_d;
    DBuf<double> u64_d, ref64;
    float *uu = u;
    if (out_mem == LSH_HOST) {
        u_d.alloc(n);
        uu = u_d.p;
    }
    double *uu64 = u64;
    if (mode != LSH_WLS_REFERENCE && (!u64 || out_mem == LSH_HOST)) {
        u64_d.alloc(n);
        uu64 = u64_d.p;
    }
    WlsPtr s = wls_create(g, (int)w, (int)h, lambda, sigma_c);
    wls_set_weights(s.get(), c, nullptr, nullptr, 0.0f);
    hipEvent_t e0, e1;
    LSH_CHECK(hipEventCreate(&e0));
    LSH_CHECK(hipEventCreate(&e1));
    WlsResult r;
    float ms = 0.0f;
    try {
        LSH_CHECK(hipEventRecord(e0, stream()));
        r = wls_solve(s.get(), dd, max_iters, tol, (int)mode, wm, uu, uu64);
        LSH_CHECK(hipEventRecord(e1, stream()));
        LSH_CHECK(hipEventSynchronize(e1));
        LSH_CHECK(hipEventElapsedTime(&ms, e0, e1));
    } catch (...) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    double err = -1.0;
    if (want_err && w > 1 && h > 1) {
        if (mode == LSH_WLS_FP64) {
            err = wls_max_abs_diff(s.get(), uu, uu64);
        } else {
            ref64.alloc(n);
            DBuf<float> tmp(n);
            wls_solve(s.get(), dd, 2000, 1e-12, WLS_FP64, nullptr, tmp.p, ref64.p);
            err = wls_max_abs_diff(s.get(), uu, ref64.p);
        }
    }
    double true_rel = (w > 1 && h > 1) ? wls_true_rel(s.get(), dd, uu) : 0.0;
    if (out_mem == LSH_HOST) {
        copy_mem(u, LSH_HOST, uu, LSH_DEVICE, n * sizeof(float));
        if (u64) copy_mem(u64, LSH_HOST, uu64, LSH_DEVICE, n * sizeof(double));
    } else if (u64 && uu64 != u64) {
        copy_mem(u64, LSH_DEVICE, uu64, LSH_DEVICE, n * sizeof(double));
    }
    sync();
    if (rep) {
        rep->iters = r.iters;
        rep->outer = r.outer;
        rep->rel = r.rel;
        rep->true_rel = true_rel;
        rep->err_fp64 = err;
        rep->ms = ms;
    }
    LSH_API_END
}

}  // extern "C"

// ------------------------------------------------ the WLS solve with no allocation
// A variant of lsh_wls_solve for callers that cannot allocate (code under HIP
// graph capture, serving loops, harnesses that forbid hipMalloc): every buffer
// of the solver lives in `scratch`, one file-scope device array per GPU, of
// LSH_WLS_NOALLOC_BYTES; grids needing more are refused.  In host code the
// array's name is no device address: the address each device holds it at is
// taken once per device (hipGetSymbolAddress, with that device current) and
// only that pointer is handed to kernels.
static __device__ __align__(256) unsigned char scratch[LSH_WLS_NOALLOC_BYTES];

namespace {
std::mutex g_scratch_mu;
std::map<int, void *> g_scratch_ptr;

void *scratch_block() {
    int dev = 0;
    LSH_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_ptr.find(dev);
    if (it != g_scratch_ptr.end()) return it->second;
    void *ptr = nullptr;
    LSH_CHECK(hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch)));
    LSH_REQUIRE(ptr, "wls no-alloc: no device address for the scratch array");
    g_scratch_ptr[dev] = ptr;
    return ptr;
}
}  // namespace

extern "C" {

char *lsh_wls_solve_noalloc(const float *guide, const float *conf, const float *d, const float *warm, uint32_t w,
                            uint32_t h, float lambda, float sigma_c, double tol, uint32_t max_iters, float *u,
                            lsh_wls_report *rep) {
    using namespace lsh;
    LSH_API_BEGIN
    LSH_REQUIRE(guide && conf && d && u, "lsh_wls_solve_noalloc: null argument");
    LSH_REQUIRE(w > 0 && h > 0, "wls: empty grid");
    LSH_REQUIRE(lambda > 0.0f, "wls: lambda must be positive");
    const size_t need = wls_scratch_bytes((int)w, (int)h);
    if (need > (size_t)LSH_WLS_NOALLOC_BYTES)
        throw Error("wls no-alloc: a " + std::to_string(w) + " x " + std::to_string(h) + " grid needs " +
                    std::to_string(need) + " bytes of scratch, above the " + std::to_string(LSH_WLS_NOALLOC_BYTES) +
                    " the library holds");
    void *blk = scratch_block();
    const auto t0 = std::chrono::steady_clock::now();
    WlsPtr s = wls_create_in(blk, LSH_WLS_NOALLOC_BYTES, guide, (int)w, (int)h, lambda, sigma_c);
    wls_set_weights(s.get(), conf, nullptr, nullptr, 0.0f);
    WlsResult r = wls_solve(s.get(), d, max_iters, tol, WLS_REFERENCE, warm, u, nullptr);
    const double true_rel = (w > 1 && h > 1) ? wls_true_rel(s.get(), d, u) : 0.0;
    LSH_CHECK(hipStreamSynchronize(stream()));
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (rep) {
        rep->iters = r.iters;
        rep->outer = r.outer;
        rep->rel = r.rel;
        rep->true_rel = true_rel;
        rep->err_fp64 = -1.0;
        rep->ms = ms;
    }
    LSH_API_END
}

}  // extern "C"
