// depth.hip -- depth from focus on gfx950 (depth.rs): the focus slice of each
// frame, guided-filter aggregation fused with the streamed peak tracker, the
// exact order statistics (noise floor, p90), the 3x3 median, the edge-aware WLS
// solve (FGS guess, then CG preconditioned by a multigrid V-cycle, Huber
// reweighting) and the guided upsampling, behind lsh_depth_from_slices.
//
// Box filters sum their window in f64 directly (the CPU path keeps an f64
// running sum; both give the exactly rounded window sum whenever the f64 sums
// are exact, which they are for these planes), then round to f32 and scale by
// the f32 reciprocal count, as the CPU path does.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <string>

#include "api.hpp"
#include "common.hpp"
#include "engine.hpp"
#include "wls.hpp"

namespace lsh {

// ---------------------------------------------------------------- focus slice
struct Taps {
    int n, halo, kind, step;
};
// The ring-difference taps as the kernel reads them: offsets into the staged
// luma window (dy * pitch + dx) and weights, in the reference's order
// (r_out <= 7: at most 15 x 15 taps); passed by value, so every lane reads
// them through the scalar unit.
#define MAX_TAPS 225
struct TapTable {
    int off[MAX_TAPS];
    float wt[MAX_TAPS];
};

// The ring difference filter of the default measure, RDF(r_in 1, r_out 3), at
// scale 1 (2 x 2 pixels a cell) with its rows as compile-time constants: in
// the reference's order (depth.rs:417-451) the taps are 7 rows, each its disk
// taps (dx = -D..D, weight wd) then its ring taps (the rest of dx = -M..M,
// weight wr), both ascending.  The host checks the order has this form before
// it selects this path.  A row's window samples for both pixels of a cell are
// read from LDS once (6 or 8 instead of 2 x (2D+1 + ring)), and each pixel's
// sum takes its taps in the reference's order.
template <int D, int M>
__device__ __forceinline__ void rdf13_row(const float *__restrict__ row, float wd, float wr, float (&a)[2]) {
    float seg[2 + 2 * M];  // seg[M + dx + j]: the sample at dx of pixel j
#pragma unroll
    for (int q = 0; q < 2 + 2 * M; q++) seg[q] = row[q - M];
#pragma unroll
    for (int dx = -D; dx <= D; dx++)
#pragma unroll
        for (int j = 0; j < 2; j++) a[j] += wd * seg[M + dx + j];
#pragma unroll
    for (int dx = -M; dx <= M; dx++)
        if (dx < -D || dx > D)
#pragma unroll
            for (int j = 0; j < 2; j++) a[j] += wr * seg[M + dx + j];
}

// the cell's two pixels of window row ly from column lx: |RDF| summed in x order
__device__ __forceinline__ float rdf13_pair(const float *__restrict__ Y, int PW, int ly, int lx, float wd, float wr) {
    float a[2] = {0.0f, 0.0f};
    rdf13_row<0, 2>(Y + (ly - 1) * PW + lx, wd, wr, a);
    rdf13_row<1, 3>(Y + ly * PW + lx, wd, wr, a);
    rdf13_row<0, 2>(Y + (ly + 1) * PW + lx, wd, wr, a);
    rdf13_row<-1, 0>(Y + (ly - 3) * PW + lx, wd, wr, a);
    rdf13_row<-1, 2>(Y + (ly - 2) * PW + lx, wd, wr, a);
    rdf13_row<-1, 2>(Y + (ly + 2) * PW + lx, wd, wr, a);
    rdf13_row<-1, 0>(Y + (ly + 3) * PW + lx, wd, wr, a);
    float rs = 0.0f;
    rs += fabsf(a[0]);
    rs += fabsf(a[1]);
    return rs;
}

// depth.rs:417-495 + :174-197: luma, ring difference filter (or SML), |.|, k x k
// block mean, one output cell per thread, the luma tile staged in LDS.
// depth.rs:1097-1100: the focus measure of the luma (depth.rs:160-165) block-
// averaged on the working grid.  A block covers 64 x (4 CY) grid cells: the
// luma of its pixels and their halo is staged in LDS once (reflect-101 at the
// borders), then each thread averages the measure over CY cells of its column.
// ONE: the input is the luma plane itself (written with the same expression by
// the warp that made the planes), not the three planes.  R13: the measure is
// RDF(1, 3) at scale 1, whose whole cells go through rdf13_pair.
template <bool ONE, bool R13>
__global__ __launch_bounds__(256) void k_focus_slice(const float *__restrict__ planes, int w, int h, int k, int CY,
                                                     const TapTable tab, Taps T, float wd, float wr,
                                                     float *__restrict__ out, int dw, int dh,
                                                     const unsigned *__restrict__ out_idx,
                                                     float *const *__restrict__ out_base) {
    if (out_base) out = out_base[0];
    if (out_idx) out += (size_t)out_idx[0] * dw * dh;
    extern __shared__ float Y[];
    const int TX = 64, TY = 4 * CY, R = T.halo;
    const int PW = TX * k + 2 * R, PH = TY * k + 2 * R;
    const int px0 = blockIdx.x * TX * k - R, py0 = blockIdx.y * TY * k - R;
    const size_t n = (size_t)w * h;
    const int tid = threadIdx.y * TX + threadIdx.x;
    // the window's size is a run-time value: loads go out eight per thread at a
    // time, each group's stores after all of its loads
    constexpr int U = 8;
    const int NQ = PW * PH;
    for (int q0 = tid; q0 < NQ; q0 += U * 256) {
        float v[U][3];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int q = q0 + 256 * u;
            if (q < NQ) {
                const int ly = q / PW, lx = q - ly * PW;
                const size_t o = (size_t)rf101(py0 + ly, h) * w + rf101(px0 + lx, w);
                v[u][0] = planes[o];
                if (!ONE) {
                    v[u][1] = planes[n + o];
                    v[u][2] = planes[2 * n + o];
                }
            }
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int q = q0 + 256 * u;
            if (q < NQ)
                Y[q] = ONE ? v[u][0] : 0.299f * v[u][0] + 0.587f * v[u][1] + 0.114f * v[u][2];  // depth.rs:160-165
        }
    }
    __syncthreads();
    const int ox = blockIdx.x * TX + threadIdx.x;
    if (ox >= dw) return;
    for (int c = 0; c < CY; c++) {
        const int oy = blockIdx.y * TY + threadIdx.y + 4 * c;
        if (oy >= dh) return;
        int x0 = ox * k, y0 = oy * k, x1 = min(x0 + k, w), y1 = min(y0 + k, h);
        float o = 0.0f;
        if (R13 && x1 - x0 == 2 && y1 - y0 == 2) {
            const int lx = x0 - px0, ly = y0 - py0;
            o += rdf13_pair(Y, PW, ly, lx, wd, wr);
            o += rdf13_pair(Y, PW, ly + 1, lx, wd, wr);
            out[(size_t)oy * dw + ox] = o / 4.0f;
            continue;
        }
        for (int y = y0; y < y1; y++) {
            float rs = 0.0f;
            int ly = y - py0;
            for (int x = x0; x < x1; x++) {
                int lx = x - px0;
                float f;
                if (T.kind == 0) {
                    const float *c0 = Y + ly * PW + lx;
                    float a = 0.0f;
                    for (int t = 0; t < T.n; t++) a += tab.wt[t] * c0[tab.off[t]];
                    f = fabsf(a);
                } else {
                    int s = T.step;
                    float cc = Y[ly * PW + lx];
                    f = fabsf(2.0f * cc - Y[ly * PW + lx - s] - Y[ly * PW + lx + s]) +
                        fabsf(2.0f * cc - Y[(ly - s) * PW + lx] - Y[(ly + s) * PW + lx]);
                }
                rs += f;
            }
            o += rs;
        }
        out[(size_t)oy * dw + ox] = o / (float)((y1 - y0) * (x1 - x0));
    }
}

// depth.rs:454-471 with the sparse convolution's order (depth.rs:417-451): rows
// grouped by dy in order of first appearance, each row's taps in list order.
static void rdf_taps_ordered(uint32_t r_in, uint32_t r_out, std::vector<int32_t> &dy, std::vector<int32_t> &dx,
                             std::vector<float> &wt) {
    long ri2 = (long)r_in * r_in, ro2 = (long)r_out * r_out, r = (long)r_out;
    std::vector<std::pair<int, int>> disk, ring;
    for (long y = -r; y <= r; y++)
        for (long x = -r; x <= r; x++) {
            long d2 = x * x + y * y;
            if (d2 <= ri2)
                disk.push_back({(int)y, (int)x});
            else if (d2 <= ro2)
                ring.push_back({(int)y, (int)x});
        }
    float wd = 1.0f / (float)disk.size(), wr = -1.0f / (float)std::max<size_t>(ring.size(), 1);
    std::vector<std::tuple<int, int, float>> taps;
    for (auto &p : disk) taps.emplace_back(p.first, p.second, wd);
    for (auto &p : ring) taps.emplace_back(p.first, p.second, wr);
    std::vector<int> rows;
    for (auto &t : taps)
        if (std::find(rows.begin(), rows.end(), std::get<0>(t)) == rows.end()) rows.push_back(std::get<0>(t));
    dy.clear();
    dx.clear();
    wt.clear();
    for (int ry : rows)
        for (auto &t : taps)
            if (std::get<0>(t) == ry) {
                dy.push_back(std::get<0>(t));
                dx.push_back(std::get<1>(t));
                wt.push_back(std::get<2>(t));
            }
}

// Whether RDF(1, 3)'s taps in the reference's order are the rows rdf13_pair
// takes (dy, disk half-width D or -1, ring half-width M), each row its disk taps
// then its ring taps, both ascending, with one disk and one ring weight.
static bool rdf13_order_holds() {
    static const int rows[7][3] = {{-1, 0, 2}, {0, 1, 3}, {1, 0, 2}, {-3, -1, 0}, {-2, -1, 2}, {2, -1, 2}, {3, -1, 0}};
    std::vector<int32_t> dy, dx;
    std::vector<float> wt;
    rdf_taps_ordered(1, 3, dy, dx, wt);
    size_t t = 0;
    for (const auto &r : rows) {
        for (int x = -r[1]; x <= r[1]; x++, t++)
            if (t >= wt.size() || dy[t] != r[0] || dx[t] != x || wt[t] != wt[0]) return false;
        for (int x = -r[2]; x <= r[2]; x++)
            if (x < -r[1] || x > r[1]) {
                if (t >= wt.size() || dy[t] != r[0] || dx[t] != x || wt[t] != wt.back()) return false;
                t++;
            }
    }
    return t == wt.size();
}

// The focus slice of planes (w x h) with a holder of the measure's taps
// (uploaded when the measure changes).
static void focus_slice_with(DBuf<int32_t> &taps_dy, DBuf<int32_t> &taps_dx, DBuf<float> &taps_wt, uint32_t &ntaps,
                             uint32_t &taps_key, size_t w, size_t h, const float *planes, const lsh_depth_params &p,
                             float *out, const unsigned *out_idx = nullptr, bool luma = false,
                             float *const *out_base = nullptr);

void focus_slice(lsh_frames *f, const float *planes, const lsh_depth_params &p, float *out, const unsigned *out_idx,
                 bool luma, float *const *out_base) {
    focus_slice_with(f->taps_dy, f->taps_dx, f->taps_wt, f->ntaps, f->taps_key, f->w, f->h, planes, p, out, out_idx,
                     luma, out_base);
}

static void focus_slice_with(DBuf<int32_t> &taps_dy, DBuf<int32_t> &taps_dx, DBuf<float> &taps_wt, uint32_t &ntaps,
                             uint32_t &taps_key, size_t w, size_t h, const float *planes, const lsh_depth_params &p,
                             float *out, const unsigned *out_idx, bool luma, float *const *out_base) {
    LSH_REQUIRE(p.scale <= 2, "focus slice: scale above 2 is not supported");
    LSH_REQUIRE(p.focus_kind <= 1, "focus slice: unknown focus measure");
    int k = 1 << p.scale;
    uint32_t key = p.focus_kind == 0 ? (p.r_in << 8 | p.r_out) : (0x10000u | p.sml_step);
    if (key != taps_key) {
        std::vector<int32_t> dy, dx;
        std::vector<float> wt;
        if (p.focus_kind == 0) {
            LSH_REQUIRE(p.r_out > p.r_in && p.r_out <= 7, "focus slice: ring radii out of range");
            rdf_taps_ordered(p.r_in, p.r_out, dy, dx, wt);
        } else {
            LSH_REQUIRE(p.sml_step <= 7, "focus slice: SML step out of range");
            dy = {0};
            dx = {0};
            wt = {0.f};
        }
        taps_dy.alloc(dy.size());
        taps_dx.alloc(dx.size());
        taps_wt.alloc(wt.size());
        copy_mem(taps_dy.p, LSH_DEVICE, dy.data(), LSH_HOST, dy.size() * sizeof(int32_t));
        copy_mem(taps_dx.p, LSH_DEVICE, dx.data(), LSH_HOST, dx.size() * sizeof(int32_t));
        copy_mem(taps_wt.p, LSH_DEVICE, wt.data(), LSH_HOST, wt.size() * sizeof(float));
        ntaps = (uint32_t)wt.size();
        taps_key = key;
    }
    Taps T;
    T.kind = (int)p.focus_kind;
    T.n = (int)ntaps;
    T.step = (int)std::max<uint32_t>(p.sml_step, 1);
    T.halo = p.focus_kind == 0 ? (int)p.r_out : T.step;
    int dw = (int)((w + k - 1) / k), dh = (int)((h + k - 1) / k);
    // cells per thread: as many (up to 4) as keep the staged window within 32 KB
    const int PW = 64 * k + 2 * T.halo;
    int CY = 4;
    while (CY > 1 && (size_t)PW * (4 * CY * k + 2 * T.halo) * sizeof(float) > 32768) CY--;
    size_t shm = (size_t)PW * (4 * CY * k + 2 * T.halo) * sizeof(float);
    TapTable tab{};
    if (p.focus_kind == 0) {  // the same taps as uploaded, as offsets into the window of pitch PW
        std::vector<int32_t> dy, dx;
        std::vector<float> wt;
        rdf_taps_ordered(p.r_in, p.r_out, dy, dx, wt);
        LSH_REQUIRE(wt.size() <= MAX_TAPS && wt.size() == ntaps, "focus slice: tap table");
        for (size_t t = 0; t < wt.size(); t++) {
            tab.off[t] = dy[t] * PW + dx[t];
            tab.wt[t] = wt[t];
        }
    }
    // the specialised RDF(1, 3) at scale 1, when the reference's order has the form it assumes
    const bool r13 = p.focus_kind == 0 && p.r_in == 1 && p.r_out == 3 && k == 2 && rdf13_order_holds();
    float wd = 0.0f, wr = 0.0f;
    if (r13) {
        wd = tab.wt[0];             // the first tap is a disk tap
        wr = tab.wt[ntaps - 1];     // the last a ring tap
    }
    const double by = (luma ? 4.0 : 12.0) * w * h + 4.0 * dw * dh;
    const dim3 grid(cdiv(dw, 64), cdiv(dh, 4 * CY)), blk(64, 4);
    const float *src = planes;
    if (luma && r13)
        launch_b("focus_slice", by, k_focus_slice<true, true>, grid, blk, shm, src, (int)w, (int)h, k, CY, tab, T, wd, wr, out, dw, dh, out_idx, out_base);
    else if (luma)
        launch_b("focus_slice", by, k_focus_slice<true, false>, grid, blk, shm, src, (int)w, (int)h, k, CY, tab, T, wd, wr, out, dw, dh, out_idx, out_base);
    else if (r13)
        launch_b("focus_slice", by, k_focus_slice<false, true>, grid, blk, shm, src, (int)w, (int)h, k, CY, tab, T, wd, wr, out, dw, dh, out_idx, out_base);
    else
        launch_b("focus_slice", by, k_focus_slice<false, false>, grid, blk, shm, src, (int)w, (int)h, k, CY, tab, T, wd, wr, out, dw, dh, out_idx, out_base);
}

// ----------------------------------------------------------------- helpers
__global__ void k_luma3(const float *__restrict__ planes, size_t n, float *__restrict__ y) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    y[i] = 0.299f * planes[i] + 0.587f * planes[n + i] + 0.114f * planes[2 * n + i];
}

// depth.rs:174-197
__global__ void k_block_mean(const float *__restrict__ src, int w, int h, int k, float *__restrict__ out, int dw,
                             int dh) {
    int ox = blockIdx.x * blockDim.x + threadIdx.x, oy = blockIdx.y * blockDim.y + threadIdx.y;
    if (ox >= dw || oy >= dh) return;
    int x0 = ox * k, y0 = oy * k, x1 = min(x0 + k, w), y1 = min(y0 + k, h);
    float o = 0.0f;
    for (int y = y0; y < y1; y++) {
        float rs = 0.0f;
        for (int x = x0; x < x1; x++) rs += src[(size_t)y * w + x];
        o += rs;
    }
    out[(size_t)oy * dw + ox] = o / (float)((y1 - y0) * (x1 - x0));
}

// depth.rs:201-221
__device__ __forceinline__ float bilinear_at(const float *__restrict__ g, int dw, int dh, int x, int y, float inv) {
    float fy = rclampf(((float)y + 0.5f) * inv - 0.5f, 0.0f, (float)(dh - 1));
    int y0 = (int)fy, y1 = min(y0 + 1, dh - 1);
    float ty = fy - (float)y0;
    float fx = rclampf(((float)x + 0.5f) * inv - 0.5f, 0.0f, (float)(dw - 1));
    int x0 = (int)fx, x1 = min(x0 + 1, dw - 1);
    float tx = fx - (float)x0;
    const float *r0 = g + (size_t)y0 * dw, *r1 = g + (size_t)y1 * dw;
    float top = r0[x0] + (r0[x1] - r0[x0]) * tx;
    float bot = r1[x0] + (r1[x1] - r1[x0]) * tx;
    return top + (bot - top) * ty;
}

__global__ void k_up_bilinear(const float *__restrict__ g, int dw, int dh, int w, int h, float inv, float lo, float hi,
                              int clamp, float *__restrict__ out) {
    int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= w || y >= h) return;
    float v = bilinear_at(g, dw, dh, x, y, inv);
    if (clamp) v = rclampf(v, lo, hi);
    out[(size_t)y * w + x] = v;
}

__global__ void k_clamp(float *__restrict__ v, size_t n, float lo, float hi) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = rclampf(v[i], lo, hi);
}

// depth.rs:1187-1196: a*Y + b with both coefficients bilinear from the grid, clamped
__global__ void k_up_guided(const float *__restrict__ a, const float *__restrict__ b, int dw, int dh,
                            const float *__restrict__ yfull, int w, int h, float inv, float maxd,
                            float *__restrict__ out) {
    int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= w || y >= h) return;
    size_t o = (size_t)y * w + x;
    float av = bilinear_at(a, dw, dh, x, y, inv), bv = bilinear_at(b, dw, dh, x, y, inv);
    out[o] = rclampf(av * yfull[o] + bv, 0.0f, maxd);
}

// -------------------------------------------------------------- box filters
// Horizontal window sums of two planes in f64, rounded to f32.  src1 is a
// product plane when mul is set: src1[i] = g[i] * src0[i] (or g[i]*g[i] when
// src0 is the guide itself).
__global__ __launch_bounds__(256) void k_box_h2_direct(const float *__restrict__ a, const float *__restrict__ b,
                                                const float *__restrict__ g, int mode, int w, int h, int r,
                                                float *__restrict__ oa, float *__restrict__ ob) {
    int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= w || y >= h) return;
    const size_t row = (size_t)y * w;
    int lo = max(x - r, 0), hi = min(x + r, w - 1);
    double sa = 0.0, sb = 0.0;
    for (int i = lo; i <= hi; i++) {
        float va = a[row + i];
        float vb = mode == 0 ? b[row + i] : (mode == 1 ? g[row + i] * va : va * va);
        sa += (double)va;
        sb += (double)vb;
    }
    oa[row + x] = (float)sa;
    ob[row + x] = (float)sb;
}

// reciprocal of the border-clipped window count (depth.rs:273-280)
__device__ __forceinline__ float inv_count(int x, int y, int w, int h, int r) {
    float cy = (float)(min(y + r + 1, h) - max(y - r, 0));
    float cx = (float)(min(x + r + 1, w) - max(x - r, 0));
    return 1.0f / (cy * cx);
}

// The per-cell peak tracker state (depth.rs:499-531).
struct Peaks {
    float *c1, *l1, *r1, *c2, *prev, *prev2, *sum, *cmin;
    uint32_t *i1;
};

__device__ __forceinline__ void px_register(const Peaks &P, size_t i, float val, uint32_t idx, float l, float r,
                                            float &c1, float &c2) {
    if (val > c1) {
        if (c1 >= 0.0f) c2 = c1;
        c1 = val;
        P.i1[i] = idx;
        P.l1[i] = l;
        P.r1[i] = r;
    } else if (val > c2) {
        c2 = val;
    }
}

// depth.rs:547-564 for one cell and value v (slice m)
__device__ __forceinline__ void peak_push(const Peaks &P, size_t i, float v, uint32_t m) {
    float prev = P.prev[i], prev2 = P.prev2[i];
    if (m >= 1) {
        bool is_peak = (m == 1 || prev >= prev2) && prev > v;
        if (is_peak) {
            float c1 = P.c1[i], c2 = P.c2[i];
            float l = m >= 2 ? prev2 : -1.0f;
            px_register(P, i, prev, m - 1, l, v, c1, c2);
            P.c1[i] = c1;
            P.c2[i] = c2;
        }
    }
    P.sum[i] += v;
    P.cmin[i] = fminf(P.cmin[i], v);
    P.prev2[i] = prev;
    P.prev[i] = v;
}

// peak_push of slice m's value v1 then slice m + 1's v2, the state kept in
// registers between the two (the same operations in the same order)
__device__ __forceinline__ void peak_push2(const Peaks &P, size_t i, float v1, float v2, uint32_t m) {
    float prev = P.prev[i], prev2 = P.prev2[i], sum = P.sum[i], cmin = P.cmin[i];
    float c1 = 0.0f, c2 = 0.0f;
    bool have = false;
    auto one = [&](float v, uint32_t mm) {
        if (mm >= 1 && (mm == 1 || prev >= prev2) && prev > v) {
            if (!have) {
                c1 = P.c1[i];
                c2 = P.c2[i];
                have = true;
            }
            px_register(P, i, prev, mm - 1, mm >= 2 ? prev2 : -1.0f, v, c1, c2);
        }
        sum += v;
        cmin = fminf(cmin, v);
        prev2 = prev;
        prev = v;
    };
    one(v1, m);
    one(v2, m + 1);
    if (have) {
        P.c1[i] = c1;
        P.c2[i] = c2;
    }
    P.sum[i] = sum;
    P.cmin[i] = cmin;
    P.prev2[i] = prev2;
    P.prev[i] = prev;
}

// Vertical window sums of two planes, then per mode:
//  0  guide statistics: oa = mean(I), ob = var(I) = max(mean(I^2) - mean(I)^2, 0)
//  1  guided filter step 1 (depth.rs:382-389): oa = a, ob = b from mean(p), mean(I p)
//  2  guided filter step 2 fused with the peak push: q = mean(a) I + mean(b), max(q, 0), push
//  3  guided coefficients: oa = mean(a), ob = mean(b)
//  4  guided filter output: oa = mean(a) I + mean(b)
// The per-output tail of the vertical pass (sa, sb the window sums at (x, y)).
__device__ __forceinline__ void box_v_tail(int x, int y, int w, int h, int r, int mode, double sa, double sb,
                                           const float *__restrict__ guide, const float *__restrict__ mean_i,
                                           const float *__restrict__ var_i, float eps, float *__restrict__ oa,
                                           float *__restrict__ ob, const Peaks &P, uint32_t m) {
    float ic = inv_count(x, y, w, h, r);
    float ma = (float)sa * ic, mb = (float)sb * ic;
    size_t o = (size_t)y * w + x;
    if (mode == 0) {
        oa[o] = ma;
        ob[o] = fmaxf(mb - ma * ma, 0.0f);
    } else if (mode == 1) {
        float cov = mb - mean_i[o] * ma;
        float a = cov / (var_i[o] + eps);
        oa[o] = a;
        ob[o] = ma - a * mean_i[o];
    } else if (mode == 2) {
        float q = ma * guide[o] + mb;
        peak_push(P, o, fmaxf(q, 0.0f), m);
    } else if (mode == 3) {
        oa[o] = ma;
        ob[o] = mb;
    } else {
        oa[o] = ma * guide[o] + mb;
    }
}

__global__ __launch_bounds__(256) void k_box_v2_direct(const float *__restrict__ ta, const float *__restrict__ tb,
                                                       int w, int h, int r, int mode, const float *__restrict__ guide,
                                                       const float *__restrict__ mean_i,
                                                       const float *__restrict__ var_i, float eps,
                                                       float *__restrict__ oa, float *__restrict__ ob, Peaks P,
                                                       uint32_t m) {
    int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= w || y >= h) return;
    int lo = max(y - r, 0), hi = min(y + r, h - 1);
    double sa = 0.0, sb = 0.0;
    for (int i = lo; i <= hi; i++) {
        sa += (double)ta[(size_t)i * w + x];
        sb += (double)tb[(size_t)i * w + x];
    }
    box_v_tail(x, y, w, h, r, mode, sa, sb, guide, mean_i, var_i, eps, oa, ob, P, m);
}

// The same two passes on tiles staged in LDS: BH_RB row segments of BH_TX
// outputs and their radius-r halo (horizontal), or BV_TX columns of BV_TY outputs and
// their 2r halo rows (vertical), loaded once with every load of a thread's
// chunk issued before it is stored.  Each thread then takes a run of
// consecutive outputs along the pass (BH_RUN in a row, BV_RUN in a column):
// the first window summed in f64 from its start, each next one by adding the
// sample that enters and subtracting the one that leaves, as the reference's
// running sum does (depth.rs:293-331).  The f64 sums of f32 samples are exact
// (see box_h2), so every window's value is the direct kernels' to the bit.
#define BH_TX 256
#define BH_RB 4   // rows per horizontal tile
#define BH_RUN 4  // outputs per thread along a row: BH_RB * BH_TX / BH_RUN = 256 threads
#define BV_TX 64
#define BV_TY 32
#define BV_RUN 8  // outputs per thread down a column: BV_TX * BV_TY / BV_RUN = 256 threads
#define BOX_TILE_MAX_R 32
__global__ __launch_bounds__(256) void k_box_h2_tile(const float *__restrict__ a, const float *__restrict__ b,
                                                     const float *__restrict__ g, int mode, int w, int h, int r,
                                                     float *__restrict__ oa, float *__restrict__ ob) {
    extern __shared__ float S[];  // [2][BH_RB][BH_TX + 2 r]
    const int L = BH_TX + 2 * r;
    float *Sa = S, *Sb = S + BH_RB * L;
    const int x0 = blockIdx.x * BH_TX, y0 = blockIdx.y * BH_RB, nr = min(BH_RB, h - y0);
    const int c0 = max(x0 - r, 0), nc = min(x0 + BH_TX - 1 + r, w - 1) - c0 + 1, n = nr * nc;
    for (int k0 = threadIdx.x; k0 < n; k0 += 4 * BH_TX) {
        float va[4], vb[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int k = k0 + u * BH_TX;
            if (k < n) {
                const int j = k / nc;
                const size_t i = (size_t)(y0 + j) * w + c0 + (k - j * nc);
                va[u] = a[i];
                vb[u] = mode == 0 ? b[i] : (mode == 1 ? g[i] * va[u] : va[u] * va[u]);
            }
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int k = k0 + u * BH_TX;
            if (k < n) {
                const int j = k / nc, q = k - j * nc;
                Sa[j * L + q] = va[u];
                Sb[j * L + q] = vb[u];
            }
        }
    }
    __syncthreads();
    constexpr int TPR = BH_TX / BH_RUN;  // threads per row
    const int j = threadIdx.x / TPR, xs = x0 + (threadIdx.x % TPR) * BH_RUN;
    if (j >= nr || xs >= w) return;
    const float *ra = Sa + j * L - c0, *rb = Sb + j * L - c0;  // indexed by the column
    double sa = 0.0, sb = 0.0;
    for (int i = max(xs - r, 0); i <= min(xs + r, w - 1); i++) {
        sa += (double)ra[i];
        sb += (double)rb[i];
    }
    float va[BH_RUN], vb[BH_RUN];
    va[0] = (float)sa;
    vb[0] = (float)sb;
#pragma unroll
    for (int u = 1; u < BH_RUN; u++) {
        const int x = xs + u;
        if (x + r <= w - 1) {
            sa += (double)ra[x + r];
            sb += (double)rb[x + r];
        }
        if (x - r - 1 >= 0) {
            sa -= (double)ra[x - r - 1];
            sb -= (double)rb[x - r - 1];
        }
        va[u] = (float)sa;
        vb[u] = (float)sb;
    }
    const size_t o = (size_t)(y0 + j) * w + xs;
    if (xs + BH_RUN <= w && (((uintptr_t)(oa + o) | (uintptr_t)(ob + o)) & 15) == 0) {
        *(float4 *)(oa + o) = make_float4(va[0], va[1], va[2], va[3]);
        *(float4 *)(ob + o) = make_float4(vb[0], vb[1], vb[2], vb[3]);
    } else {
        for (int u = 0; u < BH_RUN && xs + u < w; u++) {
            oa[o + u] = va[u];
            ob[o + u] = vb[u];
        }
    }
}

__global__ __launch_bounds__(256) void k_box_v2_tile(const float *__restrict__ ta, const float *__restrict__ tb,
                                                     int w, int h, int r, int mode, const float *__restrict__ guide,
                                                     const float *__restrict__ mean_i, const float *__restrict__ var_i,
                                                     float eps, float *__restrict__ oa, float *__restrict__ ob, Peaks P,
                                                     uint32_t m) {
    extern __shared__ float S[];  // [2][BV_TY + 2 r][BV_TX]
    float *Sa = S, *Sb = S + (BV_TY + 2 * r) * BV_TX;
    const int x0 = blockIdx.x * BV_TX, y0 = blockIdx.y * BV_TY;
    const int tx = threadIdx.x & (BV_TX - 1), ty = threadIdx.x / BV_TX, x = x0 + tx;
    const int r0 = max(y0 - r, 0), nr = min(y0 + BV_TY - 1 + r, h - 1) - r0 + 1;
    if (x < w)
        for (int q0 = ty; q0 < nr; q0 += 4 * 4) {
            float va[4], vb[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int q = q0 + 4 * u;
                if (q < nr) {
                    const size_t o = (size_t)(r0 + q) * w + x;
                    va[u] = ta[o];
                    vb[u] = tb[o];
                }
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int q = q0 + 4 * u;
                if (q < nr) {
                    Sa[q * BV_TX + tx] = va[u];
                    Sb[q * BV_TX + tx] = vb[u];
                }
            }
        }
    __syncthreads();
    const int ys = y0 + ty * BV_RUN;
    if (x >= w || ys >= h) return;
    const float *ca = Sa + tx - r0 * BV_TX, *cb = Sb + tx - r0 * BV_TX;  // indexed by the row
    double sa = 0.0, sb = 0.0;
    for (int i = max(ys - r, 0); i <= min(ys + r, h - 1); i++) {
        sa += (double)ca[i * BV_TX];
        sb += (double)cb[i * BV_TX];
    }
    box_v_tail(x, ys, w, h, r, mode, sa, sb, guide, mean_i, var_i, eps, oa, ob, P, m);
    for (int u = 1; u < BV_RUN && ys + u < h; u++) {
        const int y = ys + u;
        if (y + r <= h - 1) {
            sa += (double)ca[(y + r) * BV_TX];
            sb += (double)cb[(y + r) * BV_TX];
        }
        if (y - r - 1 >= 0) {
            sa -= (double)ca[(y - r - 1) * BV_TX];
            sb -= (double)cb[(y - r - 1) * BV_TX];
        }
        box_v_tail(x, y, w, h, r, mode, sa, sb, guide, mean_i, var_i, eps, oa, ob, P, m);
    }
}

// The guided filter's last vertical pass (mode 2 of k_box_v2_tile) for two
// slices at once, m (ta, tb) and m + 1 (tc, td), fused with both peak pushes,
// so the peak state is read and written once per pair of slices
__global__ __launch_bounds__(256) void k_box_v4_peaks_tile(const float *__restrict__ ta, const float *__restrict__ tb,
                                                           const float *__restrict__ tc, const float *__restrict__ td,
                                                           int w, int h, int r, const float *__restrict__ guide,
                                                           Peaks P, uint32_t m) {
    extern __shared__ float S[];  // [4][BV_TY + 2 r][BV_TX]
    const int RS = (BV_TY + 2 * r) * BV_TX;
    float *Sp[4] = {S, S + RS, S + 2 * RS, S + 3 * RS};
    const float *src[4] = {ta, tb, tc, td};
    const int x0 = blockIdx.x * BV_TX, y0 = blockIdx.y * BV_TY;
    const int tx = threadIdx.x & (BV_TX - 1), ty = threadIdx.x / BV_TX, x = x0 + tx;
    const int r0 = max(y0 - r, 0), nr = min(y0 + BV_TY - 1 + r, h - 1) - r0 + 1;
    if (x < w)
        for (int q0 = ty; q0 < nr; q0 += 4 * 2) {
            float v[2][4];
#pragma unroll
            for (int u = 0; u < 2; u++) {
                const int q = q0 + 4 * u;
                if (q < nr) {
                    const size_t o = (size_t)(r0 + q) * w + x;
#pragma unroll
                    for (int c = 0; c < 4; c++) v[u][c] = src[c][o];
                }
            }
#pragma unroll
            for (int u = 0; u < 2; u++) {
                const int q = q0 + 4 * u;
                if (q < nr)
#pragma unroll
                    for (int c = 0; c < 4; c++) Sp[c][q * BV_TX + tx] = v[u][c];
            }
        }
    __syncthreads();
    const int ys = y0 + ty * BV_RUN;
    if (x >= w || ys >= h) return;
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i = max(ys - r, 0); i <= min(ys + r, h - 1); i++)
#pragma unroll
        for (int c = 0; c < 4; c++) s[c] += (double)Sp[c][(i - r0) * BV_TX + tx];
    for (int u = 0; u < BV_RUN && ys + u < h; u++) {
        const int y = ys + u;
        if (u > 0) {
            if (y + r <= h - 1)
#pragma unroll
                for (int c = 0; c < 4; c++) s[c] += (double)Sp[c][(y + r - r0) * BV_TX + tx];
            if (y - r - 1 >= 0)
#pragma unroll
                for (int c = 0; c < 4; c++) s[c] -= (double)Sp[c][(y - r - 1 - r0) * BV_TX + tx];
        }
        const float ic = inv_count(x, y, w, h, r);
        const size_t o = (size_t)y * w + x;
        const float g = guide[o];
        const float q1 = ((float)s[0] * ic) * g + (float)s[1] * ic;
        const float q2 = ((float)s[2] * ic) * g + (float)s[3] * ic;
        peak_push2(P, o, fmaxf(q1, 0.0f), fmaxf(q2, 0.0f), m);
    }
}

__global__ void k_peak_push_raw(const float *__restrict__ c, size_t n, Peaks P, uint32_t m) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) peak_push(P, i, c[i], m);
}

__global__ void k_peak_init(Peaks P, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    P.c1[i] = P.l1[i] = P.r1[i] = P.c2[i] = -1.0f;
    P.prev[i] = P.prev2[i] = P.sum[i] = 0.0f;
    P.cmin[i] = INFINITY;
    P.i1[i] = 0;
}

// depth.rs:571-576
__global__ void k_peak_close(Peaks P, size_t n, uint32_t ns) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float prev = P.prev[i], prev2 = P.prev2[i];
    if (ns == 1 || prev >= prev2) {
        float c1 = P.c1[i], c2 = P.c2[i];
        float l = ns >= 2 ? prev2 : -1.0f;
        px_register(P, i, prev, ns - 1, l, -1.0f, c1, c2);
        P.c1[i] = c1;
        P.c2[i] = c2;
    }
}

// depth.rs:581-609
__global__ void k_peak_out(Peaks P, size_t n, uint32_t ns, float gate, const float *__restrict__ floor_p,
                           float *__restrict__ depth, float *__restrict__ conf, float *__restrict__ mins) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float floor_ = *floor_p;
    float inv_n = 1.0f / (float)ns;
    float c1 = P.c1[i], l = P.l1[i], r = P.r1[i];
    float delta = 0.0f;
    if (l >= 0.0f && r >= 0.0f && c1 > 0.0f) {
        float ll = logf(fmaxf(l, 1e-12f)), lc = logf(c1), lr = logf(fmaxf(r, 1e-12f));
        float den = ll - 2.0f * lc + lr;
        if (den < 0.0f) delta = rclampf(0.5f * (ll - lr) / den, -0.5f, 0.5f);
    }
    depth[i] = (float)P.i1[i] + delta;
    float cv = 0.0f;
    if (c1 > 0.0f) {
        float mean = P.sum[i] * inv_n;
        float prom = rclampf(1.0f - mean / c1, 0.0f, 1.0f);
        float c2 = P.c2[i];
        float pkr = c2 >= 0.0f ? rclampf(1.0f - c2 / c1, 0.0f, 1.0f) : 1.0f;
        float g = (gate > 0.0f && floor_ > 0.0f) ? rclampf((c1 - floor_) / (gate * floor_), 0.0f, 1.0f) : 1.0f;
        cv = prom * pkr * g;
    }
    conf[i] = cv;
    if (mins) {
        float cm = P.cmin[i];
        mins[i] = isfinite(cm) ? cm : 0.0f;
    }
}

// ------------------------------------------------- exact order statistics
// The k-th smallest (f32::total_cmp order) of every 7th element, by a
// three-pass radix select over the order-preserving 32-bit key.
__device__ __forceinline__ uint32_t okey(float v) {
    uint32_t b = __float_as_uint(v);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

__global__ void k_radix_hist(const float *__restrict__ v, size_t n, size_t stride, uint32_t prefix, int pbits,
                             int shift, int bits, unsigned *__restrict__ hist) {
    __shared__ unsigned sh[4096];
    int nb = 1 << bits;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) sh[i] = 0;
    __syncthreads();
    size_t ns = (n + stride - 1) / stride;
    for (size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x; s < ns; s += (size_t)gridDim.x * blockDim.x) {
        uint32_t key = okey(v[s * stride]);
        if (pbits > 0 && (key >> (32 - pbits)) != prefix) continue;
        atomicAdd(&sh[(key >> shift) & (nb - 1)], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nb; i += blockDim.x)
        if (sh[i]) atomicAdd(&hist[i], sh[i]);
}

static float select_kth(const float *d, size_t n, size_t stride, size_t k, DBuf<unsigned> &hist) {
    size_t ns = (n + stride - 1) / stride;
    LSH_REQUIRE(k < ns, "select: rank out of range");
    hist.ensure(4096);
    std::vector<unsigned> hh(4096);
    uint32_t prefix = 0;
    int pbits = 0;
    const int plan[3][2] = {{20, 12}, {8, 12}, {0, 8}};  // (shift, bits)
    unsigned grid = std::max(1u, std::min<unsigned>(cdiv(ns, 256), 512));
    for (auto &pl : plan) {
        int shift = pl[0], bits = pl[1], nb = 1 << bits;
        LSH_CHECK(hipMemsetAsync(hist.p, 0, nb * sizeof(unsigned), stream()));
        launch_b("radix_select", 4.0 * ns, k_radix_hist, dim3(grid), dim3(256), 0, d, n, stride, prefix, pbits, shift, bits, hist.p);
        copy_mem(hh.data(), LSH_HOST, hist.p, LSH_DEVICE, nb * sizeof(unsigned));
        size_t acc = 0;
        int bin = 0;
        for (; bin < nb; bin++) {
            if (acc + hh[bin] > k) break;
            acc += hh[bin];
        }
        LSH_REQUIRE(bin < nb, "select: histogram does not cover the rank");
        k -= acc;
        prefix = (prefix << bits) | (uint32_t)bin;
        pbits += bits;
    }
    uint32_t key = prefix;
    uint32_t b = (key & 0x80000000u) ? (key & 0x7fffffffu) : ~key;
    float out;
    std::memcpy(&out, &b, 4);
    return out;
}

// depth.rs:628-650: 3x3 median, reflect-101
__global__ void k_median3(const float *__restrict__ src, int w, int h, float *__restrict__ out) {
    int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= w || y >= h) return;
    float v[9];
    int k = 0;
    for (int dy = -1; dy <= 1; dy++) {
        const float *row = src + (size_t)rf101(y + dy, h) * w;
        for (int dx = -1; dx <= 1; dx++) v[k++] = row[rf101(x + dx, w)];
    }
    for (int i = 1; i < 9; i++) {  // insertion sort by total order
        float t = v[i];
        int j = i;
        while (j > 0 && total_key(v[j - 1]) > total_key(t)) {
            v[j] = v[j - 1];
            j--;
        }
        v[j] = t;
    }
    out[(size_t)y * w + x] = v[4];
}

// depth.rs:619-621
__global__ void k_conf_norm(float *__restrict__ conf, size_t n, float p90) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) conf[i] = fminf(conf[i] / p90, 1.0f);
}

static dim3 g2(int w, int h) { return dim3(cdiv(w, 64), cdiv(h, 4)); }
static const dim3 B2(64, 4);
static unsigned g1(size_t n) { return cdiv(n, 256); }

// The two box passes: the LDS-tiled kernels when the radius fits their tiles,
// else the direct ones (the same window sums either way).  The direct kernels
// sum each window in f64; the tiled ones slide an f64 sum over a run of a few
// windows; the reference keeps an f64 running sum along the whole row
// (depth.rs:293-331).  All three agree whenever the f64 sums are exact, which
// they are for f32 planes whose values span fewer than about 2^24 in
// magnitude over a row (the rounding to f32 then sees the same value).
static void box_h2(const char *name, double bytes, const float *a, const float *b, const float *g, int mode, int w,
                   int h, int r, float *oa, float *ob) {
    if (r <= BOX_TILE_MAX_R && cdiv(h, BH_RB) <= 65535) {
        const size_t shm = 2 * (size_t)BH_RB * (BH_TX + 2 * r) * sizeof(float);
        launch_b(name, bytes, k_box_h2_tile, dim3(cdiv(w, BH_TX), cdiv(h, BH_RB)), dim3(BH_TX), shm, a, b, g, mode, w,
                 h, r, oa, ob);
    } else {
        launch_b(name, bytes, k_box_h2_direct, g2(w, h), B2, 0, a, b, g, mode, w, h, r, oa, ob);
    }
}

static void box_v2(const char *name, double bytes, const float *ta, const float *tb, int w, int h, int r, int mode,
                   const float *guide, const float *mean_i, const float *var_i, float eps, float *oa, float *ob,
                   const Peaks &P, uint32_t m) {
    if (r <= BOX_TILE_MAX_R) {
        const size_t shm = 2 * (size_t)(BV_TY + 2 * r) * BV_TX * sizeof(float);
        const dim3 grid(cdiv(w, BV_TX), cdiv(h, BV_TY));
        launch_b(name, bytes, k_box_v2_tile, grid, dim3(256), shm, ta, tb, w, h, r, mode, guide, mean_i, var_i, eps, oa,
                 ob, P, m);
    } else {
        launch_b(name, bytes, k_box_v2_direct, g2(w, h), B2, 0, ta, tb, w, h, r, mode, guide, mean_i, var_i, eps, oa,
                 ob, P, m);
    }
}

// Guided-filter plumbing on the working grid (depth.rs:349-412)
struct GF {
    int w, h, r;
    float eps;
    const float *guide;
    DBuf<float> mean_i, var_i, t1, t2, s1, s2;
    DBuf<float> t3, t4, s3, s4;  // the second slice of a pair
};

static void gf_init(GF &g, const float *guide, int w, int h, int r, float eps) {
    g.w = w;
    g.h = h;
    g.r = r;
    g.eps = eps;
    g.guide = guide;
    size_t n = (size_t)w * h;
    for (auto *b : {&g.mean_i, &g.var_i, &g.t1, &g.t2, &g.s1, &g.s2}) b->alloc(n);
    Peaks none{};
    box_h2("box_h", 0.0, guide, nullptr, nullptr, 2, w, h, r, g.t1.p, g.t2.p);
    box_v2("box_v", 0.0, g.t1.p, g.t2.p, w, h, r, 0, nullptr, nullptr, nullptr, 0.0f, g.mean_i.p, g.var_i.p, none, 0u);
}

// step 1 of a filtering of p: sa = a, sb = b (s1, s2 unless given)
static void gf_step1(GF &g, const float *p, float *sa = nullptr, float *sb = nullptr) {
    Peaks none{};
    double M = (double)g.w * g.h;
    box_h2("box_h", 16.0 * M, p, nullptr, g.guide, 1, g.w, g.h, g.r, g.t1.p, g.t2.p);
    box_v2("box_v", 24.0 * M, g.t1.p, g.t2.p, g.w, g.h, g.r, 1, g.guide, g.mean_i.p, g.var_i.p, g.eps, sa ? sa : g.s1.p,
           sb ? sb : g.s2.p, none, 0u);
}

// the filterings of slices m (s1, s2 from step 1) and m + 1 (s3, s4) to their
// ends, their peak pushes in one pass over the peak state
static void gf_step2_pair(GF &g, const Peaks &P, uint32_t m) {
    double M = (double)g.w * g.h;
    box_h2("box_h", 16.0 * M, g.s1.p, g.s2.p, nullptr, 0, g.w, g.h, g.r, g.t1.p, g.t2.p);
    box_h2("box_h", 16.0 * M, g.s3.p, g.s4.p, nullptr, 0, g.w, g.h, g.r, g.t3.p, g.t4.p);
    LSH_REQUIRE(g.r <= 16, "depth: paired pass radius");
    const size_t shm = 4 * (size_t)(BV_TY + 2 * g.r) * BV_TX * sizeof(float);
    // four planes and the guide read; the peak state's four planes read and
    // written once for the two slices (the profile only where a peak lands)
    launch_b("box_v_peaks2", (20.0 + 32.0) * M, k_box_v4_peaks_tile, dim3(cdiv(g.w, BV_TX), cdiv(g.h, BV_TY)), dim3(256),
             shm, (const float *)g.t1.p, (const float *)g.t2.p, (const float *)g.t3.p, (const float *)g.t4.p, g.w, g.h,
             g.r, g.guide, P, m);
}

// step 2: the box means of a and b, then per mode (2 peak push, 3 coefficients into oa/ob, 4 filtered into oa)
static void gf_step2(GF &g, int mode, float *oa, float *ob, const Peaks &P, uint32_t m) {
    double M = (double)g.w * g.h;
    box_h2("box_h", 16.0 * M, g.s1.p, g.s2.p, nullptr, 0, g.w, g.h, g.r, g.t1.p, g.t2.p);
    // mode 2: t1, t2 and the guide read; the peak push reads and writes prev, prev2, sum and cmin at every cell;
    // only where the previous slice was a peak does it also touch the profile
    // (c1, c2, i1, l1, r1), which no count can know beforehand
    double by = mode == 2 ? (12.0 + 32.0) * M : 16.0 * M;
    box_v2(mode == 2 ? "box_v_peaks" : "box_v", by, g.t1.p, g.t2.p, g.w, g.h, g.r, mode, g.guide, nullptr, nullptr,
           0.0f, oa, ob, P, m);
}

struct PeakBufs {
    DBuf<float> c1, l1, r1, c2, prev, prev2, sum, cmin;
    DBuf<uint32_t> i1;
    Peaks view() { return Peaks{c1.p, l1.p, r1.p, c2.p, prev.p, prev2.p, sum.p, cmin.p, i1.p}; }
};

// ------------------------------------------------------ slices in, lines out
// The depth pass's slice iterator (depth.rs:1120, 1137-1142), pulled in frame
// order: one contiguous array, or the caller's callback.
struct SliceSource {
    virtual ~SliceSource() = default;
    // slice m, its cell count and memory; nullptr once the source has ended; throws the source's error
    virtual const float *get(uint32_t m, uint64_t &cells, lsh_mem &mem) = 0;
    // the pass has read slice m
    virtual void done(uint32_t, lsh_mem) {}
    // every slice's memory stays the caller's until the pass ends (two may be read at once)
    virtual bool stable() const { return false; }
};

struct ArraySource final : SliceSource {
    const float *base;
    lsh_mem mem;
    size_t M;
    ArraySource(const float *b, lsh_mem m, size_t cells) : base(b), mem(m), M(cells) {}
    bool stable() const override { return mem == LSH_DEVICE; }
    const float *get(uint32_t m, uint64_t &cells, lsh_mem &mm) override {
        cells = M;
        mm = mem;
        return base + (size_t)m * M;
    }
};

struct CallbackSource final : SliceSource {
    lsh_slice_fn fn;
    void *user;
    CallbackSource(lsh_slice_fn f, void *u) : fn(f), user(u) {}
    const float *get(uint32_t m, uint64_t &cells, lsh_mem &mm) override {
        const float *s = nullptr;
        cells = 0;
        mm = LSH_HOST;
        const char *e = fn(user, m, &s, &cells, &mm);
        if (e) throw Error(e);  // the source's own message, as the reference's `??` passes it on
        LSH_REQUIRE(mm == LSH_HOST || mm == LSH_DEVICE, "depth: a slice's memory is LSH_HOST or LSH_DEVICE");
        return s;
    }
    // the caller may reuse a device buffer once its slice has been read (a
    // host slice was copied, synchronously, before its kernels)
    void done(uint32_t, lsh_mem mm) override {
        if (mm == LSH_DEVICE) sync();
    }
};

// Rust's Display of an f32 (the shortest decimal that reads back as the same
// float, never in exponent form), for log lines worded as the reference's.
static std::string rs_f32(float v) {
    if (std::isnan(v)) return "NaN";
    if (std::isinf(v)) return v > 0 ? "inf" : "-inf";
    if (v == 0.0f) return std::signbit(v) ? "-0" : "0";
    char buf[64];
    int p = 1;
    for (; p < 9; p++) {
        snprintf(buf, sizeof buf, "%.*e", p - 1, (double)v);
        if (strtof(buf, nullptr) == v) break;
    }
    snprintf(buf, sizeof buf, "%.*e", p - 1, (double)v);
    std::string s(buf);
    const bool neg = s[0] == '-';
    if (neg) s.erase(0, 1);
    const size_t ep = s.find('e');
    const int ex = atoi(s.c_str() + ep + 1);
    std::string dig;
    for (size_t i = 0; i < ep; i++)
        if (s[i] >= '0' && s[i] <= '9') dig += s[i];
    while (dig.size() > 1 && dig.back() == '0') dig.pop_back();
    const int point = ex + 1;  // digits before the decimal point
    std::string out;
    if (point <= 0)
        out = "0." + std::string((size_t)-point, '0') + dig;
    else if ((size_t)point >= dig.size())
        out = dig + std::string((size_t)point - dig.size(), '0');
    else
        out = dig.substr(0, (size_t)point) + "." + dig.substr((size_t)point);
    return neg ? "-" + out : out;
}

// Rust's Debug of an f32 inside a struct: Display, with ".0" on whole numbers;
// below 1e-4 or from 1e16 on, the shortest digits in exponent form ("1e-5")
static std::string rs_f32_dbg(float v) {
    const float a = std::fabs(v);
    if (std::isfinite(v) && a != 0.0f && (a < 1e-4f || a >= 1e16f)) {
        char buf[64];
        int p = 1;
        for (; p < 9; p++) {
            snprintf(buf, sizeof buf, "%.*e", p - 1, (double)v);
            if (strtof(buf, nullptr) == v) break;
        }
        std::string s(buf);
        const size_t ep = s.find('e');
        std::string m = s.substr(0, ep);
        if (m.find('.') != std::string::npos) {
            while (m.back() == '0') m.pop_back();
            if (m.back() == '.') m.pop_back();
        }
        return m + "e" + std::to_string(atoi(s.c_str() + ep + 1));
    }
    std::string s = rs_f32(v);
    if (std::isfinite(v) && s.find('.') == std::string::npos) s += ".0";
    return s;
}

// Rust's {:.1e}: one decimal, the exponent unpadded ("1.2e-6")
static std::string rs_e1(double v) {
    char buf[64];
    snprintf(buf, sizeof buf, "%.1e", v);
    std::string s(buf);
    const size_t ep = s.find('e');
    if (ep == std::string::npos) return s;
    return s.substr(0, ep) + "e" + std::to_string(atoi(s.c_str() + ep + 1));
}

// the depth pass's progress lines (depth.rs:1124 log), timed from its start
struct DepthLog {
    lsh_log_fn fn = nullptr;
    void *user = nullptr;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    bool on() const { return fn != nullptr; }
    double secs() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
    __attribute__((format(printf, 2, 3))) void line(const char *fmt, ...) const {
        if (!fn) return;
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        fn(user, buf);
    }
};

// per-block f64 sums of a plane, in a fixed order (the log's mean confidence)
__global__ void k_sum_part(const float *__restrict__ v, size_t n, double *__restrict__ part) {
    __shared__ double s[256];
    double a = 0.0;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) a += v[i];
    s[threadIdx.x] = a;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) s[threadIdx.x] += s[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

static double plane_mean(const float *v, size_t n) {
    const unsigned nb = (unsigned)std::min<size_t>(512, std::max<size_t>(1, cdiv(n, 256)));
    DBuf<double> part(nb);
    launch_b("sum_part", 4.0 * n, k_sum_part, dim3(nb), dim3(256), 0, v, n, part.p);
    std::vector<double> h(nb);
    copy_mem(h.data(), LSH_HOST, part.p, LSH_DEVICE, nb * sizeof(double));
    double s = 0.0;
    for (double x : h) s += x;
    return s / (double)n;
}

}  // namespace lsh

using namespace lsh;

// The depth pass up to the WLS solve (depth.rs:1125-1159): the guide, the
// aggregated focus volume's peaks, the noise floor, the median and the p90
// confidence normalisation, all on the working grid.
struct DepthGrid {
    int k = 1, dw = 0, dh = 0;
    DBuf<float> yfull, guide, depth_w, conf, mins;
    float p90 = 0.0f, floor_ = 0.0f;
};

static void depth_grid(DepthGrid &G, SliceSource &src, uint32_t n, const float *fused3, lsh_mem fused_mem, uint32_t w,
                       uint32_t h, const lsh_depth_params *p, const DepthLog &log) {
    LSH_REQUIRE(n > 0, "depth: no slices");
    LSH_REQUIRE(w > 0 && h > 0, "depth: empty frame");
    LSH_REQUIRE(p->scale <= 8, "depth: scale out of range");
    const int k = 1 << p->scale;
    const int dw = (int)((w + k - 1) / k), dh = (int)((h + k - 1) / k);
    G.k = k;
    G.dw = dw;
    G.dh = dh;
    const size_t N = (size_t)w * h, M = (size_t)dw * dh;
    (void)N;
    // the guide: the fused image's luma, block-averaged (depth.rs:1129-1131)
    DBuf<float> fused_d;
    const float *fz = fused3;
    if (fused_mem == LSH_HOST) {
        fused_d.alloc(3 * N);
        copy_mem(fused_d.p, LSH_DEVICE, fused3, LSH_HOST, 3 * N * sizeof(float));
        fz = fused_d.p;
    }
    DBuf<float> &yfull = G.yfull, &guide = G.guide;
    yfull.alloc(N);
    guide.alloc(M);
    launch_b("luma", 16.0 * N, k_luma3, dim3(g1(N)), dim3(256), 0, fz, N, yfull.p);
    launch_b("block_mean", 4.0 * ((double)N + (double)dw * dh), k_block_mean, g2(dw, dh), B2, 0, (const float *)yfull.p,
             (int)w, (int)h, k, guide.p, dw, dh);
    // aggregation and the streamed peak search (depth.rs:1137-1151)
    PeakBufs PB;
    for (auto *b : {&PB.c1, &PB.l1, &PB.r1, &PB.c2, &PB.prev, &PB.prev2, &PB.sum, &PB.cmin}) b->alloc(M);
    PB.i1.alloc(M);
    Peaks P = PB.view();
    launch_b("peak_init", 36.0 * M, k_peak_init, dim3(g1(M)), dim3(256), 0, P, M);
    GF gf;
    if (p->agg_radius > 0) gf_init(gf, guide.p, dw, dh, (int)p->agg_radius, p->agg_eps);
    if (log.on()) {
        char focus[64];
        if (p->focus_kind == 0)
            snprintf(focus, sizeof focus, "Rdf { r_in: %u, r_out: %u }", p->r_in, p->r_out);
        else
            snprintf(focus, sizeof focus, "Sml { step: %u }", p->sml_step);
        sync();
        log.line("depth: %u frames, focus %s, working grid %dx%d (1/%d), aggregation r=%u eps=%s (GPU, %.2fs to start)", n,
                 focus, dw, dh, k, p->agg_radius, rs_f32(p->agg_eps).c_str(), log.secs());
    }
    DBuf<float> stage;
    auto fetch = [&](uint32_t m, lsh_mem &mem) {
        uint64_t cells = 0;
        mem = LSH_HOST;
        const float *c = src.get(m, cells, mem);
        if (!c) throw Error("depth: slice " + std::to_string(m) + " of " + std::to_string(n) + " missing");
        if (cells != M)
            throw Error("depth: slice " + std::to_string(m) + " has " + std::to_string(cells) + " cells, the working grid " +
                        std::to_string(dw) + "x" + std::to_string(dh));
        if (mem == LSH_HOST) {
            if (!stage.p) stage.alloc(M);
            copy_mem(stage.p, LSH_DEVICE, c, LSH_HOST, M * sizeof(float));
            c = stage.p;
        }
        return c;
    };
    // slices in pairs where the source keeps both in place (a device array):
    // the last pass of both filterings updates the peak state once
    // (radius up to 16: the paired pass's four staged planes then fit 64 KB of LDS)
    const bool pairs = p->agg_radius > 0 && p->agg_radius <= 16 && src.stable();
    if (pairs && n >= 2)
        for (auto *b : {&gf.t3, &gf.t4, &gf.s3, &gf.s4}) b->alloc(M);
    for (uint32_t m = 0; m < n;) {
        lsh_mem mem = LSH_HOST, mem2 = LSH_HOST;
        const float *c = fetch(m, mem);
        if (pairs && m + 1 < n) {
            const float *c2 = fetch(m + 1, mem2);
            gf_step1(gf, c);
            gf_step1(gf, c2, gf.s3.p, gf.s4.p);
            gf_step2_pair(gf, P, m);
            src.done(m, mem);
            src.done(m + 1, mem2);
            m += 2;
            continue;
        }
        if (p->agg_radius > 0) {
            gf_step1(gf, c);
            gf_step2(gf, 2, nullptr, nullptr, P, m);
        } else {
            launch_b("peak_push", 76.0 * M, k_peak_push_raw, dim3(g1(M)), dim3(256), 0, c, M, P, m);
        }
        src.done(m, mem);
        m++;
    }
    if (log.on()) {
        sync();
        log.line("depth: %u slices aggregated  (%.1fs)", n, log.secs());
    }
    // close the profiles, the noise floor (depth.rs:568-580)
    // every cell reads prev and prev2; only the cells whose last slice is a
    // rising edge also read and write their profile, which no count can know
    launch_b("peak_close", 8.0 * M, k_peak_close, dim3(g1(M)), dim3(256), 0, P, M, n);
    DBuf<unsigned> hist;
    size_t ns = (M + 6) / 7;
    float floor_ = G.floor_ = select_kth(PB.cmin.p, M, 7, ns / 2, hist);
    DBuf<float> floor_d(1), &depth_w = G.depth_w, &conf = G.conf, &mins = G.mins;
    depth_w.alloc(M);
    conf.alloc(M);
    mins.alloc(M);
    copy_mem(floor_d.p, LSH_DEVICE, &floor_, LSH_HOST, sizeof(float));
    launch_b("peak_out", 52.0 * M, k_peak_out, dim3(g1(M)), dim3(256), 0, P, M, n, p->gate, (const float *)floor_d.p, depth_w.p,
           conf.p, mins.p);
    if (p->median) {
        DBuf<float> med(M);
        launch_b("median3", 8.0 * M, k_median3, g2(dw, dh), B2, 0, (const float *)depth_w.p, dw, dh, med.p);
        depth_w = std::move(med);
    }
    // p90 normalisation (depth.rs:615-623)
    size_t kk = std::min(ns * 9 / 10, ns - 1);
    float p90 = G.p90 = select_kth(conf.p, M, 7, kk, hist);
    if (p90 > 1e-6f) launch_b("conf_norm", 8.0 * M, k_conf_norm, dim3(g1(M)), dim3(256), 0, conf.p, M, p90);
    if (log.on())
        log.line("depth: peaks found, confidence p90 %.3f, mean (normalized) %.3f  (%.1fs)", p90, plane_mean(conf.p, M),
                 log.secs());
}

// depth.rs:1119-1200 over any slice source: the grid pass, the WLS solves, the
// upsampling, with the reference's progress lines when log is on
static void depth_run(SliceSource &src, uint32_t n, const float *fused3, lsh_mem fused_mem, uint32_t w, uint32_t h,
                      const lsh_depth_params *p, float *depth, float *conf_out, lsh_mem out_mem, float *floor_plane,
                      lsh_depth_info *info, const DepthLog &log) {
    LSH_REQUIRE(n > 0, "depth: no slices");
    LSH_REQUIRE(w > 0 && h > 0, "depth: empty frame");
    LSH_REQUIRE(p->scale <= 8, "depth: scale out of range");
    DepthGrid G;
    depth_grid(G, src, n, fused3, fused_mem, w, h, p, log);
    const int k = G.k, dw = G.dw, dh = G.dh;
    const size_t N = (size_t)w * h, M = (size_t)dw * dh;
    DBuf<float> &yfull = G.yfull, &guide = G.guide, &depth_w = G.depth_w, &conf = G.conf, &mins = G.mins;
    const float p90 = G.p90, floor_ = G.floor_;
    // WLS (depth.rs:1161-1178)
    uint32_t it1 = 0, it2 = 0;
    float rel1 = 0.0f, rel2 = 0.0f;
    double trel1 = 0.0, trel2 = 0.0;
    DBuf<float> u;
    const float *dres = depth_w.p;
    if (p->lambda > 0.0f) {
        LSH_REQUIRE(p->precision <= 1, "depth: precision is 0 (reference) or 1 (refined)");
        // cg_iters caps the CG iterations summed over the refined mode's inner solves
        const int mode = p->precision == 1 ? WLS_REFINED : WLS_REFERENCE;
        WlsPtr s = wls_create(guide.p, dw, dh, p->lambda, p->sigma_c);
        wls_set_weights(s.get(), conf.p, nullptr, nullptr, 0.0f);
        u.alloc(M);
        u.zero();
        // with a learned start the first solve begins at its prediction for
        // this system (depth in frames, so dscale is the frame range), the
        // reweighted one at the first solve's result; else both at the FGS sweeps
        const float *start = nullptr;
        if (p->warm) {
            warm_predict_dev(p->warm, depth_w.p, conf.p, guide.p, dw, dh, (float)std::max<uint32_t>(n - 1, 1),
                             p->sigma_c, u.p);
            start = u.p;
        }
        WlsResult r1 = wls_solve(s.get(), depth_w.p, p->cg_iters, p->tol, mode, start, u.p, nullptr);
        it1 = it2 = r1.iters;
        rel1 = rel2 = (float)r1.rel;
        trel1 = trel2 = wls_true_rel(s.get(), depth_w.p, u.p);
        if (p->robust > 0.0f) {
            log.line("depth: robust reweighting (tau=%s frames), first solve %u iterations, residual %s",
                     rs_f32(p->robust).c_str(), it1, rs_e1(rel1).c_str());
            wls_set_weights(s.get(), conf.p, depth_w.p, u.p, p->robust);
            WlsResult r2 = wls_solve(s.get(), depth_w.p, p->cg_iters, p->tol, mode, start, u.p, nullptr);
            it2 = r2.iters;
            rel2 = (float)r2.rel;
            trel2 = wls_true_rel(s.get(), depth_w.p, u.p);
        }
        dres = u.p;
    }
    log.line("depth: WLS lambda=%s sigma_c=%s solved in %u iterations, residual %s  (%.1fs)", rs_f32(p->lambda).c_str(),
             rs_f32(p->sigma_c).c_str(), it2, rs_e1(rel2).c_str(), log.secs());
    // upsampling (depth.rs:1180-1197)
    DBuf<float> dout, cout;
    float *dd = depth, *cc = conf_out;
    if (out_mem == LSH_HOST) {
        dout.alloc(N);
        cout.alloc(N);
        dd = dout.p;
        cc = cout.p;
    }
    float maxd = (float)(n - 1), inv = 1.0f / (float)k;
    if (p->upsample_kind == 0) {
        launch_b("upsample", 4.0 * ((double)M + N), k_up_bilinear, g2(w, h), B2, 0, dres, dw, dh, (int)w, (int)h, inv,
                 0.0f, maxd, 1, dd);
    } else {
        GF ug;
        gf_init(ug, guide.p, dw, dh, (int)p->up_radius, p->up_eps);
        gf_step1(ug, dres);
        Peaks none{};
        if (k == 1) {
            gf_step2(ug, 4, dd, nullptr, none, 0);
            launch_b("clamp", 8.0 * N, k_clamp, dim3(g1(N)), dim3(256), 0, dd, N, 0.0f, maxd);
        } else {
            DBuf<float> ca(M), cb(M);
            gf_step2(ug, 3, ca.p, cb.p, none, 0);
            launch_b("upsample", 8.0 * ((double)M + N), k_up_guided, g2(w, h), B2, 0, (const float *)ca.p, (const float *)cb.p, dw, dh,
                   (const float *)yfull.p, (int)w, (int)h, inv, maxd, dd);
        }
    }
    if (k == 1)
        LSH_CHECK(hipMemcpyAsync(cc, conf.p, M * sizeof(float), hipMemcpyDeviceToDevice, stream()));
    else
        launch_b("upsample", 4.0 * ((double)M + N), k_up_bilinear, g2(w, h), B2, 0, (const float *)conf.p, dw, dh, (int)w, (int)h, inv, 0.0f, 0.0f, 0, cc);
    if (out_mem == LSH_HOST) {
        copy_mem(depth, LSH_HOST, dd, LSH_DEVICE, N * sizeof(float));
        copy_mem(conf_out, LSH_HOST, cc, LSH_DEVICE, N * sizeof(float));
    }
    if (floor_plane) copy_mem(floor_plane, out_mem, mins.p, LSH_DEVICE, M * sizeof(float));
    sync();
    if (log.on()) {
        std::string up = "Bilinear";
        if (p->upsample_kind != 0)
            up = "Guided { radius: " + std::to_string(p->up_radius) + ", eps: " + rs_f32_dbg(p->up_eps) + " }";
        log.line("depth: upsampled to %ux%u (%s)  (%.1fs)", w, h, up.c_str(), log.secs());
    }
    if (info) {
        info->dw = (uint32_t)dw;
        info->dh = (uint32_t)dh;
        info->iters1 = it1;
        info->iters2 = it2;
        info->rel1 = rel1;
        info->rel2 = rel2;
        info->p90 = p90;
        info->floor = floor_;
        info->true_rel1 = trel1;
        info->true_rel2 = trel2;
    }
}

extern "C" {

char *lsh_depth_from_slices(const float *slices, lsh_mem slices_mem, uint32_t n, const float *fused3,
                            lsh_mem fused_mem, uint32_t w, uint32_t h, const lsh_depth_params *p, float *depth,
                            float *conf_out, lsh_mem out_mem, float *floor_plane, lsh_depth_info *info) {
    LSH_API_BEGIN
    LSH_REQUIRE(slices && fused3 && p && depth && conf_out, "lsh_depth_from_slices: null argument");
    LSH_REQUIRE(p->scale <= 8, "depth: scale out of range");
    const size_t k = (size_t)1 << p->scale;
    ArraySource src(slices, slices_mem, ((w + k - 1) / k) * ((h + k - 1) / k));
    depth_run(src, n, fused3, fused_mem, w, h, p, depth, conf_out, out_mem, floor_plane, info, DepthLog{});
    LSH_API_END
}

char *lsh_depth_from_source(lsh_slice_fn next, void *next_user, uint32_t n, const float *fused3, lsh_mem fused_mem,
                            uint32_t w, uint32_t h, const lsh_depth_params *p, float *depth, float *conf,
                            lsh_mem out_mem, float *floor_plane, lsh_depth_info *info, lsh_log_fn log, void *log_user) {
    LSH_API_BEGIN
    LSH_REQUIRE(next && fused3 && p && depth && conf, "lsh_depth_from_source: null argument");
    CallbackSource src(next, next_user);
    DepthLog lg;
    lg.fn = log;
    lg.user = log_user;
    depth_run(src, n, fused3, fused_mem, w, h, p, depth, conf, out_mem, floor_plane, info, lg);
    LSH_API_END
}

char *lsh_depth_grid(const float *slices, lsh_mem slices_mem, uint32_t n, const float *fused3, lsh_mem fused_mem,
                     uint32_t w, uint32_t h, const lsh_depth_params *p, float *guide, float *depth, float *conf,
                     lsh_mem out_mem, lsh_depth_info *info) {
    LSH_API_BEGIN
    LSH_REQUIRE(slices && fused3 && p && guide && depth && conf, "lsh_depth_grid: null argument");
    LSH_REQUIRE(p->scale <= 8, "depth: scale out of range");
    const size_t k = (size_t)1 << p->scale;
    ArraySource src(slices, slices_mem, ((w + k - 1) / k) * ((h + k - 1) / k));
    DepthGrid G;
    depth_grid(G, src, n, fused3, fused_mem, w, h, p, DepthLog{});
    const size_t M = (size_t)G.dw * G.dh;
    copy_mem(guide, out_mem, G.guide.p, LSH_DEVICE, M * sizeof(float));
    copy_mem(depth, out_mem, G.depth_w.p, LSH_DEVICE, M * sizeof(float));
    copy_mem(conf, out_mem, G.conf.p, LSH_DEVICE, M * sizeof(float));
    sync();
    if (info) {
        *info = lsh_depth_info{};
        info->dw = (uint32_t)G.dw;
        info->dh = (uint32_t)G.dh;
        info->p90 = G.p90;
        info->floor = G.floor_;
    }
    LSH_API_END
}

}  // extern "C"

// ----------------------------------------------------- the weighted average
// wav.rs:97-161 as an accumulator: per aligned frame the focus slice on the
// working grid (the depth pass's measure), its box mean, the weights above the
// noise floor, their smoothing (the guided filter on the fused luma's grid
// means, or the box mean), and the weighted sums at full resolution with the
// grid weights taken there bilinearly on the fly; at the end the clamped
// quotient.  Each step is the reference's operation in its order.
namespace lsh {

// wav.rs:88-91
__global__ void k_wav_weight(const float *__restrict__ c, const float *__restrict__ floor_p, size_t n, float gate,
                             float power, float *__restrict__ out) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float f = floor_p[i];
    out[i] = powf(fmaxf(c[i] - (1.0f + gate) * f, 0.0f), power) + powf(1e-2f * f, power) + 1e-30f;
}

// wav.rs:142: the guided filter undershoots at an edge, a weight stays a weight
__global__ void k_wav_atleast(float *__restrict__ v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = fmaxf(v[i], 1e-30f);
}

// wav.rs:146-155: the grid weights at full resolution (depth.rs:201-221) and the weighted sums
__global__ void k_wav_accum(const float *__restrict__ g, int dw, int dh, int w, int h, float inv,
                            const float *__restrict__ planes, float *__restrict__ acc, float *__restrict__ wsum) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= w || y >= h) return;
    const size_t n = (size_t)w * h, o = (size_t)y * w + x;
    const float wv = bilinear_at(g, dw, dh, x, y, inv);
    acc[o] += wv * planes[o];
    acc[n + o] += wv * planes[n + o];
    acc[2 * n + o] += wv * planes[2 * n + o];
    wsum[o] += wv;
}

// wav.rs:158-160
__global__ void k_wav_finish(const float *__restrict__ acc, const float *__restrict__ wsum, size_t n,
                             float *__restrict__ out) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float s = wsum[i];
    out[i] = rclampf(acc[i] / s, 0.0f, 1.0f);
    out[n + i] = rclampf(acc[n + i] / s, 0.0f, 1.0f);
    out[2 * n + i] = rclampf(acc[2 * n + i] / s, 0.0f, 1.0f);
}

// the box mean of one grid plane (depth.rs:263-347)
static void box_mean_plane(const float *src, int w, int h, int r, float *out, float *t1, float *t2, float *spare) {
    Peaks none{};
    double M = (double)w * h;
    box_h2("box_h", 12.0 * M, src, src, nullptr, 0, w, h, r, t1, t2);
    box_v2("box_v", 16.0 * M, t1, t2, w, h, r, 3, nullptr, nullptr, nullptr, 0.0f, out, spare, none, 0u);
}

}  // namespace lsh

struct lsh_wav {
    size_t w = 0, h = 0, k = 1, dw = 0, dh = 0, count = 0;
    lsh_depth_params dp{};
    float power = 2.0f, gate = 0.5f;
    uint32_t smooth = 3;
    bool guided = false;
    lsh::GF gf;
    lsh::DBuf<float> floor, gy, acc, wsum, g, t, b1, b2, b3, slice, stage;
    lsh::DBuf<int32_t> taps_dy, taps_dx;
    lsh::DBuf<float> taps_wt;
    uint32_t ntaps = 0, taps_key = 0xffffffffu;
};

namespace lsh {
static void wav_push(lsh_wav *a, const float *planes, const float *slice) {
    const int dw = (int)a->dw, dh = (int)a->dh;
    const size_t m = a->dw * a->dh;
    const float *c = slice;
    if (!c) {
        focus_slice_with(a->taps_dy, a->taps_dx, a->taps_wt, a->ntaps, a->taps_key, a->w, a->h, planes, a->dp,
                         a->slice.p);
        c = a->slice.p;
    }
    if (a->smooth > 0) {
        box_mean_plane(c, dw, dh, (int)a->smooth, a->g.p, a->b1.p, a->b2.p, a->b3.p);
        c = a->g.p;
    }
    launch_b("wav_weight", 12.0 * m, k_wav_weight, dim3(g1(m)), dim3(256), 0, c, (const float *)a->floor.p, m, a->gate,
             a->power, a->t.p);
    const float *wg = a->t.p;
    if (a->guided) {
        Peaks none{};
        gf_step1(a->gf, a->t.p);
        gf_step2(a->gf, 4, a->g.p, nullptr, none, 0u);
        launch("wav_atleast", k_wav_atleast, dim3(g1(m)), dim3(256), 0, a->g.p, m);
        wg = a->g.p;
    } else if (a->smooth > 0) {
        box_mean_plane(a->t.p, dw, dh, (int)a->smooth, a->g.p, a->b1.p, a->b2.p, a->b3.p);
        wg = a->g.p;
    }
    const double N = (double)a->w * a->h;
    launch_b("wav_accum", 32.0 * N, k_wav_accum, g2((int)a->w, (int)a->h), B2, 0, wg, dw, dh, (int)a->w, (int)a->h,
             1.0f / (float)a->k, planes, a->acc.p, a->wsum.p);
    a->count++;
}
}  // namespace lsh

extern "C" {

char *lsh_wav_new(uint32_t w, uint32_t h, const lsh_depth_params *dp, const lsh_wav_params *wp, const float *floor_p,
                  lsh_mem floor_mem, const float *fused3, lsh_mem fused_mem, lsh_wav **out) {
    LSH_API_BEGIN
    LSH_REQUIRE(dp && wp && floor_p && out, "lsh_wav_new: null argument");
    *out = nullptr;
    LSH_REQUIRE(w > 0 && h > 0, "weighted average: empty frame");
    LSH_REQUIRE(dp->scale <= 2, "weighted average: scale above 2 is not supported");
    auto a = new lsh_wav();
    try {
        a->w = w;
        a->h = h;
        a->k = (size_t)1 << dp->scale;
        a->dw = (w + a->k - 1) / a->k;
        a->dh = (h + a->k - 1) / a->k;
        a->dp = *dp;
        a->power = std::max(wp->power, 0.0f);  // wav.rs:110-111
        a->gate = std::max(wp->gate, 0.0f);
        a->smooth = wp->smooth;
        a->guided = wp->edge && wp->smooth > 0 && fused3 != nullptr;
        const size_t m = a->dw * a->dh, N = (size_t)w * h;
        a->floor.alloc(m);
        copy_mem(a->floor.p, LSH_DEVICE, floor_p, floor_mem, m * sizeof(float));
        for (auto *b : {&a->g, &a->t, &a->b1, &a->b2, &a->b3, &a->slice}) b->alloc(m);
        a->acc.alloc(3 * N);
        a->wsum.alloc(N);
        a->acc.zero();
        a->wsum.zero();
        if (a->guided) {  // wav.rs:114-122: the fused luma's block means on the grid guide the weights
            DBuf<float> f3, y(N);
            const float *src = fused3;
            if (fused_mem == LSH_HOST) {
                f3.alloc(3 * N);
                copy_mem(f3.p, LSH_DEVICE, fused3, LSH_HOST, 3 * N * sizeof(float));
                src = f3.p;
            }
            launch("luma", k_luma3, dim3(g1(N)), dim3(256), 0, src, N, y.p);
            a->gy.alloc(m);
            launch("block_mean", k_block_mean, g2((int)a->dw, (int)a->dh), B2, 0, (const float *)y.p, (int)w, (int)h,
                   (int)a->k, a->gy.p, (int)a->dw, (int)a->dh);
            gf_init(a->gf, a->gy.p, (int)a->dw, (int)a->dh, (int)a->smooth, dp->agg_eps);
            sync();
        }
    } catch (...) {
        delete a;
        throw;
    }
    *out = a;
    LSH_API_END
}

char *lsh_wav_push(lsh_wav *a, const float *img3, lsh_mem img_mem, const float *slice, lsh_mem slice_mem) {
    LSH_API_BEGIN
    LSH_REQUIRE(a && img3, "lsh_wav_push: null argument");
    const size_t N = a->w * a->h, m = a->dw * a->dh;
    const float *planes = img3;
    if (img_mem == LSH_HOST) {
        a->stage.ensure(3 * N);
        copy_mem(a->stage.p, LSH_DEVICE, img3, LSH_HOST, 3 * N * sizeof(float));
        planes = a->stage.p;
    }
    const float *sl = slice;
    if (slice && slice_mem == LSH_HOST) {
        copy_mem(a->slice.p, LSH_DEVICE, slice, LSH_HOST, m * sizeof(float));
        sl = a->slice.p;
    }
    wav_push(a, planes, sl);
    LSH_API_END
}

char *lsh_wav_push_frames(lsh_wav *a, lsh_frames *src) {
    LSH_API_BEGIN
    LSH_REQUIRE(a && src, "lsh_wav_push_frames: null argument");
    LSH_REQUIRE(src->w == a->w && src->h == a->h, "weighted average: the frames handle's size differs");
    LSH_REQUIRE(src->processed > 0, "weighted average: the frames handle holds no frame");
    wav_push(a, src->planes.p, nullptr);
    LSH_API_END
}

char *lsh_wav_finish(lsh_wav *a, float *img3, lsh_mem mem) {
    LSH_API_BEGIN
    LSH_REQUIRE(a && img3, "lsh_wav_finish: null argument");
    LSH_REQUIRE(a->count > 0, "weighted average: no frames");
    const size_t N = a->w * a->h;
    DBuf<float> tmp;
    float *dst = img3;
    if (mem == LSH_HOST) {
        tmp.alloc(3 * N);
        dst = tmp.p;
    }
    launch_b("wav_finish", 28.0 * N, k_wav_finish, dim3(g1(N)), dim3(256), 0, (const float *)a->acc.p,
             (const float *)a->wsum.p, N, dst);
    if (mem == LSH_HOST) copy_mem(img3, LSH_HOST, tmp.p, LSH_DEVICE, 3 * N * sizeof(float));
    sync();
    LSH_API_END
}

void lsh_wav_free(lsh_wav *a) {
    if (!a) return;
    try {
        sync();
    } catch (...) {
    }
    delete a;
}

uint32_t lsh_slab_ranges(uint32_t count, uint32_t size, uint32_t overlap, uint32_t *lohi, uint32_t cap) {
    // stack.rs:76-92
    if (count == 0) return 0;
    size = std::min(std::max(size, 1u), count);
    overlap = std::min(overlap, size - 1);
    uint32_t lo = 0, ns = 0;
    for (;;) {
        const uint32_t hi = std::min(lo + size - 1, count - 1);
        if (lohi && ns < cap) {
            lohi[2 * ns] = lo;
            lohi[2 * ns + 1] = hi;
        }
        ns++;
        if (hi + 1 >= count) return ns;
        lo = hi + 1 - overlap;
    }
}

}  // extern "C"
