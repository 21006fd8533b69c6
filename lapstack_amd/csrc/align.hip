// align.hip -- registration and resampling on gfx950 (align.rs, brightness.rs,
// the per-frame part of stack.rs), behind the frames part of the C ABI.
//
// A frame arrives as interleaved RGB u16 in HBM and is never expanded to f32
// before the warp: every kernel reads the 16-bit taps and scales each by
// 1/65535 exactly as the decoder does (io.rs:89), so the values match the
// CPU path's planes bit for bit.
//
// The search is the reference's bounded Nelder-Mead, run on the host; each
// cost evaluation is one kernel over the reference level that warps the
// target level (f64 coordinates, f32 Spline4x4 taps), folds d, d^2 and the
// count per thread in f64, reduces them across the block with wavefront
// shuffles and across blocks in a fixed order in the last block to finish,
// and hands back three numbers.
#include <algorithm>

#include "api.hpp"
#include "common.hpp"
#include "engine.hpp"

namespace lsh {

void focus_slice(lsh_frames *f, const float *planes, const lsh_depth_params &p, float *out);  // depth.hip

#define INV16 (1.0f / 65535.0f)

// align.rs:562-571 on the decoded frame
__global__ void k_luma_u16(const uint16_t *__restrict__ rgb, size_t n, float *__restrict__ y) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float r = (float)rgb[3 * i] * INV16, g = (float)rgb[3 * i + 1] * INV16, b = (float)rgb[3 * i + 2] * INV16;
    y[i] = 0.299f * r + 0.587f * g + 0.114f * b;
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

__global__ __launch_bounds__(256) void k_burt(const float *__restrict__ src, int w, int h, float *__restrict__ dst,
                                              int ow, int oh) {
    int oj = blockIdx.x * blockDim.x + threadIdx.x, oi = blockIdx.y * blockDim.y + threadIdx.y;
    if (oj >= ow || oi >= oh) return;
    float k[5];
    burt_k(k);
    int cx = 2 * oj, cy = 2 * oi;
    float rn = 0.0f, cn = 0.0f;
    for (int t = 0; t < 5; t++) {
        int yy = cy + t - 2, xx = cx + t - 2;
        if (yy >= 0 && yy < h) rn += k[t];
        if (xx >= 0 && xx < w) cn += k[t];
    }
    float acc = 0.0f;
    for (int t = 0; t < 5; t++) {
        int s = cy + t - 2;
        if (s >= 0 && s < h) {
            const float *row = src + (size_t)s * w;
            float a = 0.0f;
            for (int u = 0; u < 5; u++) {
                int i = cx + u - 2;
                if (i >= 0 && i < w) a += k[u] * row[i];
            }
            acc += k[t] * a;
        }
    }
    dst[(size_t)oi * ow + oj] = acc / (rn * cn);
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

__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
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

// align.rs:581-635: one cost evaluation -> (sum d, sum d^2, count)
__global__ __launch_bounds__(256) void k_cost(const float *__restrict__ rf, const float *__restrict__ tg, int tw, int th,
                                              Inv9 inv, int affine, double *__restrict__ part,
                                              unsigned *__restrict__ ticket, double *__restrict__ out) {
    double v[3] = {0.0, 0.0, 0.0};
    size_t n = (size_t)tw * th;
    double xmax = (double)(tw - 1), ymax = (double)(th - 1);
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        int y = (int)(i / tw), x = (int)(i - (size_t)y * tw);
        double sx, sy;
        src_point(inv, affine, x, y, sx, sy);
        if (!(sx >= 0.0 && sx <= xmax && sy >= 0.0 && sy <= ymax)) continue;
        double fx = floor(sx), fy = floor(sy);
        int x0 = (int)fx, y0 = (int)fy;
        float wx[4], wy[4];
        spline4f((float)(sx - fx), wx);
        spline4f((float)(sy - fy), wy);
        float acc = 0.0f;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const float *row = tg + (size_t)min(max(y0 + j - 1, 0), th - 1) * tw;
            float r = 0.0f;
#pragma unroll
            for (int q = 0; q < 4; q++) r += wx[q] * row[min(max(x0 + q - 1, 0), tw - 1)];
            acc += wy[j] * r;
        }
        double d = (double)(rf[i] - acc);
        v[0] += d;
        v[1] += d * d;
        v[2] += 1.0;
    }
    grid_sum<3>(v, part, ticket, out);
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
__device__ __forceinline__ void interp_w(int ik, double t, float *wf) {
    double w[6];
    int N = ik <= 1 ? 2 : (ik <= 3 ? 4 : 6);
    switch (ik) {
    case 0:
        w[0] = t < 0.5 ? 1.0 : 0.0;
        w[1] = t < 0.5 ? 0.0 : 1.0;
        break;
    case 1:
        w[0] = 1.0 - t;
        w[1] = t;
        break;
    case 3:
        w[0] = ((-1.0 / 3.0 * t + 0.8) * t - 0.46666667) * t;
        w[1] = ((t - 1.8) * t - 0.2) * t + 1.0;
        w[2] = ((1.2 - t) * t + 0.8) * t;
        w[3] = ((1.0 / 3.0 * t - 0.2) * t - 0.13333334) * t;
        break;
    default: {
        double sum = 0.0;
        for (int i = 0; i < N; i++) {
            double d = fabs(t - ((double)i + 1.0 - (double)(N / 2)));
            w[i] = ik == 2 ? keys_k(d) : (ik == 4 ? spline36_k(d) : lanczos3_k(d));
            sum += w[i];
        }
        for (int i = 0; i < N; i++) w[i] /= sum;
    }
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

// The final resample of a frame (align.rs:554-559) with the as-shot fallback
// outside the frame (stack.rs:669-678), the brightness gains (brightness.rs:
// 124-133), and optionally the warped registration luma (stack.rs:527).
template <int N>
__global__ __launch_bounds__(256) void k_warp(const uint16_t *__restrict__ rgb, const float *__restrict__ yin, int w,
                                              int h, Inv9 inv, int affine, int ik, Gains g, float *__restrict__ planes,
                                              float *__restrict__ yout) {
    int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= w || y >= h) return;
    size_t n = (size_t)w * h, o = (size_t)y * w + x;
    double sx, sy;
    src_point(inv, affine, x, y, sx, sy);
    bool valid = sx >= 0.0 && sx <= (double)(w - 1) && sy >= 0.0 && sy <= (double)(h - 1);
    double fx = floor(sx), fy = floor(sy);
    int x0 = (int)fmax(fmin(fx, 1e9), -1e9), y0 = (int)fmax(fmin(fy, 1e9), -1e9);
    float wx[6], wy[6];
    interp_w(ik, sx - fx, wx);
    interp_w(ik, sy - fy, wy);
#pragma unroll
    for (int c = 0; c < 3; c++) {
        float v = valid ? warp_sample<N>(rgb, nullptr, c, w, h, x0, y0, wx, wy) : (float)rgb[3 * o + c] * INV16;
        if (g.on[c]) v *= g.g[c];
        planes[c * n + o] = v;
    }
    if (yout) yout[o] = warp_sample<N>(nullptr, yin, -1, w, h, x0, y0, wx, wy);
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
__global__ __launch_bounds__(256) void k_bright_sums(const uint16_t *__restrict__ rgb, int w, int h, Inv9 inv,
                                                     int affine, int ik, int mode, const float *__restrict__ grid,
                                                     int gw, int gh, double *__restrict__ part,
                                                     unsigned *__restrict__ ticket, double *__restrict__ out) {
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
            interp_w(ik, sx - fx, wx);
            interp_w(ik, sy - fy, wy);
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

static void warp_launch(lsh_frames *f, const uint16_t *rgb, const Inv9 &inv, int affine, const Gains &g,
                        float *yout) {
    dim3 gr(cdiv(f->w, 64), cdiv(f->h, 4)), b(64, 4);
    const float *yin = f->yt.p;
    int ik = (int)f->interp, w = (int)f->w, h = (int)f->h;
    // u16 frame read, 3 f32 planes written; with the luma: luma read, warped luma written
    double by = (double)f->w * f->h * (yout ? 26.0 : 18.0);
    switch (taps_n(f->interp)) {
    case 2: launch_b("warp", by, k_warp<2>, gr, b, 0, rgb, yin, w, h, inv, affine, ik, g, f->planes.p, yout); break;
    case 4: launch_b("warp", by, k_warp<4>, gr, b, 0, rgb, yin, w, h, inv, affine, ik, g, f->planes.p, yout); break;
    default: launch_b("warp", by, k_warp<6>, gr, b, 0, rgb, yin, w, h, inv, affine, ik, g, f->planes.p, yout); break;
    }
}

static void sums_launch(lsh_frames *f, const uint16_t *rgb, const Inv9 &inv, int affine, int mode, double *host7) {
    size_t ng = f->gw * f->gh;
    unsigned grid = std::max(1u, std::min<unsigned>(cdiv(ng, 256), 1024));
    f->red.ensure((size_t)grid * 7);
    const float *gp = f->have_grid ? f->grid.p : nullptr;
    int ik = (int)f->interp, w = (int)f->w, h = (int)f->h, gw = (int)f->gw, gh = (int)f->gh;
    switch (taps_n(f->interp)) {
    case 2: launch("bright_sums", k_bright_sums<2>, dim3(grid), dim3(256), 0, rgb, w, h, inv, affine, ik, mode, gp, gw, gh, f->red.p, f->ticket.p, f->red_out.p); break;
    case 4: launch("bright_sums", k_bright_sums<4>, dim3(grid), dim3(256), 0, rgb, w, h, inv, affine, ik, mode, gp, gw, gh, f->red.p, f->ticket.p, f->red_out.p); break;
    default: launch("bright_sums", k_bright_sums<6>, dim3(grid), dim3(256), 0, rgb, w, h, inv, affine, ik, mode, gp, gw, gh, f->red.p, f->ticket.p, f->red_out.p); break;
    }
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
static void build_reg_pyramid(lsh_frames *f, float *base) {
    for (size_t l = 1; l < f->lw.size(); l++) {
        dim3 g(cdiv(f->lw[l], 64), cdiv(f->lh[l], 4));
        launch_b("burt_reduce", 4.0 * (f->lw[l - 1] * f->lh[l - 1] + f->lw[l] * f->lh[l]), k_burt, g, dim3(64, 4), 0,
                 (const float *)(base + f->loff[l - 1]), (int)f->lw[l - 1], (int)f->lh[l - 1], base + f->loff[l],
                 (int)f->lw[l], (int)f->lh[l]);
    }
}

// One cost evaluation at registration level lvl (align.rs:926-933)
static double eval_cost(lsh_frames *f, size_t lvl, const Sim &s) {
    size_t tw = f->lw[lvl], th = f->lh[lvl], n = tw * th;
    int affine;
    Inv9 inv = inv_for(s, tw, th, affine);
    unsigned grid = std::max(1u, std::min<unsigned>(cdiv(n, 256), 1024));
    f->red.ensure((size_t)grid * 7);
    launch_b("warp_cost", 8.0 * n, k_cost, dim3(grid), dim3(256), 0, (const float *)(f->yr.p + f->loff[lvl]),
           (const float *)(f->yt.p + f->loff[lvl]), (int)tw, (int)th, inv, affine, f->red.p, f->ticket.p, f->red_out.p);
    double *pin = (double *)pinned_scratch(64);
    LSH_CHECK(hipMemcpyAsync(pin, f->red_out.p, 3 * sizeof(double), hipMemcpyDeviceToHost, stream()));
    sync();
    f->evals++;
    double sd = pin[0], sd2 = pin[1], cnt = pin[2];
    if (cnt < 16.0) return 1e9;
    double v = (sd2 - sd * sd / cnt) / cnt;
    return std::sqrt(v > 0.0 ? v : 0.0);
}

static const double SPAN[8] = {0.10, 0.10, 0.10, 0.087266462599716474, 0.05, 0.05, 0.05, 0.05};

static double clampd(double v, double lo, double hi) {
    if (v < lo) v = lo;
    if (v > hi) v = hi;
    return v;
}

// align.rs:888-894
static bool nm_converged(double (*s)[8], const double *fv, const double *tol, size_t n) {
    if (std::fabs(fv[n] - fv[0]) > 1e-4 * (1.0 + std::fabs(fv[0]))) return false;
    for (size_t i = 1; i <= n; i++)
        for (size_t k = 0; k < n; k++)
            if (!(std::fabs(s[i][k] - s[0][k]) <= tol[k])) return false;
    return true;
}

// align.rs:796-883: bounded Nelder-Mead, cost on the device
template <typename F>
static void nelder_mead(F &&f, size_t n, const double *x0, const double *lo, const double *hi, const double *step,
                        const double *tol, double *out) {
    if (n == 0) return;
    double s[9][8] = {{0}}, fv[9] = {0}, tmp[9][8], tfv[9];
    auto clampv = [&](double *v) {
        for (size_t k = 0; k < n; k++) v[k] = clampd(v[k], lo[k], hi[k]);
    };
    for (size_t k = 0; k < n; k++) s[0][k] = x0[k];
    for (size_t i = 0; i < n; i++) {
        for (size_t k = 0; k < n; k++) s[i + 1][k] = x0[k];
        s[i + 1][i] += step[i];
        clampv(s[i + 1]);
    }
    for (size_t i = 0; i <= n; i++) fv[i] = f(s[i]);
    const double alpha = 1.0, gamma = 2.0, rho = 0.5, sigma = 0.5;
    for (int it = 0; it < 200; it++) {
        size_t idx[9];
        for (size_t i = 0; i <= n; i++) idx[i] = i;
        for (size_t i = 1; i <= n; i++) {  // stable insertion sort by value
            size_t v = idx[i], j = i;
            while (j > 0 && fv[idx[j - 1]] > fv[v]) {
                idx[j] = idx[j - 1];
                j--;
            }
            idx[j] = v;
        }
        for (size_t i = 0; i <= n; i++) {
            std::memcpy(tmp[i], s[idx[i]], sizeof(tmp[i]));
            tfv[i] = fv[idx[i]];
        }
        std::memcpy(s, tmp, sizeof(s));
        std::memcpy(fv, tfv, sizeof(fv));
        if (nm_converged(s, fv, tol, n)) break;
        double c[8] = {0};
        for (size_t i = 0; i < n; i++)
            for (size_t k = 0; k < n; k++) c[k] += s[i][k];
        for (size_t k = 0; k < n; k++) c[k] /= (double)n;
        double xr[8];
        for (size_t k = 0; k < n; k++) xr[k] = c[k] + alpha * (c[k] - s[n][k]);
        clampv(xr);
        double fr = f(xr);
        if (fr < fv[0]) {
            double xe[8];
            for (size_t k = 0; k < n; k++) xe[k] = c[k] + gamma * (xr[k] - c[k]);
            clampv(xe);
            double fe = f(xe);
            if (fe < fr) {
                std::memcpy(s[n], xe, sizeof(xe));
                fv[n] = fe;
            } else {
                std::memcpy(s[n], xr, sizeof(xr));
                fv[n] = fr;
            }
        } else if (fr < fv[n - 1]) {
            std::memcpy(s[n], xr, sizeof(xr));
            fv[n] = fr;
        } else {
            double xc[8];
            for (size_t k = 0; k < n; k++) xc[k] = c[k] + rho * (s[n][k] - c[k]);
            clampv(xc);
            double fc = f(xc);
            if (fc < fv[n]) {
                std::memcpy(s[n], xc, sizeof(xc));
                fv[n] = fc;
            } else {
                for (size_t i = 1; i <= n; i++) {
                    for (size_t k = 0; k < n; k++) s[i][k] = s[0][k] + sigma * (s[i][k] - s[0][k]);
                    clampv(s[i]);
                    fv[i] = f(s[i]);
                }
            }
        }
    }
    size_t best = 0;
    for (size_t i = 1; i <= n; i++)
        if (fv[i] < fv[best]) best = i;
    for (size_t k = 0; k < n; k++) out[k] = s[best][k];
}

// align.rs:896-942 with the target pyramid in yt and the reference in yr
static Sim multiscale_align(lsh_frames *f, const Sim &init, const uint8_t *free8, size_t coarsen) {
    size_t n = f->lw.size();  // levels 0..n-1, level 0 the full-resolution plane
    size_t free_idx[8], nfree = 0;
    for (size_t k = 0; k < 8; k++)
        if (free8[k]) free_idx[nfree++] = k;
    double lo[8], hi[8];
    for (size_t k = 0; k < nfree; k++) {
        lo[k] = init.v[free_idx[k]] - SPAN[free_idx[k]];
        hi[k] = init.v[free_idx[k]] + SPAN[free_idx[k]];
    }
    Sim cur = init;
    size_t finest = std::min(coarsen, n - 1);
    double W = (double)f->w, H = (double)f->h;
    double r = std::max(W, H) / 2.0;
    const double gain[8] = {W, H, r, r, H / 2.0, H / 2.0, W / 4.0, H / 4.0};  // align.rs:766-770
    for (size_t lvl = n; lvl-- > finest;) {
        Sim snap = cur;
        auto cost = [&](const double *xf) {
            Sim v = snap;
            for (size_t k = 0; k < nfree; k++) v.v[free_idx[k]] = xf[k];
            return eval_cost(f, lvl, v);
        };
        double x0[8], step[8], tol[8], best[8];
        double px = (double)((uint64_t)1 << lvl);
        for (size_t k = 0; k < nfree; k++) {
            x0[k] = cur.v[free_idx[k]];
            step[k] = 1.0 * px / gain[free_idx[k]];  // align.rs:784-790
            tol[k] = 0.1 * px / gain[free_idx[k]];
        }
        nelder_mead(cost, nfree, x0, lo, hi, step, tol, best);
        for (size_t k = 0; k < nfree; k++) cur.v[free_idx[k]] = best[k];
    }
    return cur;
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
        launch_b("luma", 10.0 * n, k_luma_u16, dim3(cdiv(n, 256)), dim3(256), 0, src, n, f->yt.p);
        build_reg_pyramid(f, f->yt.p);
        sim = multiscale_align(f, init, free8, coarsen);
    }
    int affine = 1;
    Inv9 inv = inv_for(sim, w, h, affine);
    // brightness gains against frame 0's grid (stack.rs:681-690)
    if (!known_sim && brightness) {
        if (first) {
            dim3 gg(cdiv(f->gw, 64), cdiv(f->gh, 4));
            launch("bright_grid", k_grid, gg, dim3(64, 4), 0, src, (int)w, (int)h, (int)f->gw, (int)f->gh, f->grid.p);
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
            launch_b("luma", 10.0 * n, k_luma_u16, dim3(cdiv(n, 256)), dim3(256), 0, src, n, f->yt.p);
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

}  // extern "C"
