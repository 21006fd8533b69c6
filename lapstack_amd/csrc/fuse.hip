// fuse.hip -- the Laplacian pyramid (pyramid.rs) and the Wang-Chang region
// energy fold (fuse.rs) on gfx950, behind the Fuser part of the C ABI.
//
// Per frame: a REDUCE chain builds the Gaussian levels; then one fused kernel
// per band-pass level computes L = G - EXPAND(G_next) on an LDS tile with a
// halo, the node energy Y^2, the separable binomial window and the strict-'>'
// select into the accumulator, without writing L.  The residual G_N is kept
// per frame (tiny) and fused at finish by the deviation/entropy Pareto rule;
// the collapse then adds each level to the EXPAND of the next.
//
// Every f32 operation is in the CPU path's order (no FMA contraction), so the
// fold is bit-exact with it.
#include <algorithm>

#include "api.hpp"
#include "common.hpp"
#include "engine.hpp"
#include "fuse_tiles.hpp"

namespace lsh {

// REDUCE of one plane (blockIdx.z selects the plane) on 64 x 8 output tiles
// (the 19-row input window re-reads 19/16 of the level where 64 x 4 tiles
// re-read 11/8; 17 % faster on a 45 MP level)
constexpr int RED3_TY = 8;
__global__ __launch_bounds__(256) void k_reduce3(const float *__restrict__ src, int w, int h,
                                                 float *__restrict__ dst, int ow, int oh) {
    __shared__ ReduceLdsT<RED3_TY> S;
    reduce_tile<true>(PlaneSrc{src, (size_t)w * h}, (int)blockIdx.z, w, h, dst + (size_t)blockIdx.z * ow * oh, ow, oh,
                blockIdx.x * RED_TX, blockIdx.y * RED3_TY, S);
}

// one band-pass level of the fold on 64 x FOLD_TY tiles.  8 rows: a tile's LDS
// (22.5 KB at radius 1) then fits on a CU beside the search's block, so in
// the pipelined stack the fold also runs on the search's CUs (100-frame 45 MP
// step 682.4 -> 679.2 ms; alone it runs as fast as with 16 rows)
#ifndef FOLD_TY
#define FOLD_TY 8
#endif
template <int R>
__global__ __launch_bounds__(256) void k_fold(const float *__restrict__ G, const float *__restrict__ Cn, int w, int h,
                                              int cw, int ch, float *__restrict__ best, float *__restrict__ acc,
                                              uint16_t *__restrict__ winner, int first, unsigned idx, Wt9 wt,
                                              int chroma, float *__restrict__ en_out,
                                              const unsigned *__restrict__ idx_dev) {
    __shared__ FoldLds<R, FOLD_TY> S;
    fold_tile<R, FOLD_TY>(PlaneSrc{G, (size_t)w * h}, Cn, w, h, cw, ch, best, acc, winner, first,
                          idx_dev ? *idx_dev : idx, wt, chroma, en_out, blockIdx.x * 64, blockIdx.y * FOLD_TY, S);
}

// L = G - EXPAND(next) (pyramid.rs:208-211), or the collapse step out = L + EXPAND(next)
// (pyramid.rs:221-226) with the final clamp to [0,1] (fuse.rs:515-517).
__global__ __launch_bounds__(256) void k_expand_op(const float *__restrict__ L, const float *__restrict__ Cn, int w,
                                                   int h, int cw, int ch, float *__restrict__ out, int mode) {
    int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= w || y >= h) return;
    size_t n = (size_t)w * h, cn = (size_t)cw * ch, o = (size_t)y * w + x;
    int c = blockIdx.z;
    float e = expand_at(Cn + c * cn, cw, ch, y, x);
    float v;
    if (mode == 0) {
        v = L[c * n + o] - e;
    } else {
        v = L[c * n + o] + e;
        if (mode == 2) v = rclampf(v, 0.0f, 1.0f);
    }
    out[c * n + o] = v;
}

// fuse.rs:127-129 halo weight
__global__ void k_halo_weight(const float *__restrict__ re, size_t n, float p, float *__restrict__ out) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = expf(rclampf(p * logf((re[i] + 1e-8f) / 1e-4f), -60.0f, 60.0f));
}

// fuse.rs:461-471: the weighted coarse levels
__global__ void k_halo_acc(const float *__restrict__ lap, const float *__restrict__ wgt, size_t n, int first,
                           float *__restrict__ acc, float *__restrict__ best) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float k = wgt[i];
    if (first) {
        best[i] = k;
        for (int c = 0; c < 3; c++) acc[c * n + i] = lap[c * n + i] * k;
    } else {
        best[i] += k;
        for (int c = 0; c < 3; c++) acc[c * n + i] += k * lap[c * n + i];
    }
}

// fuse.rs:505-510
__global__ void k_halo_norm(float *__restrict__ acc, const float *__restrict__ best, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int c = 0; c < 3; c++) acc[c * n + i] /= best[i];
}

// fuse.rs:207-245: local deviation and entropy of one residual's luma, for
// every (frame, cell).  tab[k] = p*ln(p) with p = k/n, made on the host by the
// same f32 arithmetic; the sum runs in the CPU's first-occurrence order.
__global__ void k_devent(const float *__restrict__ tops, int nf, int tw, int th, int r, float qscale,
                         const float *__restrict__ tab, float *__restrict__ dev, float *__restrict__ ent) {
    size_t tsz = (size_t)tw * th;
    size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= tsz * nf) return;
    int f = (int)(gid / tsz);
    size_t cell = gid - (size_t)f * tsz;
    int i = (int)(cell / tw), j = (int)(cell - (size_t)i * tw);
    const float *tp = tops + (size_t)f * 3 * tsz;
    int nwin = (2 * r + 1) * (2 * r + 1);
    float inv_n = 1.0f / (float)nwin;
    float s = 0.0f, s2 = 0.0f;
    int q[81];
    int k = 0;
    for (int dy = -r; dy <= r; dy++) {
        int yy = rf101(i + dy, th);
        for (int dx = -r; dx <= r; dx++) {
            size_t o = (size_t)yy * tw + rf101(j + dx, tw);
            float v = 0.299f * tp[o] + 0.587f * tp[tsz + o] + 0.114f * tp[2 * tsz + o];
            s += v;
            s2 += v * v;
            float qv = rclampf(v, 0.0f, 1.0f) * qscale + 0.5f;
            q[k++] = qv > 0.0f ? (int)qv : 0;
        }
    }
    float mean = s * inv_n;
    float d = s2 * inv_n - mean * mean;
    dev[gid] = fmaxf(d, 0.0f);
    float e = 0.0f;
    for (int a = 0; a < nwin; a++) {
        bool seen = false;
        for (int b = 0; b < a; b++)
            if (q[b] == q[a]) {
                seen = true;
                break;
            }
        if (seen) continue;
        int cntq = 0;
        for (int b = a; b < nwin; b++) cntq += q[b] == q[a];
        e -= tab[cntq];
    }
    ent[gid] = e;
}

// fuse.rs:248-306: Pareto (or deviation-only, or average) fusion of the residuals
// One block per cell: every frame's (deviation, entropy) of the cell staged in
// LDS, thread a decides whether frame a is dominated (fuse.rs:283-296), then
// three threads (one per channel) add the kept frames' coefficients in frame
// order, as the reference does.
#define RES_MAXF 4096
__global__ __launch_bounds__(128) void k_residual(const float *__restrict__ tops, const float *__restrict__ dev,
                                                  const float *__restrict__ ent, int nf, size_t tsz, int rule,
                                                  float *__restrict__ out) {
    __shared__ float D[RES_MAXF], E[RES_MAXF];
    __shared__ unsigned char keep[RES_MAXF];
    __shared__ int kept_n;
    const size_t i = blockIdx.x;
    const bool all = nf == 1 || rule == 2;
    if (!all) {
        for (int a = threadIdx.x; a < nf; a += blockDim.x) {
            D[a] = dev[(size_t)a * tsz + i];
            E[a] = ent[(size_t)a * tsz + i];
        }
        __syncthreads();
        const bool use_entropy = rule == 0;
        for (int a = threadIdx.x; a < nf; a += blockDim.x) {
            const float da = D[a], ea = E[a];
            bool dominated = false;
            for (int b = 0; b < nf && !dominated; b++) {
                if (a == b) continue;
                const float db = D[b], eb = E[b];
                dominated = use_entropy ? (db >= da && eb >= ea && (db > da || eb > ea)) : (db > da || (db == da && b < a));
            }
            keep[a] = !dominated;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            int k = 0;
            for (int a = 0; a < nf; a++) k += keep[a];
            kept_n = k;
        }
        __syncthreads();
    }
    if (threadIdx.x < 3) {
        const int c = threadIdx.x;
        float o = 0.0f;
        if (all) {
            const float inv = 1.0f / (float)nf;
            for (int t = 0; t < nf; t++) o += tops[(size_t)t * 3 * tsz + (size_t)c * tsz + i] * inv;
        } else {
            const float inv = 1.0f / (float)(kept_n > 1 ? kept_n : 1);
            for (int a = 0; a < nf; a++)
                if (keep[a]) o += tops[(size_t)a * 3 * tsz + (size_t)c * tsz + i] * inv;
        }
        out[(size_t)c * tsz + i] = o;
    }
}

// fuse.rs:336-346
__global__ void k_upsample_index(const uint16_t *__restrict__ win, int dw, int dh, int w, int h, int level,
                                 float *__restrict__ out) {
    int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= w || y >= h) return;
    int sy = min(y >> level, dh - 1), sx = min(x >> level, dw - 1);
    out[(size_t)y * w + x] = (float)win[(size_t)sy * dw + sx];
}

// fuse.rs:315-329: the largest winning energy, then per-frame counts of detail cells
__global__ void k_max_nonneg(const float *__restrict__ v, size_t n, unsigned *__restrict__ out) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned m = 0;
    for (; i < n; i += (size_t)gridDim.x * blockDim.x) {
        float f = v[i];
        if (f > 0.0f) m = max(m, __float_as_uint(f));
    }
    for (int o = 32; o > 0; o >>= 1) m = max(m, (unsigned)__shfl_down((int)m, o, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

__global__ void k_share_count(const uint16_t *__restrict__ win, const float *__restrict__ best, size_t n,
                              const unsigned *__restrict__ maxbits, int frames, unsigned *__restrict__ counts) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    float floor_ = 0.01f * __uint_as_float(*maxbits);
    for (; i < n; i += (size_t)gridDim.x * blockDim.x) {
        if (best[i] > floor_ && floor_ > 0.0f) {
            int k = (int)win[i];
            if (k < frames) atomicAdd(&counts[k], 1u);
        }
    }
}

static void fold_dispatch(int r, dim3 g, const float *G, const float *Cn, int w, int h, int cw, int ch, float *best,
                          float *acc, uint16_t *winner, int first, unsigned idx, const Wt9 &wt, int chroma,
                          float *en_out, const unsigned *idx_dev) {
    dim3 b(256);
    // G_l and G_{l+1} read, best read and written, the accumulator written (fuse.rs:414-451)
    double by = 32.0 * w * h + 12.0 * cw * ch;
    switch (r) {
    case 0: launch_b("fold", by, k_fold<0>, g, b, 0, G, Cn, w, h, cw, ch, best, acc, winner, first, idx, wt, chroma, en_out, idx_dev); break;
    case 1: launch_b("fold", by, k_fold<1>, g, b, 0, G, Cn, w, h, cw, ch, best, acc, winner, first, idx, wt, chroma, en_out, idx_dev); break;
    case 2: launch_b("fold", by, k_fold<2>, g, b, 0, G, Cn, w, h, cw, ch, best, acc, winner, first, idx, wt, chroma, en_out, idx_dev); break;
    case 3: launch_b("fold", by, k_fold<3>, g, b, 0, G, Cn, w, h, cw, ch, best, acc, winner, first, idx, wt, chroma, en_out, idx_dev); break;
    case 4: launch_b("fold", by, k_fold<4>, g, b, 0, G, Cn, w, h, cw, ch, best, acc, winner, first, idx, wt, chroma, en_out, idx_dev); break;
    default: throw Error("energy_radius above 4 is not supported");
    }
}

// fuse.rs:139-149
static void binomial(size_t r, float *out) {
    size_t n = 2 * r + 1;
    std::vector<uint64_t> row(n, 1);
    for (size_t i = 1; i < n; i++)
        for (size_t j = i - 1; j >= 1; j--) row[j] += row[j - 1];
    uint64_t s = 0;
    for (auto v : row) s += v;
    float sf = (float)s;
    for (size_t i = 0; i < n; i++) out[i] = (float)row[i] / sf;
}

static size_t half_(size_t n) { return (n + 1) / 2; }

// pyramid.rs:191-200
static size_t auto_levels(size_t w, size_t h, size_t min_top) {
    size_t n = 0, cw = w, ch = h, mt = std::max<size_t>(min_top, 1);
    while ((cw > 1 || ch > 1) && std::min(half_(cw), half_(ch)) >= mt) {
        cw = half_(cw);
        ch = half_(ch);
        n++;
    }
    return std::max<size_t>(n, 1);
}

// fuse.rs:401-484 Fuser::push with level 0 at g0 (3 planes on the device)
// the residual level of the frame into tops at the frame index in device memory
__global__ void k_tops_put(const float *__restrict__ top, size_t m, const unsigned *__restrict__ idx_dev,
                           float *__restrict__ tops) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) tops[(size_t)(*idx_dev) * m + i] = top[i];
}

static void fuser_push(lsh_fuser *f, const float *g0, const unsigned *idx_dev = nullptr) {
    LSH_REQUIRE(!f->finished, "fuser: push after finish");
    LSH_REQUIRE(f->count < 65535, "fuser: more than 65535 frames");
    const size_t L = f->levels;
    const bool first = f->count == 0;
    const unsigned idx = (unsigned)f->count;
    const size_t nsel = f->halo_on ? f->guide + 1 : L;
    auto lvl = [&](size_t l) -> const float * { return l == 0 ? g0 : f->G.p + 3 * f->off[l]; };
    // the Gaussian levels 1..L
    for (size_t l = 0; l < L; l++) {
        dim3 g(cdiv(f->lw[l + 1], 64), cdiv(f->lh[l + 1], RED3_TY), 3);
        double bytes = 12.0 * (f->lw[l] * f->lh[l] + f->lw[l + 1] * f->lh[l + 1]);
        launch_b("reduce", bytes, k_reduce3, g, dim3(256), 0, lvl(l), (int)f->lw[l], (int)f->lh[l],
                 f->G.p + 3 * f->off[l + 1], (int)f->lw[l + 1], (int)f->lh[l + 1]);
    }
    if (first) f->winner.zero();
    Wt9 wt{};
    binomial(f->p.energy_radius, wt.k);
    for (size_t l = 0; l < nsel; l++) {
        dim3 g(cdiv(f->lw[l], 64), cdiv(f->lh[l], FOLD_TY));
        bool is_depth = l == f->depth_level;
        float *en_out = (f->halo_on && l == f->guide) ? f->en_guide.p : nullptr;
        fold_dispatch((int)f->p.energy_radius, g, lvl(l), lvl(l + 1), (int)f->lw[l], (int)f->lh[l], (int)f->lw[l + 1],
                      (int)f->lh[l + 1], f->best.p + f->off[l], f->acc.p + 3 * f->off[l],
                      is_depth ? f->winner.p : nullptr, first ? 1 : 0, idx, wt, f->p.use_chroma ? 1 : 0, en_out,
                      idx_dev);
    }
    if (f->halo_on) {
        size_t g = f->guide, n = f->lw[g] * f->lh[g];
        launch("halo_weight", k_halo_weight, dim3(cdiv(n, 256)), dim3(256), 0, (const float *)f->en_guide.p, n, f->hard,
               f->wgt_a.p);
        float *wa = f->wgt_a.p, *wb = f->wgt_b.p;
        for (size_t li = g + 1; li <= L; li++) {
            dim3 gr(cdiv(f->lw[li], 64), cdiv(f->lh[li], RED3_TY), 1);
            launch("reduce", k_reduce3, gr, dim3(256), 0, (const float *)wa, (int)f->lw[li - 1], (int)f->lh[li - 1], wb,
                   (int)f->lw[li], (int)f->lh[li]);
            std::swap(wa, wb);
            size_t nl = f->lw[li] * f->lh[li];
            const float *lap = lvl(li);
            if (li < L) {
                dim3 ge(cdiv(f->lw[li], 64), cdiv(f->lh[li], 4), 3);
                launch("laplacian", k_expand_op, ge, dim3(64, 4), 0, lvl(li), lvl(li + 1), (int)f->lw[li], (int)f->lh[li],
                       (int)f->lw[li + 1], (int)f->lh[li + 1], f->lap.p, 0);
                lap = f->lap.p;
            }
            launch("halo_acc", k_halo_acc, dim3(cdiv(nl, 256)), dim3(256), 0, lap, (const float *)wa, nl, first ? 1 : 0,
                   f->acc.p + 3 * f->off[li], f->best.p + f->off[li]);
        }
    } else {
        size_t tsz = f->lw[L] * f->lh[L];
        if (idx_dev) {
            LSH_REQUIRE(f->count + 1 <= f->tops_cap, "fuser: residual store not reserved for the stream");
            launch("tops_put", k_tops_put, dim3(cdiv(3 * tsz, 256)), dim3(256), 0, lvl(L), 3 * tsz, idx_dev, f->tops.p);
            f->count++;
            return;
        }
        if (f->count + 1 > f->tops_cap) {
            size_t cap = std::max<size_t>(16, 2 * f->tops_cap);
            DBuf<float> nt(cap * 3 * tsz);
            if (f->count) LSH_CHECK(hipMemcpyAsync(nt.p, f->tops.p, f->count * 3 * tsz * sizeof(float), hipMemcpyDeviceToDevice, stream()));
            f->tops = std::move(nt);
            f->tops_cap = cap;
        }
        LSH_CHECK(hipMemcpyAsync(f->tops.p + f->count * 3 * tsz, lvl(L), 3 * tsz * sizeof(float), hipMemcpyDeviceToDevice,
                                 stream()));
    }
    f->count++;
}

}  // namespace lsh

using namespace lsh;

void lsh_fuser_push_device(lsh_fuser *f, const float *planes) { fuser_push(f, planes); }

namespace lsh {
void fuser_reserve(lsh_fuser *f, size_t nframes) {
    if (f->halo_on || nframes <= f->tops_cap) return;
    size_t tsz = f->lw[f->levels] * f->lh[f->levels];
    DBuf<float> nt(nframes * 3 * tsz);
    if (f->count) copy_mem(nt.p, LSH_DEVICE, f->tops.p, LSH_DEVICE, f->count * 3 * tsz * sizeof(float));
    sync();
    f->tops = std::move(nt);
    f->tops_cap = nframes;
}
void fuser_push_dev(lsh_fuser *f, const float *g0, const unsigned *idx_dev) { fuser_push(f, g0, idx_dev); }
}  // namespace lsh

extern "C" {

char *lsh_fuser_new(uint32_t w, uint32_t h, const lsh_fuse_params *p, lsh_fuser **out) {
    LSH_API_BEGIN
    LSH_REQUIRE(out && p, "lsh_fuser_new: null argument");
    *out = nullptr;
    LSH_REQUIRE(w > 0 && h > 0, "lsh_fuser_new: empty frame");
    LSH_REQUIRE(p->energy_radius <= 4, "lsh_fuser_new: energy_radius above 4 is not supported");
    LSH_REQUIRE(p->top_radius <= 4, "lsh_fuser_new: top_radius above 4 is not supported");
    LSH_REQUIRE(p->top_rule <= 2, "lsh_fuser_new: unknown top rule");
    auto f = new lsh_fuser();
    try {
        f->w = w;
        f->h = h;
        f->p = *p;
        size_t lv = p->levels > 0 ? (size_t)p->levels : auto_levels(w, h, 32);
        f->levels = std::max<size_t>(lv, 1);
        LSH_REQUIRE(f->levels < 30, "lsh_fuser_new: too many levels");
        f->depth_level = std::min<size_t>(p->depth_level, f->levels - 1);
        f->halo_on = p->halo > 0.0f;
        f->guide = f->depth_level;
        f->hard = std::min(p->halo, 8.0f);
        size_t cw = w, ch = h;
        for (size_t l = 0; l <= f->levels; l++) {
            f->lw.push_back(cw);
            f->lh.push_back(ch);
            f->off.push_back(f->S);
            f->S += cw * ch;
            cw = half_(cw);
            ch = half_(ch);
        }
        f->G.alloc(3 * f->S);
        f->acc.alloc(3 * f->S);
        f->best.alloc(f->S);
        f->winner.alloc(f->lw[f->depth_level] * f->lh[f->depth_level]);
        if (f->halo_on) {
            size_t ng = f->lw[f->guide] * f->lh[f->guide];
            f->en_guide.alloc(ng);
            f->wgt_a.alloc(ng);
            f->wgt_b.alloc(ng);
            f->lap.alloc(3 * f->lw[std::min(f->guide + 1, f->levels)] * f->lh[std::min(f->guide + 1, f->levels)]);
        }
        // p*ln(p) per count of a (2r+1)^2 window (fuse.rs:236-240)
        size_t nwin = (2 * p->top_radius + 1) * (2 * p->top_radius + 1);
        std::vector<float> tab(nwin + 1, 0.0f);
        float inv_n = 1.0f / (float)nwin;
        for (size_t k = 1; k <= nwin; k++) {
            float pk = (float)k * inv_n;
            tab[k] = pk * std::log(pk);
        }
        f->ent_tab.alloc(nwin + 1);
        copy_mem(f->ent_tab.p, LSH_DEVICE, tab.data(), LSH_HOST, tab.size() * sizeof(float));
    } catch (...) {
        delete f;
        throw;
    }
    *out = f;
    LSH_API_END
}

char *lsh_fuser_levels(lsh_fuser *f, uint32_t *levels, uint32_t *depth_level, uint32_t *dw, uint32_t *dh) {
    LSH_API_BEGIN
    LSH_REQUIRE(f, "lsh_fuser_levels: null fuser");
    if (levels) *levels = (uint32_t)f->levels;
    if (depth_level) *depth_level = (uint32_t)f->depth_level;
    if (dw) *dw = (uint32_t)f->lw[f->depth_level];
    if (dh) *dh = (uint32_t)f->lh[f->depth_level];
    LSH_API_END
}

char *lsh_fuser_push(lsh_fuser *f, const float *img3, lsh_mem mem) {
    LSH_API_BEGIN
    LSH_REQUIRE(f && img3, "lsh_fuser_push: null argument");
    if (mem == LSH_DEVICE) {
        fuser_push(f, img3);
    } else {
        copy_mem(f->G.p, LSH_DEVICE, img3, LSH_HOST, 3 * f->w * f->h * sizeof(float));
        fuser_push(f, f->G.p);
    }
    LSH_API_END
}

char *lsh_fuser_push_frames(lsh_fuser *f, lsh_frames *src) {
    LSH_API_BEGIN
    LSH_REQUIRE(f && src, "lsh_fuser_push_frames: null argument");
    LSH_REQUIRE(src->w == f->w && src->h == f->h, "lsh_fuser_push_frames: frame size mismatch");
    LSH_REQUIRE(src->processed > 0, "lsh_fuser_push_frames: the frames handle holds no frame");
    fuser_push(f, src->planes.p);
    LSH_API_END
}

char *lsh_fuser_shares(lsh_fuser *f, float *out, uint32_t cap, uint32_t *n) {
    LSH_API_BEGIN
    LSH_REQUIRE(f && n, "lsh_fuser_shares: null argument");
    *n = 0;
    size_t nsel = f->halo_on ? f->guide + 1 : f->levels;
    if (f->count == 0 || f->depth_level >= nsel || f->finished) return nullptr;
    LSH_REQUIRE(out && cap >= f->count, "lsh_fuser_shares: output too small");
    size_t d = f->depth_level, cells = f->lw[d] * f->lh[d];
    f->cnt.ensure(f->count + 1);
    LSH_CHECK(hipMemsetAsync(f->cnt.p, 0, (f->count + 1) * sizeof(unsigned), stream()));
    unsigned *mx = f->cnt.p + f->count;
    unsigned grid = std::min<unsigned>(cdiv(cells, 256), 2048);
    launch("shares_max", k_max_nonneg, dim3(grid), dim3(256), 0, (const float *)(f->best.p + f->off[d]), cells, mx);
    launch("shares_count", k_share_count, dim3(grid), dim3(256), 0, (const uint16_t *)f->winner.p,
           (const float *)(f->best.p + f->off[d]), cells, (const unsigned *)mx, (int)f->count, f->cnt.p);
    std::vector<unsigned> counts(f->count);
    copy_mem(counts.data(), LSH_HOST, f->cnt.p, LSH_DEVICE, f->count * sizeof(unsigned));
    size_t total = 0;
    for (auto c : counts) total += c;
    for (size_t k = 0; k < f->count; k++) out[k] = total > 0 ? (float)counts[k] / (float)total : 0.0f;
    *n = (uint32_t)f->count;
    LSH_API_END
}

char *lsh_fuser_finish(lsh_fuser *f, float *img3, lsh_mem img_mem, float *winner_full, lsh_mem win_mem,
                       uint16_t *winner_level) {
    LSH_API_BEGIN
    LSH_REQUIRE(f && img3, "lsh_fuser_finish: null argument");
    LSH_REQUIRE(f->count > 0, "no frames pushed");
    LSH_REQUIRE(!f->finished, "lsh_fuser_finish: already finished");
    const size_t L = f->levels, d = f->depth_level, N = f->w * f->h;
    // the residual level into acc's level L
    float *top = f->acc.p + 3 * f->off[L];
    size_t tsz = f->lw[L] * f->lh[L];
    if (f->halo_on) {
        for (size_t li = f->guide + 1; li <= L; li++) {
            size_t nl = f->lw[li] * f->lh[li];
            launch("halo_norm", k_halo_norm, dim3(cdiv(nl, 256)), dim3(256), 0, f->acc.p + 3 * f->off[li],
                   (const float *)(f->best.p + f->off[li]), nl);
        }
    } else {
        int nf = (int)f->count;
        DBuf<float> dev, ent;
        bool pareto = nf > 1 && f->p.top_rule != 2;
        if (pareto) {
            dev.alloc(tsz * nf);
            ent.alloc(tsz * nf);
            size_t bins = std::min<size_t>(std::max<size_t>(f->p.entropy_bins, 2), 65536);
            launch_b("residual_devent", 4.0 * (double)tsz * nf * 5.0, k_devent, dim3(cdiv(tsz * nf, 128)), dim3(128), 0, (const float *)f->tops.p, nf,
                   (int)f->lw[L], (int)f->lh[L], (int)f->p.top_radius, (float)(bins - 1), (const float *)f->ent_tab.p,
                   dev.p, ent.p);
        }
        LSH_REQUIRE(nf <= RES_MAXF, "residual rule: more than 4096 frames");
        launch_b("residual_fuse", 4.0 * (double)tsz * (5.0 * nf + 3.0), k_residual, dim3((unsigned)tsz), dim3(128), 0,
                 (const float *)f->tops.p, (const float *)dev.p, (const float *)ent.p, nf, tsz, (int)f->p.top_rule, top);
        sync();
    }
    // collapse (pyramid.rs:219-229): G_l = acc_l + EXPAND(G_{l+1}); level 0 clamped straight into the output
    DBuf<float> out0;
    float *dst0 = img3;
    if (img_mem == LSH_HOST) {
        out0.alloc(3 * N);
        dst0 = out0.p;
    }
    const float *next = top;
    for (size_t l = L; l-- > 0;) {
        float *dst = l == 0 ? dst0 : f->G.p + 3 * f->off[l];
        dim3 g(cdiv(f->lw[l], 64), cdiv(f->lh[l], 4), 3);
        launch_b("collapse", 24.0 * f->lw[l] * f->lh[l] + 12.0 * f->lw[l + 1] * f->lh[l + 1], k_expand_op, g,
                 dim3(64, 4), 0, (const float *)(f->acc.p + 3 * f->off[l]), next, (int)f->lw[l], (int)f->lh[l],
                 (int)f->lw[l + 1], (int)f->lh[l + 1], dst, l == 0 ? 2 : 1);
        next = dst;
    }
    if (img_mem == LSH_HOST) copy_mem(img3, LSH_HOST, out0.p, LSH_DEVICE, 3 * N * sizeof(float));
    if (winner_full) {
        DBuf<float> wf;
        float *dst = winner_full;
        if (win_mem == LSH_HOST) {
            wf.alloc(N);
            dst = wf.p;
        }
        dim3 g(cdiv(f->w, 64), cdiv(f->h, 4));
        launch_b("winner_upsample", 2.0 * f->lw[d] * f->lh[d] + 4.0 * N, k_upsample_index, g, dim3(64, 4), 0, (const uint16_t *)f->winner.p, (int)f->lw[d],
               (int)f->lh[d], (int)f->w, (int)f->h, (int)d, dst);
        if (win_mem == LSH_HOST) copy_mem(winner_full, LSH_HOST, wf.p, LSH_DEVICE, N * sizeof(float));
    }
    if (winner_level) copy_mem(winner_level, LSH_HOST, f->winner.p, LSH_DEVICE, f->lw[d] * f->lh[d] * sizeof(uint16_t));
    sync();
    f->finished = true;
    LSH_API_END
}

void lsh_fuser_free(lsh_fuser *f) {
    if (!f) return;
    try {
        sync();
    } catch (...) {
    }
    delete f;
}

}  // extern "C"
