/*
 * lsh_oracle.c -- CPU restatement of lapstack-core's stacking-engine path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the CPU baseline); see
 * lsh_oracle.h.  Citations are to crates/lapstack-core/src/ of the reference.
 * Build: oracle/Makefile (-O3 -ffp-contract=off -fopenmp, no fast-math).
 *
 * Parallel loops run over rows (OpenMP in place of rayon's for_rows,
 * pyramid.rs:68-77); every f32 result is independent of the thread count.
 * f64 reductions (warp cost, CG dots, brightness sums) fold per-row partials
 * in row order; the reference's rayon tree order differs at the last ulp,
 * which no test depends on.
 */
#include "lsh_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define PAR_MIN (1 << 15)
#define MAXL 40

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + ts.tv_nsec * 1e-9;
}

static void *xcalloc(size_t n, size_t sz) {
    void *p = calloc(n ? n : 1, sz);
    if (!p) abort();
    return p;
}
static float *fzeros(size_t n) { return (float *)xcalloc(n, sizeof(float)); }

/* Rust f32::max / f32::min: the non-NaN operand wins */
static inline float rmaxf(float a, float b) { return fmaxf(a, b); }
static inline float rminf(float a, float b) { return fminf(a, b); }
/* Rust f32::clamp */
static inline float rclampf(float v, float lo, float hi) {
    if (v < lo) v = lo;
    if (v > hi) v = hi;
    return v;
}
static inline double rclamp(double v, double lo, double hi) {
    if (v < lo) v = lo;
    if (v > hi) v = hi;
    return v;
}
/* Rust `x as usize` from f32: saturating, NaN -> 0 */
static inline size_t f2usize(float v) {
    if (!(v > 0.0f)) return 0;
    if (v >= 18446744073709551615.0f) return (size_t)-1;
    return (size_t)v;
}
/* f32::total_cmp */
static inline int32_t total_key(float v) {
    int32_t b;
    memcpy(&b, &v, 4);
    return b ^ (int32_t)(((uint32_t)(b >> 31)) >> 1);
}
static int cmp_total(const void *a, const void *b) {
    int32_t x = total_key(*(const float *)a), y = total_key(*(const float *)b);
    return (x > y) - (x < y);
}

/* ======================================================================
 * pyramid.rs
 * ==================================================================== */

static const float KERNEL[5] = {1.0f / 16.0f, 4.0f / 16.0f, 6.0f / 16.0f, 4.0f / 16.0f, 1.0f / 16.0f};
/* pyramid.rs:124-125: 2*[K0, K2, K4] and 2*K1 */
static const float EVEN[3] = {2.0f * (1.0f / 16.0f), 2.0f * (6.0f / 16.0f), 2.0f * (1.0f / 16.0f)};
static const float ODD = 2.0f * (4.0f / 16.0f);

/* pyramid.rs:51-66 */
size_t ora_reflect(long i, size_t n) {
    if (n == 1) return 0;
    long nn = (long)n;
    for (;;) {
        if (i < 0)
            i = -i;
        else if (i >= nn)
            i = 2 * (nn - 1) - i;
        else
            return (size_t)i;
    }
}
#define RF(i, n) ora_reflect((long)(i), (n))

static inline size_t half(size_t n) { return (n + 1) / 2; }

/* pyramid.rs:86-119 */
void ora_reduce(const float *src, size_t w, size_t h, float *out) {
    size_t ow = half(w), oh = half(h);
    float *tmp = fzeros(h * ow);
#pragma omp parallel for schedule(static) if (h * ow >= PAR_MIN)
    for (size_t y = 0; y < h; y++) {
        const float *s = src + y * w;
        for (size_t i = 0; i < ow; i++) {
            long c = 2 * (long)i;
            float a = 0.0f;
            for (int t = 0; t < 5; t++) a += KERNEL[t] * s[RF(c + t - 2, w)];
            tmp[y * ow + i] = a;
        }
    }
#pragma omp parallel for schedule(static) if (oh * ow >= PAR_MIN)
    for (size_t y = 0; y < oh; y++) {
        long c = 2 * (long)y;
        const float *r[5];
        for (int t = 0; t < 5; t++) r[t] = tmp + RF(c + t - 2, h) * ow;
        for (size_t j = 0; j < ow; j++)
            out[y * ow + j] = KERNEL[0] * r[0][j] + KERNEL[1] * r[1][j] + KERNEL[2] * r[2][j] +
                              KERNEL[3] * r[3][j] + KERNEL[4] * r[4][j];
    }
    free(tmp);
}

/* pyramid.rs:128-140 */
static void expand_row(const float *coarse, size_t cw, float *out, size_t ow) {
    for (size_t i = 0; i < ow; i++) {
        long c = (long)(i / 2);
        if (i % 2 == 0)
            out[i] = EVEN[0] * coarse[RF(c - 1, cw)] + EVEN[1] * coarse[RF(c, cw)] + EVEN[2] * coarse[RF(c + 1, cw)];
        else
            out[i] = ODD * (coarse[RF(c, cw)] + coarse[RF(c + 1, cw)]);
    }
}

/* pyramid.rs:143-168 */
void ora_expand(const float *coarse, size_t cw, size_t ch, size_t ow, size_t oh, float *out) {
    float *tmp = fzeros(ch * ow);
#pragma omp parallel for schedule(static) if (ch * ow >= PAR_MIN)
    for (size_t y = 0; y < ch; y++) expand_row(coarse + y * cw, cw, tmp + y * ow, ow);
#pragma omp parallel for schedule(static) if (oh * ow >= PAR_MIN)
    for (size_t i = 0; i < oh; i++) {
        long c = (long)(i / 2);
        float *row = out + i * ow;
        if (i % 2 == 0) {
            const float *ra = tmp + RF(c - 1, ch) * ow, *rb = tmp + RF(c, ch) * ow, *rd = tmp + RF(c + 1, ch) * ow;
            for (size_t j = 0; j < ow; j++) row[j] = EVEN[0] * ra[j] + EVEN[1] * rb[j] + EVEN[2] * rd[j];
        } else {
            const float *ra = tmp + RF(c, ch) * ow, *rb = tmp + RF(c + 1, ch) * ow;
            for (size_t j = 0; j < ow; j++) row[j] = ODD * (ra[j] + rb[j]);
        }
    }
    free(tmp);
}

/* pyramid.rs:191-200 */
size_t ora_auto_levels(size_t w, size_t h, size_t min_top) {
    size_t n = 0, cw = w, ch = h;
    size_t mt = min_top > 1 ? min_top : 1;
    while ((cw > 1 || ch > 1) && (half(cw) < half(ch) ? half(cw) : half(ch)) >= mt) {
        cw = half(cw);
        ch = half(ch);
        n++;
    }
    return n > 1 ? n : 1;
}

size_t ora_pyr_size(size_t w, size_t h, size_t levels) {
    size_t s = 0;
    for (size_t l = 0; l <= levels; l++) {
        s += 3 * w * h;
        w = half(w);
        h = half(h);
    }
    return s;
}

/* A pyramid held as per-level plane pointers. */
typedef struct {
    size_t n; /* levels + 1 */
    size_t w[MAXL], h[MAXL];
    float *p[MAXL][3];
} pyr_t;

static void pyr_free(pyr_t *P) {
    for (size_t l = 0; l < P->n; l++)
        for (int c = 0; c < 3; c++) free(P->p[l][c]);
    P->n = 0;
}

/* pyramid.rs:203-216 */
static void pyr_build(const float *img3, size_t w, size_t h, size_t levels, pyr_t *P) {
    P->n = levels + 1;
    size_t n = w * h;
    float *g[3];
    for (int c = 0; c < 3; c++) {
        g[c] = fzeros(n);
        memcpy(g[c], img3 + c * n, n * sizeof(float));
    }
    size_t cw = w, ch = h;
    for (size_t l = 0; l < levels; l++) {
        size_t ow = half(cw), oh = half(ch);
        float *next[3];
        for (int c = 0; c < 3; c++) {
            next[c] = fzeros(ow * oh);
            ora_reduce(g[c], cw, ch, next[c]);
            float *up = fzeros(cw * ch);
            ora_expand(next[c], ow, oh, cw, ch, up);
            for (size_t i = 0; i < cw * ch; i++) g[c][i] -= up[i];
            free(up);
        }
        P->w[l] = cw;
        P->h[l] = ch;
        for (int c = 0; c < 3; c++) {
            P->p[l][c] = g[c];
            g[c] = next[c];
        }
        cw = ow;
        ch = oh;
    }
    P->w[levels] = cw;
    P->h[levels] = ch;
    for (int c = 0; c < 3; c++) P->p[levels][c] = g[c];
}

/* pyramid.rs:219-229: consumes P's planes */
static void pyr_collapse(pyr_t *P, float *img3) {
    size_t L = P->n - 1;
    float *g[3] = {P->p[L][0], P->p[L][1], P->p[L][2]};
    size_t gw = P->w[L], gh = P->h[L];
    for (long l = (long)L - 1; l >= 0; l--) {
        size_t lw = P->w[l], lh = P->h[l];
        for (int c = 0; c < 3; c++) {
            float *up = fzeros(lw * lh);
            ora_expand(g[c], gw, gh, lw, lh, up);
            float *lp = P->p[l][c];
            for (size_t i = 0; i < lw * lh; i++) lp[i] += up[i];
            free(up);
            free(g[c]);
            g[c] = lp;
        }
        gw = lw;
        gh = lh;
    }
    for (int c = 0; c < 3; c++) {
        memcpy(img3 + c * gw * gh, g[c], gw * gh * sizeof(float));
        free(g[c]);
    }
    P->n = 0;
}

void ora_pyr_build(const float *img3, size_t w, size_t h, size_t levels, float *packed) {
    pyr_t P;
    pyr_build(img3, w, h, levels, &P);
    size_t off = 0;
    for (size_t l = 0; l < P.n; l++)
        for (int c = 0; c < 3; c++) {
            memcpy(packed + off, P.p[l][c], P.w[l] * P.h[l] * sizeof(float));
            off += P.w[l] * P.h[l];
        }
    pyr_free(&P);
}

void ora_pyr_collapse(const float *packed, size_t w, size_t h, size_t levels, float *img3) {
    pyr_t P;
    P.n = levels + 1;
    size_t off = 0, cw = w, ch = h;
    for (size_t l = 0; l < P.n; l++) {
        P.w[l] = cw;
        P.h[l] = ch;
        for (int c = 0; c < 3; c++) {
            P.p[l][c] = fzeros(cw * ch);
            memcpy(P.p[l][c], packed + off, cw * ch * sizeof(float));
            off += cw * ch;
        }
        cw = half(cw);
        ch = half(ch);
    }
    pyr_collapse(&P, img3);
}

/* ======================================================================
 * fuse.rs
 * ==================================================================== */

#define HALO_REF 1e-4f
#define HALO_FLOOR 1e-8f
#define HALO_MAX 8.0f
#define DETAIL_FLOOR 0.01f

/* fuse.rs:127-129 */
float ora_halo_weight(float re, float p) { return expf(rclampf(p * logf((re + HALO_FLOOR) / HALO_REF), -60.0f, 60.0f)); }

/* fuse.rs:139-149 */
void ora_binomial(size_t r, float *out) {
    size_t n = 2 * r + 1;
    uint64_t *row = (uint64_t *)xcalloc(n, sizeof(uint64_t));
    for (size_t i = 0; i < n; i++) row[i] = 1;
    for (size_t i = 1; i < n; i++)
        for (size_t j = i - 1; j >= 1; j--) row[j] += row[j - 1];
    uint64_t s = 0;
    for (size_t i = 0; i < n; i++) s += row[i];
    float sf = (float)s;
    for (size_t i = 0; i < n; i++) out[i] = (float)row[i] / sf;
    free(row);
}

/* fuse.rs:152-165 */
static void node_energy(float *const l[3], size_t n, int use_chroma, float *e) {
    const float *r = l[0], *g = l[1], *b = l[2];
#pragma omp parallel for schedule(static) if (n >= PAR_MIN)
    for (size_t i = 0; i < n; i++) {
        if (use_chroma) {
            e[i] = r[i] * r[i] + g[i] * g[i] + b[i] * b[i];
        } else {
            float y = 0.299f * r[i] + 0.587f * g[i] + 0.114f * b[i];
            e[i] = y * y;
        }
    }
}

/* fuse.rs:168-197 */
void ora_window_sum(const float *src, size_t w, size_t h, const float *wt, size_t nwt, float *out) {
    size_t r = nwt / 2;
    if (r == 0) {
        memcpy(out, src, w * h * sizeof(float));
        return;
    }
    float *tmp = fzeros(w * h);
#pragma omp parallel for schedule(static) if (w * h >= PAR_MIN)
    for (size_t y = 0; y < h; y++) {
        const float *s = src + y * w;
        for (size_t j = 0; j < w; j++) {
            float a = 0.0f;
            for (size_t t = 0; t < nwt; t++) a += wt[t] * s[RF((long)j + (long)t - (long)r, w)];
            tmp[y * w + j] = a;
        }
    }
#pragma omp parallel for schedule(static) if (w * h >= PAR_MIN)
    for (size_t y = 0; y < h; y++) {
        float *row = out + y * w;
        for (size_t j = 0; j < w; j++) row[j] = 0.0f;
        for (size_t t = 0; t < nwt; t++) {
            const float *s = tmp + RF((long)y + (long)t - (long)r, h) * w;
            for (size_t j = 0; j < w; j++) row[j] += wt[t] * s[j];
        }
    }
    free(tmp);
}

static void region_energy_planes(float *const l[3], size_t w, size_t h, size_t radius, int use_chroma, float *out) {
    float *e = fzeros(w * h);
    node_energy(l, w * h, use_chroma, e);
    float wt[2 * 64 + 1];
    ora_binomial(radius, wt);
    ora_window_sum(e, w, h, wt, 2 * radius + 1, out);
    free(e);
}

/* fuse.rs:200-203 */
void ora_region_energy(const float *l3, size_t w, size_t h, size_t radius, int use_chroma, float *out) {
    float *const l[3] = {(float *)l3, (float *)l3 + w * h, (float *)l3 + 2 * w * h};
    region_energy_planes(l, w, h, radius, use_chroma, out);
}

/* fuse.rs:207-245 */
void ora_deviation_entropy(const float *y, size_t w, size_t h, size_t radius, size_t bins, float *dev, float *ent) {
    size_t n = (2 * radius + 1) * (2 * radius + 1);
    float inv_n = 1.0f / (float)n;
    if (bins < 2) bins = 2;
    if (bins > 65536) bins = 65536;
    float qscale = (float)(bins - 1);
    long r = (long)radius;
#pragma omp parallel for schedule(static) if (w * h >= PAR_MIN)
    for (size_t i = 0; i < h; i++) {
        uint32_t *hist = (uint32_t *)xcalloc(bins, sizeof(uint32_t));
        size_t *touched = (size_t *)xcalloc(n, sizeof(size_t));
        for (size_t j = 0; j < w; j++) {
            float s = 0.0f, s2 = 0.0f;
            size_t nt = 0;
            for (long dy = -r; dy <= r; dy++) {
                size_t yy = RF((long)i + dy, h);
                for (long dx = -r; dx <= r; dx++) {
                    float v = y[yy * w + RF((long)j + dx, w)];
                    s += v;
                    s2 += v * v;
                    size_t q = f2usize(rclampf(v, 0.0f, 1.0f) * qscale + 0.5f);
                    if (hist[q] == 0) touched[nt++] = q;
                    hist[q] += 1;
                }
            }
            float mean = s * inv_n;
            dev[i * w + j] = rmaxf(s2 * inv_n - mean * mean, 0.0f);
            float e = 0.0f;
            for (size_t k = 0; k < nt; k++) {
                size_t q = touched[k];
                float p = (float)hist[q] * inv_n;
                e -= p * logf(p);
                hist[q] = 0;
            }
            ent[i * w + j] = e;
        }
        free(hist);
        free(touched);
    }
}

/* fuse.rs:248-306; tops: n residuals of 3*w*h */
void ora_fuse_residuals(const float *tops, size_t n, size_t w, size_t h, const ora_fuse_params *p, float *out3) {
    size_t sz = w * h;
    memset(out3, 0, 3 * sz * sizeof(float));
    if (n == 1 || p->top_rule == 2) {
        float inv = 1.0f / (float)n;
        for (size_t t = 0; t < n; t++)
            for (int c = 0; c < 3; c++)
                for (size_t i = 0; i < sz; i++) out3[c * sz + i] += tops[t * 3 * sz + c * sz + i] * inv;
        return;
    }
    float *dev = fzeros(n * sz), *ent = fzeros(n * sz), *y = fzeros(sz);
    for (size_t t = 0; t < n; t++) {
        const float *tp = tops + t * 3 * sz;
        for (size_t i = 0; i < sz; i++) y[i] = 0.299f * tp[i] + 0.587f * tp[sz + i] + 0.114f * tp[2 * sz + i];
        ora_deviation_entropy(y, w, h, p->top_radius, p->entropy_bins, dev + t * sz, ent + t * sz);
    }
    int use_entropy = p->top_rule == 0;
    uint8_t *keep = (uint8_t *)xcalloc(n, 1);
    for (size_t i = 0; i < sz; i++) {
        size_t kept = 0;
        for (size_t a = 0; a < n; a++) {
            float da = dev[a * sz + i], ea = ent[a * sz + i];
            int dominated = 0;
            for (size_t b = 0; b < n; b++) {
                if (a == b) continue;
                float db = dev[b * sz + i], eb = ent[b * sz + i];
                int d = use_entropy ? (db >= da && eb >= ea && (db > da || eb > ea)) : (db > da || (db == da && b < a));
                if (d) {
                    dominated = 1;
                    break;
                }
            }
            keep[a] = !dominated;
            kept += !dominated;
        }
        float inv = 1.0f / (float)(kept > 1 ? kept : 1);
        for (size_t a = 0; a < n; a++)
            if (keep[a])
                for (int c = 0; c < 3; c++) out3[c * sz + i] += tops[a * 3 * sz + c * sz + i] * inv;
    }
    free(keep);
    free(dev);
    free(ent);
    free(y);
}

/* fuse.rs:315-329 */
void ora_winner_shares(const uint16_t *win, const float *best, size_t cells, size_t frames, float *out) {
    float mx = 0.0f;
    for (size_t i = 0; i < cells; i++) mx = rmaxf(mx, best[i]);
    float floor_ = DETAIL_FLOOR * mx;
    size_t *counts = (size_t *)xcalloc(frames, sizeof(size_t));
    size_t total = 0;
    for (size_t i = 0; i < cells; i++) {
        if (best[i] > floor_ && floor_ > 0.0f) {
            size_t k = f2usize(roundf((float)win[i]));
            if (k < frames) {
                counts[k]++;
                total++;
            }
        }
    }
    for (size_t k = 0; k < frames; k++) out[k] = total > 0 ? (float)counts[k] / (float)total : 0.0f;
    free(counts);
}

/* fuse.rs:336-346 */
static void upsample_index(const uint16_t *win, size_t dw, size_t dh, size_t w, size_t h, size_t level, float *out) {
    size_t scale = (size_t)1 << level;
#pragma omp parallel for schedule(static) if (w * h >= PAR_MIN)
    for (size_t y = 0; y < h; y++) {
        size_t sy = y / scale < dh - 1 ? y / scale : dh - 1;
        for (size_t x = 0; x < w; x++) {
            size_t sx = x / scale < dw - 1 ? x / scale : dw - 1;
            out[y * w + x] = (float)win[sy * dw + sx];
        }
    }
}

/* fuse.rs:349-521 */
struct ora_fuser {
    size_t w, h, levels, depth_level, count;
    ora_fuse_params p;
    int halo_on;
    size_t guide;
    float hard;
    size_t lw[MAXL], lh[MAXL];
    float *acc[MAXL][3];
    float *best[MAXL];
    uint16_t *winner;
    float **tops;
    size_t ntops, captops;
};

ora_fuser *ora_fuser_new(size_t w, size_t h, const ora_fuse_params *p) {
    ora_fuser *f = (ora_fuser *)xcalloc(1, sizeof(ora_fuser));
    f->w = w;
    f->h = h;
    f->p = *p;
    size_t lv = p->levels > 0 ? (size_t)p->levels : ora_auto_levels(w, h, 32);
    f->levels = lv > 1 ? lv : 1;
    f->depth_level = p->depth_level < f->levels - 1 ? p->depth_level : f->levels - 1;
    /* fuse.rs:120-122 halo_guide */
    f->halo_on = p->halo > 0.0f;
    f->guide = f->depth_level;
    f->hard = p->halo < HALO_MAX ? p->halo : HALO_MAX;
    size_t cw = w, ch = h;
    for (size_t l = 0; l <= f->levels; l++) {
        f->lw[l] = cw;
        f->lh[l] = ch;
        cw = half(cw);
        ch = half(ch);
    }
    return f;
}

size_t ora_fuser_levels(const ora_fuser *f) { return f->levels; }
void ora_fuser_level_dims(const ora_fuser *f, size_t l, size_t *w, size_t *h) {
    *w = f->lw[l];
    *h = f->lh[l];
}

static size_t select_levels(const ora_fuser *f) { return f->halo_on ? f->guide + 1 : f->levels; }

/* fuse.rs:401-484 */
void ora_fuser_push(ora_fuser *f, const float *img3) {
    pyr_t P;
    pyr_build(img3, f->w, f->h, f->levels, &P);
    size_t nsel = select_levels(f);
    uint16_t idx = (uint16_t)f->count;
    int first = f->count == 0;
    float *en[MAXL];
    for (size_t l = 0; l < nsel; l++) {
        en[l] = fzeros(P.w[l] * P.h[l]);
        region_energy_planes(P.p[l], P.w[l], P.h[l], f->p.energy_radius, f->p.use_chroma, en[l]);
    }
    if (first) {
        f->winner = (uint16_t *)xcalloc(P.w[f->depth_level] * P.h[f->depth_level], sizeof(uint16_t));
        for (size_t l = 0; l < nsel; l++) {
            f->best[l] = fzeros(P.w[l] * P.h[l]);
            memcpy(f->best[l], en[l], P.w[l] * P.h[l] * sizeof(float));
        }
    } else {
        for (size_t li = 0; li < nsel; li++) {
            size_t n = P.w[li] * P.h[li];
            float *best = f->best[li], *a0 = f->acc[li][0], *a1 = f->acc[li][1], *a2 = f->acc[li][2];
            const float *n0 = P.p[li][0], *n1 = P.p[li][1], *n2 = P.p[li][2], *e = en[li];
            int is_depth = li == f->depth_level;
            uint16_t *win = f->winner;
#pragma omp parallel for schedule(static) if (n >= PAR_MIN)
            for (size_t k = 0; k < n; k++) {
                if (e[k] > best[k]) {
                    best[k] = e[k];
                    a0[k] = n0[k];
                    a1[k] = n1[k];
                    a2[k] = n2[k];
                    if (is_depth) win[k] = idx;
                }
            }
        }
    }
    if (f->halo_on) {
        size_t g = f->guide;
        size_t ww = P.w[g], wh = P.h[g];
        float *wgt = fzeros(ww * wh);
        for (size_t k = 0; k < ww * wh; k++) wgt[k] = ora_halo_weight(en[g][k], f->hard);
        for (size_t li = g + 1; li <= f->levels; li++) {
            float *nw = fzeros(half(ww) * half(wh));
            ora_reduce(wgt, ww, wh, nw);
            free(wgt);
            wgt = nw;
            ww = half(ww);
            wh = half(wh);
            size_t n = ww * wh;
            if (first) {
                for (int c = 0; c < 3; c++)
                    for (size_t k = 0; k < n; k++) P.p[li][c][k] *= wgt[k];
                f->best[li] = fzeros(n);
                memcpy(f->best[li], wgt, n * sizeof(float));
            } else {
                for (size_t k = 0; k < n; k++) f->best[li][k] += wgt[k];
                for (int c = 0; c < 3; c++)
                    for (size_t k = 0; k < n; k++) f->acc[li][c][k] += wgt[k] * P.p[li][c][k];
            }
        }
        free(wgt);
        if (first) {
            for (size_t l = 0; l <= f->levels; l++)
                for (int c = 0; c < 3; c++) {
                    f->acc[l][c] = P.p[l][c];
                    P.p[l][c] = NULL;
                }
        }
    } else {
        size_t L = f->levels;
        size_t tsz = P.w[L] * P.h[L];
        if (f->ntops == f->captops) {
            f->captops = f->captops ? 2 * f->captops : 16;
            f->tops = (float **)realloc(f->tops, f->captops * sizeof(float *));
        }
        float *top = fzeros(3 * tsz);
        for (int c = 0; c < 3; c++) memcpy(top + c * tsz, P.p[L][c], tsz * sizeof(float));
        f->tops[f->ntops++] = top;
        if (first) {
            for (size_t l = 0; l < L; l++)
                for (int c = 0; c < 3; c++) {
                    f->acc[l][c] = P.p[l][c];
                    P.p[l][c] = NULL;
                }
        }
    }
    for (size_t l = 0; l < nsel; l++) free(en[l]);
    pyr_free(&P);
    f->count++;
}

/* fuse.rs:489-494 */
size_t ora_fuser_shares(const ora_fuser *f, float *out) {
    if (f->count == 0 || f->depth_level >= select_levels(f)) return 0;
    size_t d = f->depth_level;
    ora_winner_shares(f->winner, f->best[d], f->lw[d] * f->lh[d], f->count, out);
    return f->count;
}

/* fuse.rs:499-520 */
void ora_fuser_finish(ora_fuser *f, float *img3, float *winner_full, uint16_t *winner_level) {
    size_t L = f->levels, d = f->depth_level;
    pyr_t P;
    P.n = L + 1;
    for (size_t l = 0; l <= L; l++) {
        P.w[l] = f->lw[l];
        P.h[l] = f->lh[l];
    }
    if (f->halo_on) {
        for (size_t li = f->guide + 1; li <= L; li++)
            for (int c = 0; c < 3; c++)
                for (size_t k = 0; k < f->lw[li] * f->lh[li]; k++) f->acc[li][c][k] /= f->best[li][k];
        for (size_t l = 0; l <= L; l++)
            for (int c = 0; c < 3; c++) {
                P.p[l][c] = f->acc[l][c];
                f->acc[l][c] = NULL;
            }
    } else {
        size_t tsz = f->lw[L] * f->lh[L];
        float *tops = fzeros(f->ntops * 3 * tsz);
        for (size_t t = 0; t < f->ntops; t++) memcpy(tops + t * 3 * tsz, f->tops[t], 3 * tsz * sizeof(float));
        float *res = fzeros(3 * tsz);
        ora_fuse_residuals(tops, f->ntops, f->lw[L], f->lh[L], &f->p, res);
        free(tops);
        for (size_t l = 0; l < L; l++)
            for (int c = 0; c < 3; c++) {
                P.p[l][c] = f->acc[l][c];
                f->acc[l][c] = NULL;
            }
        for (int c = 0; c < 3; c++) {
            P.p[L][c] = fzeros(tsz);
            memcpy(P.p[L][c], res + c * tsz, tsz * sizeof(float));
        }
        free(res);
    }
    pyr_collapse(&P, img3);
    for (size_t k = 0; k < 3 * f->w * f->h; k++) img3[k] = rclampf(img3[k], 0.0f, 1.0f);
    if (winner_full) upsample_index(f->winner, f->lw[d], f->lh[d], f->w, f->h, d, winner_full);
    if (winner_level) memcpy(winner_level, f->winner, f->lw[d] * f->lh[d] * sizeof(uint16_t));
}

void ora_fuser_free(ora_fuser *f) {
    if (!f) return;
    for (size_t l = 0; l < MAXL; l++) {
        for (int c = 0; c < 3; c++) free(f->acc[l][c]);
        free(f->best[l]);
    }
    for (size_t t = 0; t < f->ntops; t++) free(f->tops[t]);
    free(f->tops);
    free(f->winner);
    free(f);
}

/* ======================================================================
 * align.rs
 * ==================================================================== */

static const double SPAN[8] = {0.10, 0.10, 0.10, 0.087266462599716474, 0.05, 0.05, 0.05, 0.05};
static const double SIM_ID[8] = {0, 0, 1, 0, 1, 0, 0, 0};

static int sim_is_id(const double *s) {
    for (int k = 0; k < 8; k++)
        if (s[k] != SIM_ID[k]) return 0;
    return 1;
}

/* sin and cos of a rotation, correctly rounded (align.rs:243 takes f64::cos / f64::sin, the platform's
 * libm, whose last bit varies between libraries: glibc 2.35 is one ulp from the correctly rounded value
 * on about 0.3 % of arguments in [-0.3, 0.3]).  For |x| <= pi/4 the Taylor series is summed term by
 * term in double-double arithmetic (Dekker's product through fma, Knuth's two-sum): t_k = -t_{k-1} x^2 /
 * ((2k)(2k+1)) for the sine from t_0 = x, and t_k = -t_{k-1} x^2 / ((2k-1)(2k)) for the cosine from 1,
 * until a term falls below 2^-110 of the sum, then rounded once; beyond pi/4 libm. */
typedef struct { double h, l; } ora_dd;
static inline ora_dd dd_fast(double a, double b) { ora_dd r; r.h = a + b; r.l = b - (r.h - a); return r; }
static inline ora_dd dd_sum(ora_dd a, ora_dd b) {
    double s = a.h + b.h, bb = s - a.h, e = (a.h - (s - bb)) + (b.h - bb);
    return dd_fast(s, e + a.l + b.l);
}
static inline ora_dd dd_mul(ora_dd a, ora_dd b) {
    double p = a.h * b.h, e = fma(a.h, b.h, -p);
    return dd_fast(p, e + (a.h * b.l + a.l * b.h));
}
static inline ora_dd dd_div_d(ora_dd a, double b) {
    double q1 = a.h / b, p = q1 * b, e = fma(q1, b, -p);
    double r = ((a.h - p) - e) + a.l;
    return dd_fast(q1, r / b);
}
static double dd_series(double x, int odd) {
    ora_dd x2 = {x * x, fma(x, x, -(x * x))};
    ora_dd t = {odd ? x : 1.0, 0.0}, s = t;
    for (int k = 1; k < 40; k++) {
        double den = odd ? (double)(2 * k) * (double)(2 * k + 1) : (double)(2 * k - 1) * (double)(2 * k);
        t = dd_div_d(dd_mul(t, x2), -den);
        s = dd_sum(s, t);
        if (fabs(t.h) < 0x1p-110 * fabs(s.h)) break;
    }
    return s.h;
}
void ora_sincos(double x, double *s, double *c) {
    if (!(fabs(x) <= 0.78539816339744828)) {
        *s = sin(x);
        *c = cos(x);
        return;
    }
    *s = dd_series(x, 1);
    *c = dd_series(x, 0);
}

/* align.rs:241-252 */
void ora_sim_matrix(const double *s, size_t w, size_t h, double *m) {
    double cx = (double)w / 2.0, cy = (double)h / 2.0;
    double c, sn;
    ora_sincos(s[3], &sn, &c);
    double sc = s[2];
    double a = sc * c, b = sc * (c * s[5] - sn * s[4]), d = sc * sn, e = sc * (sn * s[5] + c * s[4]);
    double gx = s[6] / (double)w, gy = s[7] / (double)h;
    double a2 = a + cx * gx, b2 = b + cx * gy, d2 = d + cy * gx, e2 = e + cy * gy;
    double tx = cx + s[0] * (double)w - (a2 * cx + b2 * cy);
    double ty = cy + s[1] * (double)h - (d2 * cx + e2 * cy);
    m[0] = a2; m[1] = b2; m[2] = tx;
    m[3] = d2; m[4] = e2; m[5] = ty;
    m[6] = gx; m[7] = gy; m[8] = 1.0 - gx * cx - gy * cy;
}

/* align.rs:258-286 */
void ora_inverse(const double *m, double *inv) {
    if (m[6] == 0.0 && m[7] == 0.0 && m[8] == 1.0) {
        double a = m[0], b = m[1], tx = m[2], d = m[3], e = m[4], ty = m[5];
        double det = a * e - b * d;
        double ia = e / det, ib = -b / det, id = -d / det, ie = a / det;
        inv[0] = ia; inv[1] = ib; inv[2] = -(ia * tx + ib * ty);
        inv[3] = id; inv[4] = ie; inv[5] = -(id * tx + ie * ty);
        inv[6] = 0.0; inv[7] = 0.0; inv[8] = 1.0;
        return;
    }
    double t[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            /* cofactor (j, i) */
            int r0 = (j + 1) % 3, r1 = (j + 2) % 3, c0 = (i + 1) % 3, c1 = (i + 2) % 3;
            t[i * 3 + j] = m[r0 * 3 + c0] * m[r1 * 3 + c1] - m[r0 * 3 + c1] * m[r1 * 3 + c0];
        }
    double k = t[8];
    for (int i = 0; i < 9; i++) inv[i] = t[i] / k;
}

/* align.rs:291-294 */
void ora_map(const double *inv, double x, double y, double *sx, double *sy) {
    double d = inv[6] * x + inv[7] * y + inv[8];
    *sx = (inv[0] * x + inv[1] * y + inv[2]) / d;
    *sy = (inv[3] * x + inv[4] * y + inv[5]) / d;
}

/* align.rs:562-571 and depth.rs:160-165 */
void ora_luma(const float *img3, size_t n, float *y) {
    const float *r = img3, *g = img3 + n, *b = img3 + 2 * n;
#pragma omp parallel for schedule(static) if (n >= PAR_MIN)
    for (size_t i = 0; i < n; i++) y[i] = 0.299f * r[i] + 0.587f * g[i] + 0.114f * b[i];
}

/* align.rs:416-474: interpolation weights in f64 */
static void spline4(double t, double *w) {
    w[0] = ((-1.0 / 3.0 * t + 0.8) * t - 0.46666667) * t;
    w[1] = ((t - 1.8) * t - 0.2) * t + 1.0;
    w[2] = ((1.2 - t) * t + 0.8) * t;
    w[3] = ((1.0 / 3.0 * t - 0.2) * t - 0.13333334) * t;
}
static double keys_k(double d) {
    if (d < 1.0) return (1.5 * d - 2.5) * d * d + 1.0;
    if (d < 2.0) return ((-0.5 * d + 2.5) * d - 4.0) * d + 2.0;
    return 0.0;
}
static double spline36_k(double d) {
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
static double lanczos3_k(double d) {
    if (d < 1e-9) return 1.0;
    if (d < 3.0) {
        double a = 3.14159265358979323846 * d, b = 3.14159265358979323846 * d / 3.0;
        return sin(a) / a * (sin(b) / b);
    }
    return 0.0;
}
static void taps_of(double (*k)(double), int N, double t, double *w) {
    double sum = 0.0;
    for (int i = 0; i < N; i++) {
        w[i] = k(fabs(t - ((double)i + 1.0 - (double)(N / 2))));
        sum += w[i];
    }
    for (int i = 0; i < N; i++) w[i] /= sum;
}
static int interp_taps(uint32_t interp) {
    switch (interp) {
    case 0: case 1: return 2;
    case 2: case 3: return 4;
    default: return 6;
    }
}
/* align.rs:542-552 */
static void interp_weights(uint32_t interp, double t, double *w) {
    switch (interp) {
    case 0:
        if (t < 0.5) { w[0] = 1.0; w[1] = 0.0; } else { w[0] = 0.0; w[1] = 1.0; }
        break;
    case 1: w[0] = 1.0 - t; w[1] = t; break;
    case 2: taps_of(keys_k, 4, t, w); break;
    case 3: spline4(t, w); break;
    case 4: taps_of(spline36_k, 6, t, w); break;
    default: taps_of(lanczos3_k, 6, t, w); break;
    }
}

static inline long lfloor(double v) { return (long)floor(v); }
static inline long lclamp(long v, long lo, long hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* align.rs:489-538 */
void ora_warp_plane(const float *src, size_t w, size_t h, const double *sim, size_t ow, size_t oh, uint32_t interp,
                    float *out, uint8_t *valid) {
    double m[9], inv[9];
    ora_sim_matrix(sim, w, h, m);
    ora_inverse(m, inv);
    int N = interp_taps(interp);
    long start = 1 - N / 2;
    double xmax = (double)(w - 1), ymax = (double)(h - 1);
    int affine = inv[6] == 0.0 && inv[7] == 0.0 && inv[8] == 1.0;
#pragma omp parallel for schedule(static) if (ow * oh >= PAR_MIN)
    for (size_t y = 0; y < oh; y++) {
        double yf = (double)y;
        double ax0 = inv[1] * yf + inv[2], ay0 = inv[4] * yf + inv[5], adx = inv[0], ady = inv[3];
        for (size_t x = 0; x < ow; x++) {
            double sx, sy;
            if (affine) {
                sx = ax0 + adx * (double)x;
                sy = ay0 + ady * (double)x;
            } else {
                ora_map(inv, (double)x, yf, &sx, &sy);
            }
            if (valid) valid[y * ow + x] = (sx >= 0.0 && sx <= xmax && sy >= 0.0 && sy <= ymax);
            long x0 = lfloor(sx), y0 = lfloor(sy);
            double wxd[6], wyd[6];
            interp_weights(interp, sx - (double)x0, wxd);
            interp_weights(interp, sy - (double)y0, wyd);
            float wx[6], wy[6];
            for (int i = 0; i < N; i++) {
                wx[i] = (float)wxd[i];
                wy[i] = (float)wyd[i];
            }
            float acc = 0.0f;
            long xs = x0 + start, ys = y0 + start;
            for (int j = 0; j < N; j++) {
                long yy = lclamp(ys + j, 0, (long)h - 1);
                float r = 0.0f;
                for (int i = 0; i < N; i++) {
                    long xx = lclamp(xs + i, 0, (long)w - 1);
                    r += wx[i] * src[(size_t)yy * w + (size_t)xx];
                }
                acc += wy[j] * r;
            }
            out[y * ow + x] = acc;
        }
    }
}

/* align.rs:639-646 */
static void spline4f(float t, float *w) {
    w[0] = ((-1.0f / 3.0f * t + 0.8f) * t - 0.46666667f) * t;
    w[1] = ((t - 1.8f) * t - 0.2f) * t + 1.0f;
    w[2] = ((1.2f - t) * t + 0.8f) * t;
    w[3] = ((1.0f / 3.0f * t - 0.2f) * t - 0.13333334f) * t;
}

/* align.rs:581-635 */
double ora_warp_cost(const float *rf, size_t aw, const float *tgt, size_t tw, size_t th, const double *sim) {
    double m[9], inv[9];
    ora_sim_matrix(sim, tw, th, m);
    ora_inverse(m, inv);
    double xmax = (double)(tw - 1), ymax = (double)(th - 1);
    size_t ah = th; /* the reference level has the target's dimensions */
    int affine = inv[6] == 0.0 && inv[7] == 0.0 && inv[8] == 1.0;
    double *psd = (double *)xcalloc(ah, sizeof(double)), *psd2 = (double *)xcalloc(ah, sizeof(double));
    size_t *pcnt = (size_t *)xcalloc(ah, sizeof(size_t));
#pragma omp parallel for schedule(static) if (aw * ah >= PAR_MIN)
    for (size_t y = 0; y < ah; y++) {
        double sd = 0.0, sd2 = 0.0;
        size_t cnt = 0;
        double yf = (double)y;
        double ax0 = inv[1] * yf + inv[2], ay0 = inv[4] * yf + inv[5];
        for (size_t x = 0; x < aw; x++) {
            float r = rf[y * aw + x];
            double sx, sy;
            if (affine) {
                sx = ax0 + inv[0] * (double)x;
                sy = ay0 + inv[3] * (double)x;
            } else {
                ora_map(inv, (double)x, yf, &sx, &sy);
            }
            if (!(sx >= 0.0 && sx <= xmax && sy >= 0.0 && sy <= ymax)) continue;
            long x0 = lfloor(sx), y0 = lfloor(sy);
            float wx[4], wy[4];
            spline4f((float)(sx - (double)x0), wx);
            spline4f((float)(sy - (double)y0), wy);
            float acc = 0.0f;
            for (int j = 0; j < 4; j++) {
                long yy = lclamp(y0 + j - 1, 0, (long)th - 1);
                float rr = 0.0f;
                for (int i = 0; i < 4; i++) {
                    long xx = lclamp(x0 + i - 1, 0, (long)tw - 1);
                    rr += wx[i] * tgt[(size_t)yy * tw + (size_t)xx];
                }
                acc += wy[j] * rr;
            }
            double d = (double)(r - acc);
            sd += d;
            sd2 += d * d;
            cnt++;
        }
        psd[y] = sd;
        psd2[y] = sd2;
        pcnt[y] = cnt;
    }
    double sd = 0.0, sd2 = 0.0;
    size_t cnt = 0;
    for (size_t y = 0; y < ah; y++) {
        sd += psd[y];
        sd2 += psd2[y];
        cnt += pcnt[y];
    }
    free(psd);
    free(psd2);
    free(pcnt);
    if (cnt < 16) return 1e9;
    double n = (double)cnt;
    double v = (sd2 - sd * sd / n) / n;
    return sqrt(v > 0.0 ? v : 0.0);
}

/* align.rs:650-655 */
static void burt_kernel(float *k) {
    float a = 0.33f, b = 0.25f, c = 0.25f - a / 2.0f;
    k[0] = c; k[1] = b; k[2] = a; k[3] = b; k[4] = c;
}
/* align.rs:658-671 */
static void norm_full(size_t n, const float *k, float *out) {
    for (size_t i = 0; i < n; i++) {
        float s = 0.0f;
        for (int t = 0; t < 5; t++) {
            long x = (long)i + t - 2;
            if (x >= 0 && (size_t)x < n) s += k[t];
        }
        out[i] = s;
    }
}

/* align.rs:674-734 */
void ora_reduce_burt(const float *src, size_t w, size_t h, float *out) {
    float k[5];
    burt_kernel(k);
    size_t ow = (w + 1) / 2, oh = (h + 1) / 2;
    float *cn = fzeros(w), *rn = fzeros(h);
    norm_full(w, k, cn);
    norm_full(h, k, rn);
    float *htmp = fzeros(h * ow);
#pragma omp parallel for schedule(static) if (h * ow >= PAR_MIN)
    for (size_t y = 0; y < h; y++) {
        const float *s = src + y * w;
        for (size_t oj = 0; oj < ow; oj++) {
            long cx = 2 * (long)oj;
            float a = 0.0f;
            for (int t = 0; t < 5; t++) {
                long i = cx + t - 2;
                if (i >= 0 && (size_t)i < w) a += k[t] * s[i];
            }
            htmp[y * ow + oj] = a;
        }
    }
#pragma omp parallel for schedule(static) if (oh * ow >= PAR_MIN)
    for (size_t oi = 0; oi < oh; oi++) {
        long cy = 2 * (long)oi;
        float r = rn[cy];
        float *row = out + oi * ow;
        for (size_t oj = 0; oj < ow; oj++) row[oj] = 0.0f;
        for (int t = 0; t < 5; t++) {
            long s = cy + t - 2;
            if (s >= 0 && (size_t)s < h) {
                float kt = k[t];
                const float *hr = htmp + (size_t)s * ow;
                for (size_t oj = 0; oj < ow; oj++) row[oj] += kt * hr[oj];
            }
        }
        for (size_t oj = 0; oj < ow; oj++) row[oj] /= r * cn[2 * oj];
    }
    free(cn);
    free(rn);
    free(htmp);
}

/* align.rs:741-754: levels while h > 64 && w > 8 */
size_t ora_gauss_count(size_t w, size_t h, size_t *dims) {
    size_t n = 0;
    while (h > 64 && w > 8) {
        w = (w + 1) / 2;
        h = (h + 1) / 2;
        if (dims) {
            dims[2 * n] = w;
            dims[2 * n + 1] = h;
        }
        n++;
    }
    return n;
}

void ora_gauss_levels(const float *y, size_t w, size_t h, float *packed) {
    const float *cur = y;
    size_t cw = w, ch = h;
    float *out = packed;
    while (ch > 64 && cw > 8) {
        ora_reduce_burt(cur, cw, ch, out);
        cw = (cw + 1) / 2;
        ch = (ch + 1) / 2;
        cur = out;
        out += cw * ch;
    }
}

/* align.rs:766-790 */
static void param_gain(size_t w, size_t h, double *g) {
    double W = (double)w, H = (double)h;
    double r = (W > H ? W : H) / 2.0;
    g[0] = W; g[1] = H; g[2] = r; g[3] = r; g[4] = H / 2.0; g[5] = H / 2.0; g[6] = W / 4.0; g[7] = H / 4.0;
}
#define STEP_PX 1.0
#define TOL_PX 0.1
void ora_level_steps(const size_t *free_idx, size_t nfree, size_t lvl, size_t w, size_t h, double *step, double *tol) {
    double g[8];
    param_gain(w, h, g);
    double px = (double)((uint64_t)1 << lvl);
    for (size_t k = 0; k < nfree; k++) {
        step[k] = STEP_PX * px / g[free_idx[k]];
        tol[k] = TOL_PX * px / g[free_idx[k]];
    }
}

/* align.rs:888-894 */
static int nm_converged(double (*s)[8], const double *fv, const double *tol, size_t n) {
    if (fabs(fv[n] - fv[0]) > 1e-4 * (1.0 + fabs(fv[0]))) return 0;
    for (size_t i = 1; i <= n; i++)
        for (size_t k = 0; k < n; k++)
            if (!(fabs(s[i][k] - s[0][k]) <= tol[k])) return 0;
    return 1;
}

/* align.rs:796-883 */
void ora_nelder_mead(ora_cost_fn f, void *ctx, size_t n, const double *x0, const double *lo, const double *hi,
                     const double *step, const double *tol, double *out, uint32_t *evals) {
    uint32_t ne = 0;
    if (n == 0) {
        if (evals) *evals = 0;
        return;
    }
    double s[9][8] = {{0}}, fv[9] = {0}, tmp[9][8], tfv[9];
#define CLAMPV(v)                                                   \
    for (size_t k_ = 0; k_ < n; k_++) (v)[k_] = rclamp((v)[k_], lo[k_], hi[k_]);
    for (size_t k = 0; k < n; k++) s[0][k] = x0[k];
    for (size_t i = 0; i < n; i++) {
        for (size_t k = 0; k < n; k++) s[i + 1][k] = x0[k];
        s[i + 1][i] += step[i];
        CLAMPV(s[i + 1]);
    }
    for (size_t i = 0; i <= n; i++) {
        fv[i] = f(s[i], ctx);
        ne++;
    }
    const double alpha = 1.0, gamma = 2.0, rho = 0.5, sigma = 0.5;
    for (int it = 0; it < 200; it++) {
        /* stable sort of the vertices by value (slice::sort_by) */
        size_t idx[9];
        for (size_t i = 0; i <= n; i++) idx[i] = i;
        for (size_t i = 1; i <= n; i++) {
            size_t v = idx[i];
            size_t j = i;
            while (j > 0 && fv[idx[j - 1]] > fv[v]) {
                idx[j] = idx[j - 1];
                j--;
            }
            idx[j] = v;
        }
        for (size_t i = 0; i <= n; i++) {
            memcpy(tmp[i], s[idx[i]], sizeof(tmp[i]));
            tfv[i] = fv[idx[i]];
        }
        memcpy(s, tmp, sizeof(s));
        memcpy(fv, tfv, sizeof(fv));
        if (nm_converged(s, fv, tol, n)) break;
        double c[8] = {0};
        for (size_t i = 0; i < n; i++)
            for (size_t k = 0; k < n; k++) c[k] += s[i][k];
        for (size_t k = 0; k < n; k++) c[k] /= (double)n;
        double xr[8];
        for (size_t k = 0; k < n; k++) xr[k] = c[k] + alpha * (c[k] - s[n][k]);
        CLAMPV(xr);
        double fr = f(xr, ctx);
        ne++;
        if (fr < fv[0]) {
            double xe[8];
            for (size_t k = 0; k < n; k++) xe[k] = c[k] + gamma * (xr[k] - c[k]);
            CLAMPV(xe);
            double fe = f(xe, ctx);
            ne++;
            if (fe < fr) {
                memcpy(s[n], xe, sizeof(xe));
                fv[n] = fe;
            } else {
                memcpy(s[n], xr, sizeof(xr));
                fv[n] = fr;
            }
        } else if (fr < fv[n - 1]) {
            memcpy(s[n], xr, sizeof(xr));
            fv[n] = fr;
        } else {
            double xc[8];
            for (size_t k = 0; k < n; k++) xc[k] = c[k] + rho * (s[n][k] - c[k]);
            CLAMPV(xc);
            double fc = f(xc, ctx);
            ne++;
            if (fc < fv[n]) {
                memcpy(s[n], xc, sizeof(xc));
                fv[n] = fc;
            } else {
                for (size_t i = 1; i <= n; i++) {
                    for (size_t k = 0; k < n; k++) s[i][k] = s[0][k] + sigma * (s[i][k] - s[0][k]);
                    CLAMPV(s[i]);
                    fv[i] = f(s[i], ctx);
                    ne++;
                }
            }
        }
    }
#undef CLAMPV
    size_t best = 0;
    for (size_t i = 1; i <= n; i++)
        if (fv[i] < fv[best]) best = i;
    for (size_t k = 0; k < n; k++) out[k] = s[best][k];
    if (evals) *evals = ne;
}

typedef struct {
    const float *a;
    size_t aw;
    const float *t;
    size_t tw, th;
    double cur[8];
    size_t free_idx[8], nfree;
    double *trace; /* every evaluated cost in order (nullable) */
    size_t trace_cap, *trace_n;
} align_ctx;

static double align_cost(const double *xf, void *vctx) {
    align_ctx *c = (align_ctx *)vctx;
    double v[8];
    memcpy(v, c->cur, sizeof(v));
    for (size_t k = 0; k < c->nfree; k++) v[c->free_idx[k]] = xf[k];
    double cost = ora_warp_cost(c->a, c->aw, c->t, c->tw, c->th, v);
    if (c->trace_n) {
        if (*c->trace_n < c->trace_cap) c->trace[*c->trace_n] = cost;
        (*c->trace_n)++;
    }
    return cost;
}

void ora_multiscale_align(const float *rf, const float *tg, size_t w, size_t h, const double *init,
                          const uint8_t *free8, size_t coarsen, double *out, uint32_t *evals) {
    ora_multiscale_align_trace(rf, tg, w, h, init, free8, coarsen, out, evals, NULL, NULL, 0, NULL);
}

/* align.rs:896-942; lev_evals: the evaluations per level (level 0 = full resolution, MAXL entries);
 * trace: the costs in the order nelder_mead evaluates them (up to trace_cap; *trace_n the count) */
void ora_multiscale_align_trace(const float *rf, const float *tg, size_t w, size_t h, const double *init,
                                const uint8_t *free8, size_t coarsen, double *out, uint32_t *evals,
                                uint32_t *lev_evals, double *trace, size_t trace_cap, size_t *trace_n) {
    if (trace_n) *trace_n = 0;
    if (lev_evals)
        for (int l = 0; l < ORA_MAXL; l++) lev_evals[l] = 0;
    size_t dims[2 * MAXL];
    size_t nl = ora_gauss_count(w, h, dims);
    size_t tot = 0;
    for (size_t l = 0; l < nl; l++) tot += dims[2 * l] * dims[2 * l + 1];
    float *pref = fzeros(tot), *ptgt = fzeros(tot);
    ora_gauss_levels(rf, w, h, pref);
    ora_gauss_levels(tg, w, h, ptgt);
    size_t n = nl + 1;
    double cur[8];
    memcpy(cur, init, sizeof(cur));
    size_t free_idx[8], nfree = 0;
    for (size_t k = 0; k < 8; k++)
        if (free8[k]) free_idx[nfree++] = k;
    double lo[8], hi[8];
    for (size_t k = 0; k < nfree; k++) {
        lo[k] = init[free_idx[k]] - SPAN[free_idx[k]];
        hi[k] = init[free_idx[k]] + SPAN[free_idx[k]];
    }
    size_t finest = coarsen < n - 1 ? coarsen : n - 1;
    uint32_t tot_evals = 0;
    for (long lvl = (long)n - 1; lvl >= (long)finest; lvl--) {
        align_ctx c;
        if (lvl == 0) {
            c.a = rf;
            c.t = tg;
            c.aw = w;
            c.tw = w;
            c.th = h;
        } else {
            size_t off = 0;
            for (long l = 0; l < lvl - 1; l++) off += dims[2 * l] * dims[2 * l + 1];
            c.a = pref + off;
            c.t = ptgt + off;
            c.aw = dims[2 * (lvl - 1)];
            c.tw = c.aw;
            c.th = dims[2 * (lvl - 1) + 1];
        }
        memcpy(c.cur, cur, sizeof(cur));
        memcpy(c.free_idx, free_idx, sizeof(free_idx));
        c.nfree = nfree;
        c.trace = trace;
        c.trace_cap = trace_cap;
        c.trace_n = trace_n;
        double x0[8], step[8], tol[8], best[8];
        for (size_t k = 0; k < nfree; k++) x0[k] = cur[free_idx[k]];
        ora_level_steps(free_idx, nfree, (size_t)lvl, w, h, step, tol);
        uint32_t ne = 0;
        ora_nelder_mead(align_cost, &c, nfree, x0, lo, hi, step, tol, best, &ne);
        tot_evals += ne;
        if (lev_evals && lvl < ORA_MAXL) lev_evals[lvl] = ne;
        for (size_t k = 0; k < nfree; k++) cur[free_idx[k]] = best[k];
    }
    memcpy(out, cur, sizeof(cur));
    if (evals) *evals = tot_evals;
    free(pref);
    free(ptgt);
}

/* align.rs:344-403 */
void ora_common_area(const double *sims, size_t n, size_t w, size_t h, uint32_t interp, size_t *rect) {
    rect[0] = 0; rect[1] = 0; rect[2] = w; rect[3] = h;
    if (w == 0 || h == 0) return;
    double *invs = (double *)xcalloc(9 * (n ? n : 1), sizeof(double));
    size_t ni = 0;
    for (size_t i = 0; i < n; i++) {
        if (sim_is_id(sims + 8 * i)) continue;
        double m[9];
        ora_sim_matrix(sims + 8 * i, w, h, m);
        ora_inverse(m, invs + 9 * ni);
        ni++;
    }
    if (ni == 0) {
        free(invs);
        return;
    }
    double mg = (double)(interp_taps(interp) / 2);
    if (interp == 0) mg = 0.0; /* nearest: taps() = 1 */
    double xmax = (double)(w - 1) - mg, ymax = (double)(h - 1) - mg;
    long long *rlo = (long long *)xcalloc(h, sizeof(long long)), *rhi = (long long *)xcalloc(h, sizeof(long long));
    for (size_t y = 0; y < h; y++) {
        double lo = 0.0, hi = (double)(w - 1);
        double yf = (double)y;
        for (size_t i = 0; i < ni; i++) {
            const double *inv = invs + 9 * i;
            double g = inv[6], d = inv[7] * yf + inv[8];
            if (d <= 0.0 || g * (double)(w - 1) + d <= 0.0) {
                lo = 1.0;
                hi = 0.0;
                break;
            }
            double A[2] = {inv[0], inv[3]}, K[2] = {inv[1] * yf + inv[2], inv[4] * yf + inv[5]};
            double SMIN[2] = {mg, mg}, SMAX[2] = {xmax, ymax};
            for (int ax = 0; ax < 2; ax++) {
                double a = A[ax], k = K[ax], smin = SMIN[ax], smax = SMAX[ax];
                double P[2] = {a - smin * g, smax * g - a}, Q[2] = {k - smin * d, smax * d - k};
                for (int t = 0; t < 2; t++) {
                    double p = P[t], q = Q[t];
                    if (fabs(p) < 1e-12) {
                        if (q < 0.0) {
                            lo = 1.0;
                            hi = 0.0;
                        }
                    } else if (p > 0.0) {
                        double v = -q / p;
                        lo = lo > v ? lo : v;
                    } else {
                        double v = -q / p;
                        hi = hi < v ? hi : v;
                    }
                }
            }
        }
        if (lo <= hi) {
            rlo[y] = (long long)ceil(lo);
            rhi[y] = (long long)floor(hi);
        } else {
            rlo[y] = 1;
            rhi[y] = 0;
        }
    }
    size_t bx = 0, by = 0, bw = 0, bh = 0;
    for (size_t y0 = 0; y0 < h; y0++) {
        long long lo = rlo[y0], hi = rhi[y0];
        for (size_t y1 = y0; y1 < h; y1++) {
            lo = lo > rlo[y1] ? lo : rlo[y1];
            hi = hi < rhi[y1] ? hi : rhi[y1];
            if (lo > hi) break;
            size_t rw = (size_t)(hi - lo + 1), rh = y1 - y0 + 1;
            if (rw * rh > bw * bh) {
                bx = (size_t)lo;
                by = y0;
                bw = rw;
                bh = rh;
            }
        }
    }
    if (bw * bh > 0) {
        rect[0] = bx; rect[1] = by; rect[2] = bw; rect[3] = bh;
    }
    free(rlo);
    free(rhi);
    free(invs);
}

/* ======================================================================
 * brightness.rs
 * ==================================================================== */

#define BSTEP 4
#define GAIN_MIN 0.25f
#define GAIN_MAX 4.0f

/* brightness.rs:42-53 */
void ora_bright_grid(const float *img3, size_t w, size_t h, float *grid3) {
    size_t gw = (w + BSTEP - 1) / BSTEP, gh = (h + BSTEP - 1) / BSTEP, n = w * h;
    for (int c = 0; c < 3; c++)
        for (size_t y = 0, gy = 0; y < h; y += BSTEP, gy++)
            for (size_t x = 0, gx = 0; x < w; x += BSTEP, gx++) grid3[c * gw * gh + gy * gw + gx] = img3[c * n + y * w + x];
}

/* brightness.rs:67-90 */
static void bright_sums(const float *frame3, const float *grid3, size_t w, size_t h, const double *sim, double *sf,
                        double *so, size_t *cnt) {
    size_t gw = (w + BSTEP - 1) / BSTEP, gh = (h + BSTEP - 1) / BSTEP, n = w * h;
    int use_inv = !sim_is_id(sim);
    double inv[9];
    if (use_inv) {
        double m[9];
        ora_sim_matrix(sim, w, h, m);
        ora_inverse(m, inv);
    }
    for (int c = 0; c < 3; c++) sf[c] = so[c] = 0.0;
    size_t k = 0;
    for (size_t y = 0; y < h; y += BSTEP)
        for (size_t x = 0; x < w; x += BSTEP) {
            if (use_inv) {
                double sx, sy;
                ora_map(inv, (double)x, (double)y, &sx, &sy);
                if (!(sx >= 0.0 && sx <= (double)(w - 1) && sy >= 0.0 && sy <= (double)(h - 1))) continue;
            }
            size_t i = y * w + x;
            for (int c = 0; c < 3; c++) {
                sf[c] += (double)frame3[c * n + i];
                if (grid3) so[c] += (double)grid3[c * gw * gh + (y / BSTEP) * gw + x / BSTEP];
            }
            k++;
        }
    *cnt = k;
}

/* brightness.rs:93-98 */
void ora_bright_ratio(const double *ref, const double *frm, size_t n, float *g3) {
    for (int c = 0; c < 3; c++) {
        if (n < 64) {
            g3[c] = 1.0f;
            continue;
        }
        float g = frm[c] > 1e-9 ? (float)(ref[c] / frm[c]) : 1.0f;
        g3[c] = rclampf(g, GAIN_MIN, GAIN_MAX);
    }
}

/* brightness.rs:57-61 */
void ora_bright_gains(const float *grid3, const float *frame3, size_t w, size_t h, const double *sim, float *g3) {
    double sf[3], sr[3];
    size_t n;
    bright_sums(frame3, grid3, w, h, sim, sf, sr, &n);
    ora_bright_ratio(sr, sf, n, g3);
}

/* brightness.rs:107-110 */
void ora_bright_means(const float *img3, size_t w, size_t h, double *m3) {
    double sf[3], so[3];
    size_t n;
    bright_sums(img3, NULL, w, h, SIM_ID, sf, so, &n);
    for (int c = 0; c < 3; c++) m3[c] = n == 0 ? 0.0 : sf[c] / (double)n;
}

/* brightness.rs:114-117 */
void ora_bright_gains_to(const double *ref_means, const float *img3, size_t w, size_t h, float *g3) {
    double m[3];
    ora_bright_means(img3, w, h, m);
    ora_bright_ratio(ref_means, m, (size_t)-1, g3);
}

/* brightness.rs:119-133 */
void ora_bright_apply(float *img3, size_t n, const float *g) {
    int unity = 1;
    for (int c = 0; c < 3; c++)
        if (!(fabsf(g[c] - 1.0f) < 1e-6f)) unity = 0;
    if (unity) return;
    for (int c = 0; c < 3; c++)
        if (fabsf(g[c] - 1.0f) > 1e-6f)
            for (size_t i = 0; i < n; i++) img3[c * n + i] *= g[c];
}

/* ======================================================================
 * depth.rs
 * ==================================================================== */

static size_t blocks(size_t n, size_t k) { return (n + k - 1) / k; }

/* depth.rs:174-197 */
void ora_block_mean(const float *src, size_t w, size_t h, size_t k, float *out) {
    size_t dw = blocks(w, k), dh = blocks(h, k);
    if (k == 1) {
        memcpy(out, src, w * h * sizeof(float));
        return;
    }
#pragma omp parallel for schedule(static) if (dw * dh >= PAR_MIN)
    for (size_t oy = 0; oy < dh; oy++) {
        float *row = out + oy * dw;
        size_t y0 = oy * k, y1 = y0 + k < h ? y0 + k : h;
        for (size_t ox = 0; ox < dw; ox++) row[ox] = 0.0f;
        for (size_t y = y0; y < y1; y++) {
            const float *s = src + y * w;
            for (size_t ox = 0; ox < dw; ox++) {
                size_t x0 = ox * k, x1 = x0 + k < w ? x0 + k : w;
                float sum = 0.0f;
                for (size_t x = x0; x < x1; x++) sum += s[x];
                row[ox] += sum;
            }
        }
        for (size_t ox = 0; ox < dw; ox++) {
            size_t xe = ox * k + k < w ? ox * k + k : w;
            float cnt = (float)((y1 - y0) * (xe - ox * k));
            row[ox] /= cnt;
        }
    }
}

/* depth.rs:201-221 */
void ora_upsample_bilinear(const float *g, size_t dw, size_t dh, size_t w, size_t h, size_t k, float *out) {
    float inv = 1.0f / (float)k;
#pragma omp parallel for schedule(static) if (w * h >= PAR_MIN)
    for (size_t y = 0; y < h; y++) {
        float fy = rclampf(((float)y + 0.5f) * inv - 0.5f, 0.0f, (float)(dh - 1));
        size_t y0 = f2usize(fy), y1 = y0 + 1 < dh - 1 ? y0 + 1 : dh - 1;
        float ty = fy - (float)y0;
        const float *r0 = g + y0 * dw, *r1 = g + y1 * dw;
        for (size_t x = 0; x < w; x++) {
            float fx = rclampf(((float)x + 0.5f) * inv - 0.5f, 0.0f, (float)(dw - 1));
            size_t x0 = f2usize(fx), x1 = x0 + 1 < dw - 1 ? x0 + 1 : dw - 1;
            float tx = fx - (float)x0;
            float top = r0[x0] + (r0[x1] - r0[x0]) * tx;
            float bot = r1[x0] + (r1[x1] - r1[x0]) * tx;
            out[y * w + x] = top + (bot - top) * ty;
        }
    }
}

/* depth.rs:263-347: border-clipped box filter with f64 running sums */
typedef struct {
    size_t w, h, r;
    float *inv_count;
    float *tmp;
} boxf_t;

static void boxf_init(boxf_t *b, size_t w, size_t h, size_t r) {
    b->w = w;
    b->h = h;
    b->r = r;
    b->inv_count = fzeros(w * h);
    b->tmp = fzeros(w * h);
    for (size_t y = 0; y < h; y++) {
        float cy = (float)(((y + r + 1) < h ? (y + r + 1) : h) - (y > r ? y - r : 0));
        for (size_t x = 0; x < w; x++) {
            float cx = (float)(((x + r + 1) < w ? (x + r + 1) : w) - (x > r ? x - r : 0));
            b->inv_count[y * w + x] = 1.0f / (cy * cx);
        }
    }
}
static void boxf_free(boxf_t *b) {
    free(b->inv_count);
    free(b->tmp);
}

static void boxf_sum(boxf_t *b, const float *src, float *out) {
    size_t w = b->w, h = b->h, r = b->r;
    float *tmp = b->tmp;
#pragma omp parallel for schedule(static) if (w * h >= PAR_MIN)
    for (size_t y = 0; y < h; y++) {
        const float *s = src + y * w;
        float *row = tmp + y * w;
        double acc = 0.0;
        size_t e = r + 1 < w ? r + 1 : w;
        for (size_t x = 0; x < e; x++) acc += (double)s[x];
        for (size_t x = 0; x < w; x++) {
            row[x] = (float)acc;
            if (x + r + 1 < w) acc += (double)s[x + r + 1];
            if (x >= r) acc -= (double)s[x - r];
        }
    }
    size_t nb = blocks(w, 64);
#pragma omp parallel for schedule(static) if (w * h >= PAR_MIN)
    for (size_t bi = 0; bi < nb; bi++) {
        size_t x0 = bi * 64, x1 = x0 + 64 < w ? x0 + 64 : w, bw = x1 - x0;
        double acc[64] = {0};
        size_t e = r + 1 < h ? r + 1 : h;
        for (size_t y = 0; y < e; y++)
            for (size_t j = 0; j < bw; j++) acc[j] += (double)tmp[y * w + x0 + j];
        for (size_t y = 0; y < h; y++) {
            for (size_t j = 0; j < bw; j++) out[y * w + x0 + j] = (float)acc[j];
            if (y + r + 1 < h)
                for (size_t j = 0; j < bw; j++) acc[j] += (double)tmp[(y + r + 1) * w + x0 + j];
            if (y >= r)
                for (size_t j = 0; j < bw; j++) acc[j] -= (double)tmp[(y - r) * w + x0 + j];
        }
    }
}

static void boxf_mean(boxf_t *b, const float *src, float *out) {
    boxf_sum(b, src, out);
    size_t n = b->w * b->h;
#pragma omp parallel for schedule(static) if (n >= PAR_MIN)
    for (size_t i = 0; i < n; i++) out[i] *= b->inv_count[i];
}

void ora_box_mean(const float *src, size_t w, size_t h, size_t r, float *out) {
    boxf_t b;
    boxf_init(&b, w, h, r);
    boxf_mean(&b, src, out);
    boxf_free(&b);
}

/* depth.rs:352-412 */
typedef struct {
    size_t w, h;
    const float *guide;
    float *mean_i, *var_i;
    boxf_t bf;
    float eps;
    float *s[4];
} guided_t;

static void guided_init(guided_t *g, const float *guide, size_t w, size_t h, size_t r, float eps) {
    size_t n = w * h;
    g->w = w;
    g->h = h;
    g->guide = guide;
    g->eps = eps;
    boxf_init(&g->bf, w, h, r);
    g->mean_i = fzeros(n);
    boxf_mean(&g->bf, guide, g->mean_i);
    g->var_i = fzeros(n);
    float *ii = fzeros(n);
    for (size_t i = 0; i < n; i++) ii[i] = guide[i] * guide[i];
    boxf_mean(&g->bf, ii, g->var_i);
    for (size_t i = 0; i < n; i++) g->var_i[i] = rmaxf(g->var_i[i] - g->mean_i[i] * g->mean_i[i], 0.0f);
    g->s[0] = ii;
    for (int k = 1; k < 4; k++) g->s[k] = fzeros(n);
}
static void guided_free(guided_t *g) {
    boxf_free(&g->bf);
    free(g->mean_i);
    free(g->var_i);
    for (int k = 0; k < 4; k++) free(g->s[k]);
}
static void guided_coeffs_into(guided_t *g, const float *p) {
    size_t n = g->w * g->h;
    float *s0 = g->s[0], *s1 = g->s[1], *s2 = g->s[2], *s3 = g->s[3];
    boxf_mean(&g->bf, p, s0);
#pragma omp parallel for schedule(static) if (n >= PAR_MIN)
    for (size_t i = 0; i < n; i++) s1[i] = g->guide[i] * p[i];
    boxf_mean(&g->bf, s1, s2);
#pragma omp parallel for schedule(static) if (n >= PAR_MIN)
    for (size_t k = 0; k < n; k++) {
        float cov = s2[k] - g->mean_i[k] * s0[k];
        float a = cov / (g->var_i[k] + g->eps);
        s2[k] = a;
        s3[k] = s0[k] - a * g->mean_i[k];
    }
    boxf_mean(&g->bf, s2, s1);
    boxf_mean(&g->bf, s3, s0);
}
static void guided_filter_into(guided_t *g, const float *p, float *out) {
    guided_coeffs_into(g, p);
    size_t n = g->w * g->h;
    const float *a = g->s[1], *b = g->s[0];
#pragma omp parallel for schedule(static) if (n >= PAR_MIN)
    for (size_t k = 0; k < n; k++) out[k] = a[k] * g->guide[k] + b[k];
}

void ora_guided_filter(const float *guide, const float *p, size_t w, size_t h, size_t r, float eps, float *out) {
    guided_t g;
    guided_init(&g, guide, w, h, r, eps);
    guided_filter_into(&g, p, out);
    guided_free(&g);
}

void ora_guided_coeffs(const float *guide, const float *p, size_t w, size_t h, size_t r, float eps, float *a, float *b) {
    guided_t g;
    guided_init(&g, guide, w, h, r, eps);
    guided_coeffs_into(&g, p);
    memcpy(a, g.s[1], w * h * sizeof(float));
    memcpy(b, g.s[0], w * h * sizeof(float));
    guided_free(&g);
}

/* depth.rs:454-471 */
size_t ora_rdf_taps(size_t r_in, size_t r_out, int32_t *dy, int32_t *dx, float *wt) {
    long ri2 = (long)(r_in * r_in), ro2 = (long)(r_out * r_out), r = (long)r_out;
    size_t nd = 0, nr = 0;
    for (long y = -r; y <= r; y++)
        for (long x = -r; x <= r; x++) {
            long d2 = x * x + y * y;
            if (d2 <= ri2)
                nd++;
            else if (d2 <= ro2)
                nr++;
        }
    float wd = 1.0f / (float)nd, wr = -1.0f / (float)(nr > 1 ? nr : 1);
    size_t k = 0;
    for (int pass = 0; pass < 2; pass++)
        for (long y = -r; y <= r; y++)
            for (long x = -r; x <= r; x++) {
                long d2 = x * x + y * y;
                int in_disk = d2 <= ri2, in_ring = !in_disk && d2 <= ro2;
                if ((pass == 0 && in_disk) || (pass == 1 && in_ring)) {
                    if (dy) {
                        dy[k] = (int32_t)y;
                        dx[k] = (int32_t)x;
                        wt[k] = pass == 0 ? wd : wr;
                    }
                    k++;
                }
            }
    return k;
}

/* depth.rs:417-451: sparse convolution, rows grouped by dy in order of first appearance */
static void conv_sparse(const float *src, size_t w, size_t h, const int32_t *tdy, const int32_t *tdx, const float *twt,
                        size_t nt, float *out) {
    int32_t rdy[64];
    size_t nrows = 0, rcount[64] = {0}, ridx[64][64];
    for (size_t t = 0; t < nt; t++) {
        size_t r = 0;
        while (r < nrows && rdy[r] != tdy[t]) r++;
        if (r == nrows) rdy[nrows++] = tdy[t];
        ridx[r][rcount[r]++] = t;
    }
#pragma omp parallel for schedule(static) if (w * h >= PAR_MIN)
    for (size_t y = 0; y < h; y++) {
        float *acc = out + y * w;
        for (size_t x = 0; x < w; x++) acc[x] = 0.0f;
        for (size_t r = 0; r < nrows; r++) {
            const float *s = src + RF((long)y + rdy[r], h) * w;
            for (size_t q = 0; q < rcount[r]; q++) {
                size_t t = ridx[r][q];
                long dx = tdx[t];
                float wt = twt[t];
                for (size_t x = 0; x < w; x++) acc[x] += wt * s[RF((long)x + dx, w)];
            }
        }
    }
}

/* depth.rs:474-495 */
void ora_focus_measure(const float *y, size_t w, size_t h, const ora_depth_params *p, float *out) {
    if (p->focus_kind == 0) {
        int32_t dy[256], dx[256];
        float wt[256];
        size_t nt = ora_rdf_taps(p->r_in, p->r_out, dy, dx, wt);
        conv_sparse(y, w, h, dy, dx, wt, nt, out);
        for (size_t i = 0; i < w * h; i++) out[i] = fabsf(out[i]);
    } else {
        long s = p->sml_step > 1 ? (long)p->sml_step : 1;
#pragma omp parallel for schedule(static) if (w * h >= PAR_MIN)
        for (size_t yy = 0; yy < h; yy++) {
            size_t up = RF((long)yy - s, h), dn = RF((long)yy + s, h);
            const float *c = y + yy * w, *u = y + up * w, *d = y + dn * w;
            for (size_t x = 0; x < w; x++) {
                float l = c[RF((long)x - s, w)], r = c[RF((long)x + s, w)];
                out[yy * w + x] = fabsf(2.0f * c[x] - l - r) + fabsf(2.0f * c[x] - u[x] - d[x]);
            }
        }
    }
}

/* depth.rs:1097-1100 */
void ora_focus_slice(const float *img3, size_t w, size_t h, const ora_depth_params *p, float *out) {
    float *y = fzeros(w * h), *f = fzeros(w * h);
    ora_luma(img3, w * h, y);
    ora_focus_measure(y, w, h, p, f);
    ora_block_mean(f, w, h, (size_t)1 << p->scale, out);
    free(y);
    free(f);
}

/* depth.rs:499-611 */
struct ora_peaks {
    size_t m, n;
    float *c1, *l1, *r1, *c2, *prev, *prev2, *sum, *cmin;
    uint32_t *i1;
};

ora_peaks *ora_peaks_new(size_t m) {
    ora_peaks *t = (ora_peaks *)xcalloc(1, sizeof(ora_peaks));
    t->m = m;
    t->c1 = fzeros(m); t->l1 = fzeros(m); t->r1 = fzeros(m); t->c2 = fzeros(m);
    t->prev = fzeros(m); t->prev2 = fzeros(m); t->sum = fzeros(m); t->cmin = fzeros(m);
    t->i1 = (uint32_t *)xcalloc(m, sizeof(uint32_t));
    for (size_t i = 0; i < m; i++) {
        t->c1[i] = t->l1[i] = t->r1[i] = t->c2[i] = -1.0f;
        t->cmin[i] = INFINITY;
    }
    return t;
}

static inline void px_register(ora_peaks *t, size_t i, float val, uint32_t idx, float l, float r) {
    if (val > t->c1[i]) {
        if (t->c1[i] >= 0.0f) t->c2[i] = t->c1[i];
        t->c1[i] = val;
        t->i1[i] = idx;
        t->l1[i] = l;
        t->r1[i] = r;
    } else if (val > t->c2[i]) {
        t->c2[i] = val;
    }
}

void ora_peaks_push(ora_peaks *t, const float *c) {
    size_t m = t->n;
#pragma omp parallel for schedule(static) if (t->m >= PAR_MIN)
    for (size_t i = 0; i < t->m; i++) {
        float v = c[i];
        if (m >= 1) {
            int is_peak = (m == 1 || t->prev[i] >= t->prev2[i]) && t->prev[i] > v;
            if (is_peak) {
                float l = m >= 2 ? t->prev2[i] : -1.0f;
                px_register(t, i, t->prev[i], (uint32_t)(m - 1), l, v);
            }
        }
        t->sum[i] += v;
        t->cmin[i] = rminf(t->cmin[i], v);
        t->prev2[i] = t->prev[i];
        t->prev[i] = v;
    }
    t->n++;
}

void ora_peaks_finish(ora_peaks *t, float gate, float *depth, float *conf, float *mins, float *floor_out) {
    size_t n = t->n, M = t->m;
    for (size_t i = 0; i < M; i++)
        if (n == 1 || t->prev[i] >= t->prev2[i]) {
            float l = n >= 2 ? t->prev2[i] : -1.0f;
            px_register(t, i, t->prev[i], (uint32_t)(n - 1), l, -1.0f);
        }
    size_t ns = (M + 6) / 7;
    float *smp = fzeros(ns);
    for (size_t i = 0, k = 0; i < M; i += 7, k++) smp[k] = t->cmin[i];
    qsort(smp, ns, sizeof(float), cmp_total);
    float floor_ = smp[ns / 2];
    free(smp);
    float inv_n = 1.0f / (float)n;
#pragma omp parallel for schedule(static) if (M >= PAR_MIN)
    for (size_t i = 0; i < M; i++) {
        float c1 = t->c1[i], l = t->l1[i], r = t->r1[i];
        float delta = 0.0f;
        if (l >= 0.0f && r >= 0.0f && c1 > 0.0f) {
            float ll = logf(rmaxf(l, 1e-12f)), lc = logf(c1), lr = logf(rmaxf(r, 1e-12f));
            float den = ll - 2.0f * lc + lr;
            if (den < 0.0f) delta = rclampf(0.5f * (ll - lr) / den, -0.5f, 0.5f);
        }
        depth[i] = (float)t->i1[i] + delta;
        if (c1 > 0.0f) {
            float mean = t->sum[i] * inv_n;
            float prom = rclampf(1.0f - mean / c1, 0.0f, 1.0f);
            float pkr = t->c2[i] >= 0.0f ? rclampf(1.0f - t->c2[i] / c1, 0.0f, 1.0f) : 1.0f;
            float g = (gate > 0.0f && floor_ > 0.0f) ? rclampf((c1 - floor_) / (gate * floor_), 0.0f, 1.0f) : 1.0f;
            conf[i] = prom * pkr * g;
        } else {
            conf[i] = 0.0f;
        }
        if (mins) mins[i] = isfinite(t->cmin[i]) ? t->cmin[i] : 0.0f;
    }
    if (floor_out) *floor_out = floor_;
}

void ora_peaks_free(ora_peaks *t) {
    if (!t) return;
    free(t->c1); free(t->l1); free(t->r1); free(t->c2);
    free(t->prev); free(t->prev2); free(t->sum); free(t->cmin); free(t->i1);
    free(t);
}

/* depth.rs:615-623 */
float ora_normalize_conf(float *conf, size_t n) {
    size_t ns = (n + 6) / 7;
    float *smp = fzeros(ns);
    for (size_t i = 0, k = 0; i < n; i += 7, k++) smp[k] = conf[i];
    size_t k = ns * 9 / 10 < ns - 1 ? ns * 9 / 10 : ns - 1;
    qsort(smp, ns, sizeof(float), cmp_total);
    float p90 = smp[k];
    free(smp);
    if (p90 > 1e-6f)
        for (size_t i = 0; i < n; i++) conf[i] = rminf(conf[i] / p90, 1.0f);
    return p90;
}

/* depth.rs:628-650 */
void ora_median3(const float *src, size_t w, size_t h, float *out) {
#pragma omp parallel for schedule(static) if (w * h >= PAR_MIN)
    for (size_t y = 0; y < h; y++) {
        const float *rows[3];
        for (int k = 0; k < 3; k++) rows[k] = src + RF((long)y + k - 1, h) * w;
        for (size_t x = 0; x < w; x++) {
            size_t xs[3] = {RF((long)x - 1, w), x, RF((long)x + 1, w)};
            float v[9];
            int k = 0;
            for (int r = 0; r < 3; r++)
                for (int q = 0; q < 3; q++) v[k++] = rows[r][xs[q]];
            /* insertion sort by total order */
            for (int i = 1; i < 9; i++) {
                float t = v[i];
                int j = i;
                while (j > 0 && total_key(v[j - 1]) > total_key(t)) {
                    v[j] = v[j - 1];
                    j--;
                }
                v[j] = t;
            }
            out[y * w + x] = v[4];
        }
    }
}

/* depth.rs:653-672 */
static void edge_weights(const float *guide, size_t w, size_t h, float sigma, float *ax, float *ay) {
    float inv = -1.0f / rmaxf(sigma, 1e-6f);
    memset(ax, 0, w * h * sizeof(float));
    memset(ay, 0, w * h * sizeof(float));
#pragma omp parallel for schedule(static) if (w * h >= PAR_MIN)
    for (size_t y = 0; y < h; y++) {
        const float *g = guide + y * w;
        for (size_t x = 0; x + 1 < w; x++) ax[y * w + x] = expf(fabsf(g[x] - g[x + 1]) * inv);
        if (y + 1 < h)
            for (size_t x = 0; x < w; x++) ay[y * w + x] = expf(fabsf(g[x] - guide[(y + 1) * w + x]) * inv);
    }
}

/* depth.rs:677-702 */
void ora_solve_rows(float *u, const float *f, const float *wd, const float *a, size_t w, size_t h, float lam) {
#pragma omp parallel for schedule(static) if (w * h >= PAR_MIN)
    for (size_t y = 0; y < h; y++) {
        size_t o = y * w;
        float *cp = fzeros(w), *dp = fzeros(w);
        float prev_a = 0.0f;
        for (size_t i = 0; i < w; i++) {
            float ai = i + 1 < w ? a[o + i] : 0.0f;
            float diag = wd[o + i] + lam * (prev_a + ai);
            float lower = -lam * prev_a;
            float upper = -lam * ai;
            float m = diag - lower * (i > 0 ? cp[i - 1] : 0.0f);
            if (fabsf(m) < 1e-12f) m = 1e-12f;
            cp[i] = upper / m;
            dp[i] = (wd[o + i] * f[o + i] - lower * (i > 0 ? dp[i - 1] : 0.0f)) / m;
            prev_a = ai;
        }
        float *row = u + o;
        row[w - 1] = dp[w - 1];
        for (size_t i = w - 1; i-- > 0;) row[i] = dp[i] - cp[i] * row[i + 1];
        free(cp);
        free(dp);
    }
}

/* depth.rs:707-746 */
void ora_solve_cols(float *u, const float *f, const float *wd, const float *a, size_t w, size_t h, float lam) {
#pragma omp parallel for schedule(static) if (w * h >= PAR_MIN)
    for (size_t x = 0; x < w; x++) {
        float *cp = fzeros(h), *dp = fzeros(h);
        for (size_t y = 0; y < h; y++) {
            size_t i = y * w + x;
            float prev_a = y > 0 ? a[i - w] : 0.0f;
            float ai = y + 1 < h ? a[i] : 0.0f;
            float diag = wd[i] + lam * (prev_a + ai);
            float lower = -lam * prev_a;
            float upper = -lam * ai;
            float c0 = y > 0 ? cp[y - 1] : 0.0f, d0 = y > 0 ? dp[y - 1] : 0.0f;
            float m = diag - lower * c0;
            if (fabsf(m) < 1e-12f) m = 1e-12f;
            cp[y] = upper / m;
            dp[y] = (wd[i] * f[i] - lower * d0) / m;
        }
        u[(h - 1) * w + x] = dp[h - 1];
        for (size_t y = h - 1; y-- > 0;) u[y * w + x] = dp[y] - cp[y] * u[(y + 1) * w + x];
        free(cp);
        free(dp);
    }
}

/* depth.rs:753-898 */
#define MG_OMEGA 0.8f
#define MG_COARSE 32
#define MG_MIN 16

typedef struct {
    size_t w, h;
    float *wd, *ax, *ay, *dinv, *x, *b, *r;
} mglev_t;

static void mg_alloc(mglev_t *L, size_t w, size_t h) {
    size_t n = w * h;
    L->w = w;
    L->h = h;
    L->wd = fzeros(n); L->ax = fzeros(n); L->ay = fzeros(n); L->dinv = fzeros(n);
    L->x = fzeros(n); L->b = fzeros(n); L->r = fzeros(n);
}
static void mg_free(mglev_t *L) {
    free(L->wd); free(L->ax); free(L->ay); free(L->dinv); free(L->x); free(L->b); free(L->r);
}

/* mode 0: out = Ax; 1: out = b - Ax; 2: out = x + w*dinv*(b - Ax) */
static void mg_apply(const mglev_t *L, const float *x, float *out, int mode) {
    size_t w = L->w, h = L->h;
    const float *wd = L->wd, *ax = L->ax, *ay = L->ay;
#pragma omp parallel for schedule(static) if (w * h >= PAR_MIN)
    for (size_t y = 0; y < h; y++)
        for (size_t i = 0; i < w; i++) {
            size_t k = y * w + i;
            float xk = x[k];
            float v = wd[k] * xk;
            if (i > 0) v += ax[k - 1] * (xk - x[k - 1]);
            if (i + 1 < w) v += ax[k] * (xk - x[k + 1]);
            if (y > 0) v += ay[k - w] * (xk - x[k - w]);
            if (y + 1 < h) v += ay[k] * (xk - x[k + w]);
            if (mode == 0)
                out[k] = v;
            else if (mode == 1)
                out[k] = L->b[k] - v;
            else
                out[k] = x[k] + MG_OMEGA * L->dinv[k] * (L->b[k] - v);
        }
}

static void mg_set_dinv(mglev_t *L) {
    size_t w = L->w, h = L->h;
#pragma omp parallel for schedule(static) if (w * h >= PAR_MIN)
    for (size_t y = 0; y < h; y++)
        for (size_t i = 0; i < w; i++) {
            size_t k = y * w + i;
            float v = L->wd[k];
            if (i > 0) v += L->ax[k - 1];
            if (i + 1 < w) v += L->ax[k];
            if (y > 0) v += L->ay[k - w];
            if (y + 1 < h) v += L->ay[k];
            L->dinv[k] = 1.0f / rmaxf(v, 1e-12f);
        }
}

static void mg_jacobi(mglev_t *L, int zero_start) {
    size_t n = L->w * L->h;
    if (zero_start) {
        for (size_t k = 0; k < n; k++) L->x[k] = MG_OMEGA * L->dinv[k] * L->b[k];
        return;
    }
    mg_apply(L, L->x, L->r, 2);
    float *t = L->x;
    L->x = L->r;
    L->r = t;
}

static void mg_coarsen_wd(const mglev_t *F, mglev_t *C) {
    size_t w = F->w, h = F->h, cw = C->w, ch = C->h;
#pragma omp parallel for schedule(static) if (cw * ch >= PAR_MIN)
    for (size_t yy = 0; yy < ch; yy++) {
        float *row = C->wd + yy * cw;
        for (size_t xx = 0; xx < cw; xx++) row[xx] = 0.0f;
        size_t ye = 2 * yy + 2 < h ? 2 * yy + 2 : h;
        for (size_t y = 2 * yy; y < ye; y++)
            for (size_t xx = 0; xx < cw; xx++) {
                size_t xe = 2 * xx + 2 < w ? 2 * xx + 2 : w;
                float s = 0.0f;
                for (size_t x = 2 * xx; x < xe; x++) s += F->wd[y * w + x];
                row[xx] += s;
            }
    }
}

static void mg_coarsen(const mglev_t *F, mglev_t *C) {
    size_t w = F->w, h = F->h, cw = half(w), ch = half(h);
    mg_alloc(C, cw, ch);
    mg_coarsen_wd(F, C);
    for (size_t yy = 0; yy < ch; yy++) {
        size_t ye = 2 * yy + 2 < h ? 2 * yy + 2 : h;
        for (size_t y = 2 * yy; y < ye; y++)
            for (size_t xx = 0; xx + 1 < cw; xx++) C->ax[yy * cw + xx] += F->ax[y * w + 2 * xx + 1];
        if (yy + 1 < ch) {
            size_t y = 2 * yy + 1;
            for (size_t xx = 0; xx < cw; xx++) {
                size_t xe = 2 * xx + 2 < w ? 2 * xx + 2 : w;
                float s = 0.0f;
                for (size_t x = 2 * xx; x < xe; x++) s += F->ay[y * w + x];
                C->ay[yy * cw + xx] = s;
            }
        }
    }
}

/* depth.rs:901-921 */
static void prolong_add(float *x, size_t w, size_t h, const float *xc, size_t cw) {
#pragma omp parallel for schedule(static) if (w * h >= PAR_MIN)
    for (size_t y = 0; y < h; y++)
        for (size_t i = 0; i < w; i++) x[y * w + i] += xc[(y / 2) * cw + i / 2];
}
static void restrict_(const float *r, size_t w, size_t h, float *bc, size_t cw, size_t ch) {
#pragma omp parallel for schedule(static) if (cw * ch >= PAR_MIN)
    for (size_t yy = 0; yy < ch; yy++) {
        float *row = bc + yy * cw;
        for (size_t xx = 0; xx < cw; xx++) row[xx] = 0.0f;
        size_t ye = 2 * yy + 2 < h ? 2 * yy + 2 : h;
        for (size_t y = 2 * yy; y < ye; y++)
            for (size_t xx = 0; xx < cw; xx++) {
                size_t xe = 2 * xx + 2 < w ? 2 * xx + 2 : w;
                float s = 0.0f;
                for (size_t x = 2 * xx; x < xe; x++) s += r[y * w + x];
                row[xx] += s;
            }
    }
}

/* depth.rs:924-947 */
static void vcycle(mglev_t *L, size_t nl) {
    if (nl == 0) return;
    mglev_t *top = &L[0];
    if (nl == 1) {
        for (int k = 0; k < MG_COARSE; k++) mg_jacobi(top, k == 0);
        return;
    }
    mg_jacobi(top, 1);
    mg_apply(top, top->x, top->r, 1);
    restrict_(top->r, top->w, top->h, L[1].b, L[1].w, L[1].h);
    vcycle(L + 1, nl - 1);
    prolong_add(top->x, top->w, top->h, L[1].x, L[1].w);
    mg_jacobi(top, 0);
}

/* depth.rs:958-1078 */
struct ora_wls {
    size_t w, h, nl;
    float lambda;
    float *ex, *ey;
    mglev_t L[MAXL];
    float *r, *z, *p, *ap, *f;
};

ora_wls *ora_wls_new(const float *guide, size_t w, size_t h, float lambda, float sigma_c) {
    ora_wls *s = (ora_wls *)xcalloc(1, sizeof(ora_wls));
    size_t n = w * h;
    s->w = w;
    s->h = h;
    s->lambda = lambda;
    s->ex = fzeros(n);
    s->ey = fzeros(n);
    edge_weights(guide, w, h, sigma_c, s->ex, s->ey);
    mg_alloc(&s->L[0], w, h);
    for (size_t i = 0; i < n; i++) {
        s->L[0].ax[i] = lambda * s->ex[i];
        s->L[0].ay[i] = lambda * s->ey[i];
    }
    s->nl = 1;
    while ((s->L[s->nl - 1].w > s->L[s->nl - 1].h ? s->L[s->nl - 1].w : s->L[s->nl - 1].h) > MG_MIN) {
        mg_coarsen(&s->L[s->nl - 1], &s->L[s->nl]);
        s->nl++;
    }
    s->r = fzeros(n); s->z = fzeros(n); s->p = fzeros(n); s->ap = fzeros(n); s->f = fzeros(n);
    return s;
}

size_t ora_wls_levels(const ora_wls *s) { return s->nl; }

void ora_wls_set_weights(ora_wls *s, const float *conf) {
    size_t n = s->w * s->h;
    for (size_t i = 0; i < n; i++) s->L[0].wd[i] = rmaxf(conf[i], 0.0f) + 1e-4f;
    for (size_t l = 0; l < s->nl; l++) {
        if (l > 0) mg_coarsen_wd(&s->L[l - 1], &s->L[l]);
        mg_set_dinv(&s->L[l]);
    }
}

static double dot(const float *a, const float *b, size_t w, size_t h) {
    double *part = (double *)xcalloc(h, sizeof(double));
#pragma omp parallel for schedule(static) if (w * h >= PAR_MIN)
    for (size_t y = 0; y < h; y++) {
        double s = 0.0;
        for (size_t x = 0; x < w; x++) s += (double)a[y * w + x] * (double)b[y * w + x];
        part[y] = s;
    }
    double s = 0.0;
    for (size_t y = 0; y < h; y++) s += part[y];
    free(part);
    return s;
}

static void precond(ora_wls *s, const float *r, float *z) {
    size_t n = s->w * s->h;
    memcpy(s->L[0].b, r, n * sizeof(float));
    vcycle(s->L, s->nl);
    memcpy(z, s->L[0].x, n * sizeof(float));
}

void ora_wls_vcycle(ora_wls *s, const float *r, float *z) { precond(s, r, z); }
void ora_wls_apply(const ora_wls *s, const float *x, float *ax) { mg_apply(&s->L[0], x, ax, 0); }

uint32_t ora_wls_solve(ora_wls *s, const float *d, uint32_t max_iters, float tol, float *u, float *rel_out) {
    size_t w = s->w, h = s->h, n = w * h;
    float lambda = s->lambda;
    const float *wd = s->L[0].wd;
    if (w == 1 || h == 1) {
        if (h == 1)
            ora_solve_rows(u, d, wd, s->ex, w, h, lambda);
        else
            ora_solve_cols(u, d, wd, s->ey, w, h, lambda);
        if (rel_out) *rel_out = 0.0f;
        return 0;
    }
    const int T = 3;
    for (size_t i = 0; i < n; i++) s->p[i] = 1.0f;
    const float *ones = s->p;
    float *f = s->f;
    memcpy(f, d, n * sizeof(float));
    for (int t = 1; t <= T; t++) {
        float lam_t = 1.5f * lambda * powf(4.0f, (float)(T - t)) / (powf(4.0f, (float)T) - 1.0f);
        const float *wt = t == 1 ? wd : ones;
        ora_solve_rows(u, f, wt, s->ex, w, h, lam_t);
        memcpy(f, u, n * sizeof(float));
        ora_solve_cols(u, f, ones, s->ey, w, h, lam_t);
        memcpy(f, u, n * sizeof(float));
    }
    float *b = s->f;
    for (size_t i = 0; i < n; i++) b[i] = wd[i] * d[i];
    double bnorm = sqrt(dot(b, b, w, h));
    if (bnorm < 1e-30) bnorm = 1e-30;
    memcpy(s->L[0].b, b, n * sizeof(float)); /* mg_apply mode 1 reads L->b */
    mg_apply(&s->L[0], u, s->r, 1);
    precond(s, s->r, s->z);
    memcpy(s->p, s->z, n * sizeof(float));
    double rz = dot(s->r, s->z, w, h);
    float rel = (float)(sqrt(dot(s->r, s->r, w, h)) / bnorm);
    uint32_t iters = 0;
    for (uint32_t it = 0; it < max_iters; it++) {
        if (rel < tol) break;
        iters++;
        mg_apply(&s->L[0], s->p, s->ap, 0);
        double pap = dot(s->p, s->ap, w, h);
        if (pap <= 0.0) break;
        float alpha = (float)(rz / pap);
        for (size_t i = 0; i < n; i++) u[i] += alpha * s->p[i];
        for (size_t i = 0; i < n; i++) s->r[i] -= alpha * s->ap[i];
        rel = (float)(sqrt(dot(s->r, s->r, w, h)) / bnorm);
        if (rel < tol) break;
        precond(s, s->r, s->z);
        double rz_new = dot(s->r, s->z, w, h);
        float beta = (float)(rz_new / rz);
        rz = rz_new;
        for (size_t i = 0; i < n; i++) s->p[i] = s->z[i] + beta * s->p[i];
    }
    if (rel_out) *rel_out = rel;
    return iters;
}

void ora_wls_free(ora_wls *s) {
    if (!s) return;
    for (size_t l = 0; l < s->nl; l++) mg_free(&s->L[l]);
    free(s->ex); free(s->ey); free(s->r); free(s->z); free(s->p); free(s->ap); free(s->f);
    free(s);
}

/* depth.rs:1168: the Huber IRLS weights, conf * min(tau / |d - u|, 1) */
void ora_huber_weights(const float *conf, const float *d, const float *u, size_t n, float tau, float *out) {
    for (size_t i = 0; i < n; i++) out[i] = conf[i] * rminf(tau / fabsf(d[i] - u[i]), 1.0f);
}

/* depth.rs:1125-1159: the aggregation and the streamed peak search over the n
 * slices (guide = block-mean fused luma), the median, the p90 normalisation;
 * guide, depth, conf (dw*dh each) are the WLS system's inputs; mins the per-cell
 * minima, *floor the noise floor, *p90 the normalisation.  Returns the aggregation's
 * and the peak search's wall time in t2 (nullable). */
static void depth_grid_(const float *slices, size_t n, const float *fused3, size_t w, size_t h,
                        const ora_depth_params *p, float *y_full, float *guide, float *depth_w, float *conf,
                        float *mins, float *floor_, float *p90, double *t2) {
    size_t k = (size_t)1 << p->scale;
    size_t dw = blocks(w, k), dh = blocks(h, k), M = dw * dh;
    ora_luma(fused3, w * h, y_full);
    ora_block_mean(y_full, w, h, k, guide);
    double t_agg = now_s();
    guided_t gf;
    guided_init(&gf, guide, dw, dh, p->agg_radius, p->agg_eps);
    float *agg = fzeros(M);
    ora_peaks *tr = ora_peaks_new(M);
    for (size_t m = 0; m < n; m++) {
        const float *c = slices + m * M;
        if (p->agg_radius > 0) {
            guided_filter_into(&gf, c, agg);
            for (size_t i = 0; i < M; i++) agg[i] = rmaxf(agg[i], 0.0f);
            ora_peaks_push(tr, agg);
        } else {
            ora_peaks_push(tr, c);
        }
    }
    guided_free(&gf);
    free(agg);
    t_agg = now_s() - t_agg;
    double t_peaks = now_s();
    ora_peaks_finish(tr, p->gate, depth_w, conf, mins, floor_);
    ora_peaks_free(tr);
    if (p->median) {
        float *m2 = fzeros(M);
        ora_median3(depth_w, dw, dh, m2);
        memcpy(depth_w, m2, M * sizeof(float));
        free(m2);
    }
    *p90 = ora_normalize_conf(conf, M);
    if (t2) {
        t2[0] = t_agg;
        t2[1] = now_s() - t_peaks;
    }
}

int ora_depth_grid(const float *slices, size_t n, const float *fused3, size_t w, size_t h, const ora_depth_params *p,
                   float *guide, float *depth, float *conf, float *p90) {
    if (n == 0) return -1;
    size_t k = (size_t)1 << p->scale, M = blocks(w, k) * blocks(h, k);
    float *y_full = fzeros(w * h), *mins = fzeros(M), floor_;
    depth_grid_(slices, n, fused3, w, h, p, y_full, guide, depth, conf, mins, &floor_, p90, NULL);
    free(y_full);
    free(mins);
    return 0;
}

/* depth.rs:1119-1200 */
int ora_depth_from_slices(const float *slices, size_t n, const float *fused3, size_t w, size_t h,
                          const ora_depth_params *p, float *depth, float *conf_out, float *floor_plane,
                          ora_depth_out *info) {
    if (n == 0) return -1;
    size_t k = (size_t)1 << p->scale;
    size_t dw = blocks(w, k), dh = blocks(h, k), M = dw * dh;
    float *y_full = fzeros(w * h), *guide = fzeros(M);
    float *depth_w = fzeros(M), *conf = fzeros(M), *mins = fzeros(M), floor_, p90;
    double t2[2];
    depth_grid_(slices, n, fused3, w, h, p, y_full, guide, depth_w, conf, mins, &floor_, &p90, t2);
    double t_agg = t2[0], t_peaks = t2[1];
    double t_wls = now_s();
    uint32_t it1 = 0, it2 = 0;
    float rel1 = 0.0f, rel2 = 0.0f;
    if (p->lambda > 0.0f) {
        ora_wls *s = ora_wls_new(guide, dw, dh, p->lambda, p->sigma_c);
        ora_wls_set_weights(s, conf);
        float *u = fzeros(M);
        it1 = ora_wls_solve(s, depth_w, p->cg_iters, p->tol, u, &rel1);
        rel2 = rel1;
        it2 = it1;
        if (p->robust > 0.0f) {
            float *w2 = fzeros(M);
            ora_huber_weights(conf, depth_w, u, M, p->robust, w2);
            ora_wls_set_weights(s, w2);
            free(w2);
            it2 = ora_wls_solve(s, depth_w, p->cg_iters, p->tol, u, &rel2);
        }
        ora_wls_free(s);
        free(depth_w);
        depth_w = u;
    }
    t_wls = now_s() - t_wls;
    double t_up = now_s();
    float max_d = (float)(n - 1);
    if (p->upsample_kind == 0) {
        ora_upsample_bilinear(depth_w, dw, dh, w, h, k, depth);
    } else if (k == 1) {
        ora_guided_filter(guide, depth_w, dw, dh, p->up_radius, p->up_eps, depth);
    } else {
        float *a = fzeros(M), *b = fzeros(M), *au = fzeros(w * h), *bu = fzeros(w * h);
        ora_guided_coeffs(guide, depth_w, dw, dh, p->up_radius, p->up_eps, a, b);
        ora_upsample_bilinear(a, dw, dh, w, h, k, au);
        ora_upsample_bilinear(b, dw, dh, w, h, k, bu);
        for (size_t i = 0; i < w * h; i++) depth[i] = au[i] * y_full[i] + bu[i];
        free(a); free(b); free(au); free(bu);
    }
    for (size_t i = 0; i < w * h; i++) depth[i] = rclampf(depth[i], 0.0f, max_d);
    if (k == 1)
        memcpy(conf_out, conf, M * sizeof(float));
    else
        ora_upsample_bilinear(conf, dw, dh, w, h, k, conf_out);
    if (floor_plane) memcpy(floor_plane, mins, M * sizeof(float));
    if (info) {
        info->dw = (uint32_t)dw;
        info->dh = (uint32_t)dh;
        info->iters1 = it1;
        info->iters2 = it2;
        info->rel1 = rel1;
        info->rel2 = rel2;
        info->p90 = p90;
        info->floor = floor_;
        info->agg_s = t_agg;
        info->peaks_s = t_peaks;
        info->wls_s = t_wls;
        info->up_s = now_s() - t_up;
    }
    free(depth_w); free(conf); free(mins); free(guide); free(y_full);
    return 0;
}

/* ---- wav.rs: the weighted average ---- */

/* wav.rs:88-91: a cell's weight, the contrast above (1 + gate) x the noise
 * floor raised to power, plus the even weight of a cell nothing rises above */
float ora_wav_weight(float c, float f, float gate, float power) {
    return powf(rmaxf(c - (1.0f + gate) * f, 0.0f), power) + powf(1e-2f * f, power) + 1e-30f;
}

/* wav.rs:97-161 weighted_average over n aligned frames (n x 3*w*h planes):
 * per frame the focus slice (luma, focus measure, block mean), its box mean
 * (smooth > 0), the weights, their smoothing (the guided filter on the fused
 * luma's grid means when edge and a guide are given, else the box mean),
 * the bilinear upsample, and the weighted sums; then the clamped quotient.
 * floor: dw*dh (the depth pass's noise floor).  guide3: the fused image or NULL. */
void ora_weighted_average(const float *frames, size_t n, size_t w, size_t h, const ora_depth_params *dp, float power,
                          size_t smooth, float gate, int edge, const float *floor, const float *guide3, float *out3) {
    const size_t k = (size_t)1 << dp->scale, dw = blocks(w, k), dh = blocks(h, k), m = dw * dh, N = w * h;
    power = rmaxf(power, 0.0f);
    gate = rmaxf(gate, 0.0f);
    const int have_bf = smooth > 0, have_gf = edge && smooth > 0 && guide3 != NULL;
    boxf_t bf;
    if (have_bf) boxf_init(&bf, dw, dh, smooth);
    guided_t gf;
    float *gy = NULL;
    if (have_gf) {
        float *y = fzeros(N);
        ora_luma(guide3, N, y);
        gy = fzeros(m);
        ora_block_mean(y, w, h, k, gy);
        free(y);
        guided_init(&gf, gy, dw, dh, smooth, dp->agg_eps);
    }
    float *acc = fzeros(3 * N), *wsum = fzeros(N), *g = fzeros(m), *t = fzeros(m), *wf = fzeros(N);
    for (size_t q = 0; q < n; q++) {
        const float *f = frames + q * 3 * N;
        ora_focus_slice(f, w, h, dp, g);
        if (have_bf) {
            boxf_mean(&bf, g, t);
            memcpy(g, t, m * sizeof(float));
        }
        for (size_t i = 0; i < m; i++) t[i] = ora_wav_weight(g[i], floor[i], gate, power);
        if (have_gf) {
            guided_filter_into(&gf, t, g);
            for (size_t i = 0; i < m; i++) g[i] = rmaxf(g[i], 1e-30f);  /* a weight stays a weight */
        } else if (have_bf) {
            boxf_mean(&bf, t, g);
        } else {
            memcpy(g, t, m * sizeof(float));
        }
        ora_upsample_bilinear(g, dw, dh, w, h, k, wf);
        for (size_t c = 0; c < 3; c++) {
            float *a = acc + c * N;
            const float *s = f + c * N;
#pragma omp parallel for schedule(static) if (N >= PAR_MIN)
            for (size_t i = 0; i < N; i++) a[i] += wf[i] * s[i];
        }
        for (size_t i = 0; i < N; i++) wsum[i] += wf[i];
    }
    for (size_t c = 0; c < 3; c++)
        for (size_t i = 0; i < N; i++) out3[c * N + i] = rclampf(acc[c * N + i] / wsum[i], 0.0f, 1.0f);
    free(acc);
    free(wsum);
    free(g);
    free(t);
    free(wf);
    if (have_bf) boxf_free(&bf);
    if (have_gf) {
        guided_free(&gf);
        free(gy);
    }
}

/* stack.rs:76-92: the frame ranges [lo, hi] of the slabs; returns their count
 * (at most cap written, two entries each) */
size_t ora_slab_ranges(size_t count, size_t size, size_t overlap, size_t *out2, size_t cap) {
    if (count == 0) return 0;
    size = size < 1 ? 1 : (size > count ? count : size);
    overlap = overlap < size - 1 ? overlap : size - 1;
    size_t lo = 0, ns = 0;
    for (;;) {
        size_t hi = lo + size - 1 < count - 1 ? lo + size - 1 : count - 1;
        if (ns < cap) {
            out2[2 * ns] = lo;
            out2[2 * ns + 1] = hi;
        }
        ns++;
        if (hi + 1 >= count) return ns;
        lo = hi + 1 - overlap;
    }
}

/* ======================================================================
 * io.rs and stack.rs
 * ==================================================================== */

/* io.rs:51-58 and :89: v * (1/65535 as f32) */
void ora_u16_to_planes(const uint16_t *rgb, size_t n, float *img3) {
    const float inv16 = 1.0f / 65535.0f;
    for (int c = 0; c < 3; c++)
        for (size_t i = 0; i < n; i++) img3[c * n + i] = (float)rgb[3 * i + c] * inv16;
}

/* io.rs:287 */
void ora_planes_to_u16(const float *img3, size_t n, uint16_t *rgb) {
    for (size_t i = 0; i < n; i++)
        for (int c = 0; c < 3; c++) {
            float v = rclampf(img3[c * n + i], 0.0f, 1.0f) * 65535.0f + 0.5f;
            rgb[3 * i + c] = (uint16_t)f2usize(v < 65535.0f ? v : 65535.0f);
        }
}

/* stack.rs:740-826 (fuse_range), :513-530 (register), :638-706 (AlignedFrames::get),
 * :361-419 (LazyFrames brightness), :833-853 (fuse_and_depth) */
int ora_stack_run(const uint16_t *frames, size_t n, size_t w, size_t h, const ora_fuse_params *fp,
                  const ora_stack_params *sp, const ora_depth_params *dp, float *img3_out, float *winner_full,
                  double *sims, float *gains, float *depth, float *conf, float *slices_out, ora_depth_out *dinfo,
                  ora_stack_times *times, uint32_t *evals) {
    if (n == 0) return -1;
    size_t N = w * h;
    double t_all = now_s();
    ora_stack_times T = {0};
    ora_fuser *fu = ora_fuser_new(w, h, fp);
    size_t M = 0;
    float *slices = NULL;
    if (sp->depth) {
        size_t k = (size_t)1 << dp->scale;
        M = blocks(w, k) * blocks(h, k);
        slices = fzeros(n * M);
    }
    float *img = fzeros(3 * N), *y = fzeros(N), *prev_ref = NULL, *grid = NULL;
    double last_sim[8], ref_means[3];
    memcpy(last_sim, SIM_ID, sizeof(last_sim));
    int aligned = sp->align && n > 1;
    if (evals) memset(evals, 0, n * (1 + ORA_MAXL) * sizeof(uint32_t));
    for (size_t i = 0; i < n; i++) {
        ora_u16_to_planes(frames + i * 3 * N, N, img);
        double sim[8];
        float g[3] = {1.0f, 1.0f, 1.0f};
        memcpy(sim, SIM_ID, sizeof(sim));
        if (aligned) {
            double t0 = now_s();
            ora_luma(img, N, y);
            if (prev_ref) {
                ora_multiscale_align_trace(prev_ref, y, w, h, last_sim, sp->free8, sp->coarsen, sim, NULL,
                                           evals ? evals + i * (1 + ORA_MAXL) + 1 : NULL, NULL, 0, NULL);
                if (evals)
                    for (int l = 0; l < ORA_MAXL; l++) evals[i * (1 + ORA_MAXL)] += evals[i * (1 + ORA_MAXL) + 1 + l];
            }
            if (i + 1 < n) {
                if (!prev_ref) prev_ref = fzeros(N);
                if (sim_is_id(sim))
                    memcpy(prev_ref, y, N * sizeof(float));
                else
                    ora_warp_plane(y, w, h, sim, w, h, sp->interp, prev_ref, NULL);
            }
            memcpy(last_sim, sim, sizeof(sim));
            T.register_s += now_s() - t0;
            t0 = now_s();
            if (!sim_is_id(sim)) {
                float *warped = fzeros(3 * N);
                uint8_t *valid = (uint8_t *)xcalloc(N, 1);
                ora_warp_plane(img, w, h, sim, w, h, sp->interp, warped, valid);
                ora_warp_plane(img + N, w, h, sim, w, h, sp->interp, warped + N, NULL);
                ora_warp_plane(img + 2 * N, w, h, sim, w, h, sp->interp, warped + 2 * N, NULL);
                for (int c = 0; c < 3; c++)
                    for (size_t k = 0; k < N; k++)
                        if (!valid[k]) warped[c * N + k] = img[c * N + k];
                memcpy(img, warped, 3 * N * sizeof(float));
                free(warped);
                free(valid);
            }
            if (sp->brightness) {
                if (i == 0) {
                    grid = fzeros(3 * blocks(w, BSTEP) * blocks(h, BSTEP));
                    ora_bright_grid(img, w, h, grid);
                } else {
                    ora_bright_gains(grid, img, w, h, sim, g);
                    ora_bright_apply(img, N, g);
                }
            }
            T.warp_s += now_s() - t0;
        } else if (sp->brightness) {
            double t0 = now_s();
            if (i == 0) {
                ora_bright_means(img, w, h, ref_means);
            } else {
                ora_bright_gains_to(ref_means, img, w, h, g);
                ora_bright_apply(img, N, g);
            }
            T.warp_s += now_s() - t0;
        }
        if (sims) memcpy(sims + 8 * i, sim, sizeof(sim));
        if (gains) memcpy(gains + 3 * i, g, sizeof(g));
        double t0 = now_s();
        ora_fuser_push(fu, img);
        T.fold_s += now_s() - t0;
        if (sp->depth) {
            t0 = now_s();
            ora_focus_slice(img, w, h, dp, slices + i * M);
            T.slice_s += now_s() - t0;
        }
    }
    double t0 = now_s();
    ora_fuser_finish(fu, img3_out, winner_full, NULL);
    ora_fuser_free(fu);
    T.finish_s = now_s() - t0;
    int rc = 0;
    if (sp->depth) {
        t0 = now_s();
        rc = ora_depth_from_slices(slices, n, img3_out, w, h, dp, depth, conf, NULL, dinfo);
        T.depth_s = now_s() - t0;
        if (slices_out) memcpy(slices_out, slices, n * M * sizeof(float));
    }
    T.total_s = now_s() - t_all;
    if (times) *times = T;
    free(slices); free(img); free(y); free(prev_ref); free(grid);
    return rc;
}
