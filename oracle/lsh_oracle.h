/*
 * lsh_oracle.h -- CPU restatement of lapstack-core's stacking-engine path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the HIP backend is
 * compared against and the CPU baseline the benchmark times beside it.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it;
 * the product library (lapstack_amd/liblsh.so) never links or calls it.
 *
 * Every function restates, operation for operation, the reference CPU code
 * cited beside it (paths relative to crates/lapstack-core/src/ of
 * ragtux/LAPSTACK).  Floating-point order is kept (left-to-right sums, no FMA
 * contraction: the library is built with -ffp-contract=off), f64 accumulators
 * stay f64 where the reference uses them, and Rust's saturating `as` casts and
 * `total_cmp` order are reproduced.
 *
 * Planes are row-major f32; an RGB image is three consecutive planes
 * (R, G, B), each w*h.  A Sim is 8 doubles:
 * xoff, yoff, scale, rot, aspect, shear, px, py  (align.rs:199-253).
 */
#ifndef LSH_ORACLE_H
#define LSH_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* fuse.rs:69-107 FuseParams. levels <= 0 means auto_levels(w, h, 32). */
typedef struct ora_fuse_params {
    int32_t levels;
    uint32_t energy_radius;
    uint32_t top_rule;      /* 0 DevEntropy, 1 Deviation, 2 Average (fuse.rs:48-56) */
    uint32_t top_radius;
    uint32_t entropy_bins;
    uint32_t use_chroma;
    uint32_t depth_level;
    float halo;
} ora_fuse_params;

/* depth.rs:68-109 DepthParams, plus the PCG tolerance (1e-5 in depth.rs:1054). */
typedef struct ora_depth_params {
    uint32_t scale;
    uint32_t focus_kind;    /* 0 RDF, 1 SML (depth.rs:53-58) */
    uint32_t r_in, r_out, sml_step;
    uint32_t agg_radius;
    float agg_eps;
    float lambda;
    float sigma_c;
    uint32_t cg_iters;
    uint32_t median;
    uint32_t upsample_kind; /* 0 bilinear, 1 guided (depth.rs:62-66) */
    uint32_t up_radius;
    float up_eps;
    float gate;
    float robust;
    float tol;
} ora_depth_params;

/* ---- pyramid.rs ---- */
size_t ora_reflect(long i, size_t n);
void ora_reduce(const float *src, size_t w, size_t h, float *out);
void ora_expand(const float *coarse, size_t cw, size_t ch, size_t ow, size_t oh, float *out);
size_t ora_auto_levels(size_t w, size_t h, size_t min_top);
/* Laplacian pyramid, packed: level l holds 3 planes of w_l*h_l, levels 0..levels (the last is G_N). */
size_t ora_pyr_size(size_t w, size_t h, size_t levels);
void ora_pyr_build(const float *img3, size_t w, size_t h, size_t levels, float *packed);
void ora_pyr_collapse(const float *packed, size_t w, size_t h, size_t levels, float *img3);

/* ---- fuse.rs ---- */
void ora_binomial(size_t r, float *out);
void ora_window_sum(const float *src, size_t w, size_t h, const float *wt, size_t nwt, float *out);
void ora_region_energy(const float *l3, size_t w, size_t h, size_t radius, int use_chroma, float *out);
void ora_deviation_entropy(const float *y, size_t w, size_t h, size_t radius, size_t bins, float *dev, float *ent);
void ora_fuse_residuals(const float *tops, size_t n, size_t w, size_t h, const ora_fuse_params *p, float *out3);
void ora_winner_shares(const uint16_t *win, const float *best, size_t cells, size_t frames, float *out);
float ora_halo_weight(float re, float p);

typedef struct ora_fuser ora_fuser;
ora_fuser *ora_fuser_new(size_t w, size_t h, const ora_fuse_params *p);
size_t ora_fuser_levels(const ora_fuser *f);
void ora_fuser_level_dims(const ora_fuser *f, size_t l, size_t *w, size_t *h);
void ora_fuser_push(ora_fuser *f, const float *img3);
/* returns the number of shares written (0 when none) */
size_t ora_fuser_shares(const ora_fuser *f, float *out);
/* fused image (clamped), full-res winner map (nullable), winner at depth_level (nullable) */
void ora_fuser_finish(ora_fuser *f, float *img3, float *winner_full, uint16_t *winner_level);
void ora_fuser_free(ora_fuser *f);

/* ---- align.rs ---- */
/* sin and cos, correctly rounded for |x| <= pi/4 (the rotation of every Sim matrix) */
void ora_sincos(double x, double *s, double *c);
void ora_sim_matrix(const double *sim, size_t w, size_t h, double *m9);
void ora_inverse(const double *m9, double *inv9);
void ora_map(const double *inv9, double x, double y, double *sx, double *sy);
void ora_luma(const float *img3, size_t n, float *y);
void ora_warp_plane(const float *src, size_t w, size_t h, const double *sim, size_t ow, size_t oh,
                    uint32_t interp, float *out, uint8_t *valid);
double ora_warp_cost(const float *rf, size_t aw, const float *tgt, size_t tw, size_t th, const double *sim);
void ora_reduce_burt(const float *src, size_t w, size_t h, float *out);
/* number of registration levels above the plane, and their packed dims */
size_t ora_gauss_count(size_t w, size_t h, size_t *dims /* 2 per level, nullable */);
void ora_gauss_levels(const float *y, size_t w, size_t h, float *packed);
void ora_level_steps(const size_t *free_idx, size_t nfree, size_t lvl, size_t w, size_t h, double *step, double *tol);
typedef double (*ora_cost_fn)(const double *x, void *ctx);
void ora_nelder_mead(ora_cost_fn f, void *ctx, size_t n, const double *x0, const double *lo, const double *hi,
                     const double *step, const double *tol, double *out, uint32_t *evals);
#define ORA_MAXL 16
void ora_multiscale_align_trace(const float *rf, const float *tg, size_t w, size_t h, const double *init,
                                const uint8_t *free8, size_t coarsen, double *out, uint32_t *evals,
                                uint32_t *lev_evals, double *trace, size_t trace_cap, size_t *trace_n);
void ora_multiscale_align(const float *rf, const float *tg, size_t w, size_t h, const double *init,
                          const uint8_t *free8, size_t coarsen, double *out, uint32_t *evals);
void ora_common_area(const double *sims, size_t n, size_t w, size_t h, uint32_t interp, size_t *rect4);

/* ---- brightness.rs ---- */
void ora_bright_grid(const float *img3, size_t w, size_t h, float *grid3 /* 3*ceil(w/4)*ceil(h/4) */);
void ora_bright_gains(const float *grid3, const float *frame3, size_t w, size_t h, const double *sim, float *g3);
void ora_bright_ratio(const double *ref, const double *frm, size_t n, float *g3);
void ora_bright_means(const float *img3, size_t w, size_t h, double *m3);
void ora_bright_gains_to(const double *ref_means, const float *img3, size_t w, size_t h, float *g3);
void ora_bright_apply(float *img3, size_t n, const float *g3);

/* ---- depth.rs ---- */
void ora_block_mean(const float *src, size_t w, size_t h, size_t k, float *out);
void ora_upsample_bilinear(const float *g, size_t dw, size_t dh, size_t w, size_t h, size_t k, float *out);
void ora_box_mean(const float *src, size_t w, size_t h, size_t r, float *out);
void ora_guided_filter(const float *guide, const float *p, size_t w, size_t h, size_t r, float eps, float *out);
void ora_guided_coeffs(const float *guide, const float *p, size_t w, size_t h, size_t r, float eps, float *a, float *b);
size_t ora_rdf_taps(size_t r_in, size_t r_out, int32_t *dy, int32_t *dx, float *wt);
void ora_focus_measure(const float *y, size_t w, size_t h, const ora_depth_params *p, float *out);
void ora_focus_slice(const float *img3, size_t w, size_t h, const ora_depth_params *p, float *out);

typedef struct ora_peaks ora_peaks;
ora_peaks *ora_peaks_new(size_t m);
void ora_peaks_push(ora_peaks *t, const float *c);
void ora_peaks_finish(ora_peaks *t, float gate, float *depth, float *conf, float *mins, float *floor_out);
void ora_peaks_free(ora_peaks *t);
float ora_normalize_conf(float *conf, size_t n);
void ora_median3(const float *src, size_t w, size_t h, float *out);
void ora_solve_rows(float *u, const float *f, const float *wd, const float *a, size_t w, size_t h, float lam);
void ora_solve_cols(float *u, const float *f, const float *wd, const float *a, size_t w, size_t h, float lam);

typedef struct ora_wls ora_wls;
ora_wls *ora_wls_new(const float *guide, size_t w, size_t h, float lambda, float sigma_c);
size_t ora_wls_levels(const ora_wls *s);
void ora_wls_set_weights(ora_wls *s, const float *conf);
/* solve for d into u (u holds the start on entry only for the degenerate path); returns iterations */
uint32_t ora_wls_solve(ora_wls *s, const float *d, uint32_t max_iters, float tol, float *u, float *rel);
/* one V-cycle applied to r (tests: symmetry of the preconditioner) */
void ora_wls_vcycle(ora_wls *s, const float *r, float *z);
void ora_wls_apply(const ora_wls *s, const float *x, float *ax);
void ora_wls_free(ora_wls *s);

typedef struct ora_depth_out {
    uint32_t dw, dh;
    uint32_t iters1, iters2;
    float rel1, rel2;
    float p90;
    float floor;
    double agg_s, peaks_s, wls_s, up_s; /* wall time of the phases */
} ora_depth_out;
/* slices: n planes of dw*dh.  depth/conf: w*h.  floor_plane: dw*dh (nullable). */
/* depth.rs:1168 Huber IRLS weights: out = conf * min(tau / |d - u|, 1) */
void ora_huber_weights(const float *conf, const float *d, const float *u, size_t n, float tau, float *out);
/* depth.rs:1125-1159, the depth pass up to the WLS solve: guide, peak depth (after the median) and
 * p90-normalised confidence, dw*dh each; *p90 the normalisation.  Returns 0 on success. */
int ora_depth_grid(const float *slices, size_t n, const float *fused3, size_t w, size_t h, const ora_depth_params *p,
                   float *guide, float *depth, float *conf, float *p90);
int ora_depth_from_slices(const float *slices, size_t n, const float *fused3, size_t w, size_t h,
                          const ora_depth_params *p, float *depth, float *conf, float *floor_plane,
                          ora_depth_out *info);

/* ---- wav.rs ---- */
float ora_wav_weight(float c, float f, float gate, float power);
/* frames: n x 3*w*h; floor: the depth pass's grid floor; guide3: the fused image or NULL */
void ora_weighted_average(const float *frames, size_t n, size_t w, size_t h, const ora_depth_params *dp, float power,
                          size_t smooth, float gate, int edge, const float *floor, const float *guide3, float *out3);
/* stack.rs:76-92; returns the number of slabs, writes up to cap (lo, hi) pairs */
size_t ora_slab_ranges(size_t count, size_t size, size_t overlap, size_t *out2, size_t cap);

/* ---- stack.rs: the per-frame loop (fuse_range + AlignedFrames + fuse_and_depth) ---- */
typedef struct ora_stack_params {
    uint32_t align;         /* 0: frames are registered already (LazyFrames path) */
    uint8_t free8[8];
    uint32_t coarsen;
    uint32_t interp;
    uint32_t brightness;
    uint32_t depth;         /* 0: report the winner map instead */
} ora_stack_params;

typedef struct ora_stack_times {
    double register_s, warp_s, fold_s, slice_s, finish_s, depth_s, total_s;
} ora_stack_times;

/* frames: n interleaved RGB u16 images (h*w*3 each).  Outputs (host):
 * img3 3*w*h; winner_full w*h; sims n*8; gains n*3; depth, conf w*h (only with depth);
 * slices n*dw*dh (nullable).  Returns 0 on success. */
int ora_stack_run(const uint16_t *frames, size_t n, size_t w, size_t h, const ora_fuse_params *fp,
                  const ora_stack_params *sp, const ora_depth_params *dp, float *img3, float *winner_full,
                  double *sims, float *gains, float *depth, float *conf, float *slices,
                  ora_depth_out *dinfo, ora_stack_times *times,
                  uint32_t *evals /* nullable: per frame 1 + ORA_MAXL: the search's total, then per level */);

/* io.rs:89 (u16 -> f32 planes) and io.rs:287 (f32 -> u16) */
void ora_u16_to_planes(const uint16_t *rgb, size_t n, float *img3);
void ora_planes_to_u16(const float *img3, size_t n, uint16_t *rgb);

#ifdef __cplusplus
}
#endif
#endif
