/*
 * lsh.h -- C ABI of the MI355X (gfx950) stacking-engine backend.
 *
 * The backend fills the engine's backend seam as a third engine beside the
 * CPU and CUDA ones.  Reference citations are to crates/lapstack-core/src/ of
 * ragtux/LAPSTACK:
 *   Backend enum and probe ........ align.rs:27-49, gpu.rs:307-310
 *   Fuser contract (AnyFuser) ..... stack.rs:122-201, fuse.rs:349-521, gpu.rs:432-563
 *   Frames contract (GpuFrames) ... stack.rs:206-230, gpu.rs:636-900
 *   Depth contract ................ depth.rs:1119-1200, gpu.rs:1327-1468
 *
 * Conventions follow the reference's own C-ABI plugin (crates/lapstack-raw/
 * src/ffi.rs): an ABI version check, every call returns NULL on success or a
 * library-owned error string the caller releases with lsh_free_str, outputs
 * are caller-provided at documented sizes, and nothing unwinds across the
 * boundary (every C++ exception is caught and returned as a string).  A size
 * mismatch is an error string, never an abort.
 *
 * Memory: every buffer argument comes with an lsh_mem saying where it lives
 * (host memory, or device memory of the current HIP device).  Images are
 * three consecutive row-major f32 planes (R, G, B) of w*h, values in [0, 1]
 * (the reference's Img3, pyramid.rs:18-24).  Frames in 16-bit form are
 * interleaved RGB u16 (h*w*3), the decoder's layout (io.rs:80-97).
 *
 * Threading: handles are single-thread-affine.  All work of the library runs
 * in order on one HIP stream per device (lsh_stream), so a frames handle, a
 * fuser and a depth pass hand device buffers to each other without waits.
 */
#ifndef LSH_H
#define LSH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LSH_ABI_VERSION 3u

typedef enum lsh_mem { LSH_HOST = 0, LSH_DEVICE = 1 } lsh_mem;

/* Interpolation kernels, in the order of align.rs:146 Interp::ALL (Interp::id). */
enum { LSH_NEAREST = 0, LSH_BILINEAR = 1, LSH_BICUBIC = 2, LSH_SPLINE4X4 = 3, LSH_SPLINE6X6 = 4, LSH_LANCZOS3 = 5 };

/* fuse.rs:69-107 FuseParams.  levels <= 0: auto_levels(w, h, 32) (pyramid.rs:191).
 * top_rule: 0 DevEntropy, 1 Deviation, 2 Average (fuse.rs:48-56). */
typedef struct lsh_fuse_params {
    int32_t levels;
    uint32_t energy_radius;
    uint32_t top_rule;
    uint32_t top_radius;
    uint32_t entropy_bins;
    uint32_t use_chroma;
    uint32_t depth_level;
    float halo;
} lsh_fuse_params;

/* depth.rs:68-109 DepthParams.  focus_kind: 0 RDF{r_in, r_out}, 1 SML{sml_step};
 * upsample_kind: 0 bilinear, 1 guided{up_radius, up_eps}.  tol is the PCG's
 * relative-residual stop, fixed at 1e-5 in the reference (depth.rs:1054).
 * precision (no counterpart in the reference): the WLS solves' mode, 0 the
 * reference's FP32 MG-PCG (LSH_WLS_REFERENCE), 1 that MG-PCG inside FP64
 * iterative refinement (LSH_WLS_REFINED), whose FP64 true residual meets tol. */
typedef struct lsh_depth_params {
    uint32_t scale;
    uint32_t focus_kind;
    uint32_t r_in, r_out, sml_step;
    uint32_t agg_radius;
    float agg_eps;
    float lambda;
    float sigma_c;
    uint32_t cg_iters;
    uint32_t median;
    uint32_t upsample_kind;
    uint32_t up_radius;
    float up_eps;
    float gate;
    float robust;
    float tol;
    uint32_t precision;
    /* nullable (no counterpart in the reference): a learned warm start
     * (lsh_warm_*); the first WLS solve then starts from its prediction for
     * the system, the reweighted solve from the first solve's result, instead
     * of the FGS sweeps (depth.rs:1019-1034); MG-PCG still takes each to tol */
    struct lsh_warm *warm;
} lsh_depth_params;

/* depth.rs:140-156 DepthMap's scalars plus the WLS report (depth.rs:1160,
 * 1179): per solve the iterations, the solver's relative residual, and the
 * true relative residual ||W d - A u|| / ||W d|| of its FP32 result in FP64. */
typedef struct lsh_depth_info {
    uint32_t dw, dh;
    uint32_t iters1, iters2;
    float rel1, rel2;
    float p90;
    float floor;
    double true_rel1, true_rel2;
} lsh_depth_info;

typedef struct lsh_frames lsh_frames;
typedef struct lsh_fuser lsh_fuser;
typedef struct lsh_warm lsh_warm;

/* ---- library ----------------------------------------------------------- */

/* ABI version; a client refuses a mismatch (raw.rs:31-32, ffi.rs:66-71). */
uint32_t lsh_abi(void);
/* Backend probe (gpu.rs:307-310 available): NULL when a gfx950 device is usable. */
char *lsh_available(void);
/* Release a string returned by the library (ffi.rs:143-152 lapstack_raw_free_str). */
void lsh_free_str(char *s);
/* The library's HIP stream on the current device (a hipStream_t), for callers
 * that order their own work against it. */
char *lsh_stream(void **stream_out);
/* Wait for all queued work of the library on the current device. */
char *lsh_sync(void);
/* The library's device memory on the current device: bytes in blocks handed
 * out, and bytes in released blocks it keeps for reuse (at most an eighth of
 * the device's memory, at least 4 GiB). */
char *lsh_mem_stats(uint64_t *live_bytes, uint64_t *cached_bytes);
/* Return every kept block of the current device to it (waits for the work
 * that last used them), for callers that need the memory elsewhere. */
char *lsh_mem_release(void);

/* Per-kernel timing with HIP events on the library's stream: enable, read
 * (names separated by '\n' into names; per kernel the total ms, the total
 * algorithmic HBM bytes -- each array a launch must read or write, once, in
 * its stored format -- and the launch count), reset.  Off by default; timing
 * adds an event pair per launch. */
char *lsh_prof_enable(uint32_t on);
char *lsh_prof_read(char *names, size_t names_cap, double *total_ms, double *total_bytes, uint64_t *counts,
                    uint32_t cap, uint32_t *n);
char *lsh_prof_reset(void);

/* ---- frames: registration, warp, brightness, focus slice ---------------- */

/* gpu.rs:636 GpuFrames::new.  interp: the final resampling kernel (LSH_*). */
char *lsh_frames_new(uint32_t w, uint32_t h, uint32_t interp, lsh_frames **out);

/* gpu.rs:751-886 GpuFrames::process, with the CPU path's semantics
 * (stack.rs:513-530 register, :638-706 AlignedFrames::get):
 *   rgb           one frame, interleaved RGB u16 (h*w*3) in rgb_mem;
 *   guess         the previous frame's Sim (8 doubles: xoff, yoff, scale, rot,
 *                 aspect, shear, px, py; align.rs:199-253);
 *   free8         which Sim parameters the search moves (align.rs:115-119);
 *   coarsen       stop the search this many levels short of full resolution;
 *   known_sim, known_gains  (both NULL, or both set) a re-pass's transform and gains;
 *   brightness    bring the frame to frame 0's brightness (brightness.rs:57-61);
 *   measure       when non-NULL, the focus slice (depth.rs:1097) is written to
 *                 slice_out (dw*dh floats, dw = ceil(w/2^scale)) in slice_mem.
 * The first frame processed sits at the identity and becomes the brightness
 * reference.  Afterward the handle holds the frame (lsh_fuser_push_frames). */
char *lsh_frames_process(lsh_frames *f, const uint16_t *rgb, lsh_mem rgb_mem, const double *guess,
                         const uint8_t *free8, uint32_t coarsen, const double *known_sim, const float *known_gains,
                         uint32_t brightness, const lsh_depth_params *measure, double *sim_out, float *gains_out,
                         float *slice_out, lsh_mem slice_mem);

/* stack.rs:361-419 LazyFrames::take for registered stacks (no alignment): the
 * frame as decoded (io.rs:89), with brightness gains to frame 0's channel means
 * (brightness.rs:107-117) when asked, and its focus slice when measure != NULL. */
char *lsh_frames_ingest(lsh_frames *f, const uint16_t *rgb, lsh_mem rgb_mem, uint32_t brightness,
                        const lsh_depth_params *measure, float *gains_out, float *slice_out, lsh_mem slice_mem);

/* gpu.rs:889-900 planes/download: the held frame's device planes, or a copy. */
char *lsh_frames_planes(lsh_frames *f, const float **planes_dev);
char *lsh_frames_download(lsh_frames *f, float *img3, lsh_mem mem);
/* Cost evaluations the last lsh_frames_process search ran (the reference logs
 * the same count per frame pair, docs/engine.md:73-75). */
uint64_t lsh_frames_last_evals(lsh_frames *f);
/* The last search per registration level (level 0 = full resolution): cost
 * evaluations, and the device wall-clock ticks (hipDeviceAttributeWallClockRate)
 * the search spent there; up to cap levels.  With cap above 16: evals[16 + l]
 * the collective passes at level l (each pass evaluates up to 9 points), and
 * ticks[16..20) the ticks per phase (matrix, pixels, combine, steering), and
 * ticks[20 + 8 l + q] those of phase q at level l (q 4-6: the combine's
 * publishing, waiting and folding; q 7: the steering's branch, before an
 * iteration's set-up; cap 148 for all); the phases
 * are timed only while lsh_prof_enable is on or LSH_SEARCH_PHASES=1 is set in
 * the environment (zero otherwise: the clock reads would lengthen every pass). */
char *lsh_frames_search_stats(lsh_frames *f, uint64_t *evals, uint64_t *ticks, uint32_t cap);
/* No counterpart in the reference (whose search is a host loop it can log per
 * evaluation, align.rs:796-883): trace the costs the following searches count
 * -- every cost the reference's nelder_mead evaluates, in its order, level by
 * level coarse to fine, up to cap of them (0 turns tracing off) -- and read the
 * last search's (*n the count; up to cap into costs). */
char *lsh_frames_search_trace(lsh_frames *f, uint32_t cap);
char *lsh_frames_search_costs(lsh_frames *f, double *costs, uint32_t cap, uint32_t *n);
/* The blocks that took part in the last lsh_frames_process search (*n) and,
 * per block, a digest of every partial-sum total it folded: the blocks steer
 * one search each from the costs they fold, so all digests are equal. */
char *lsh_frames_search_blocks(lsh_frames *f, uint64_t *digests, uint32_t cap, uint32_t *n);
void lsh_frames_free(lsh_frames *f);

/* ---- fuser: Laplacian pyramid, region-energy fold, residual rule, collapse */

/* fuse.rs:371 Fuser::new / gpu.rs GpuFuser::new. */
char *lsh_fuser_new(uint32_t w, uint32_t h, const lsh_fuse_params *p, lsh_fuser **out);
/* fuse.rs:351 Fuser::levels. */
char *lsh_fuser_levels(lsh_fuser *f, uint32_t *levels, uint32_t *depth_level, uint32_t *dw, uint32_t *dh);
/* fuse.rs:401 Fuser::push: one aligned frame (3 planes). */
char *lsh_fuser_push(lsh_fuser *f, const float *img3, lsh_mem mem);
/* gpu.rs:451 GpuFuser::push_device: the frame a frames handle holds. */
char *lsh_fuser_push_frames(lsh_fuser *f, lsh_frames *src);
/* fuse.rs:489 Fuser::shares: up to cap values (n = frames folded, or 0). */
char *lsh_fuser_shares(lsh_fuser *f, float *out, uint32_t cap, uint32_t *n);
/* fuse.rs:499 Fuser::finish: fused image (clamped to [0,1]) into img3 (img_mem);
 * the winner map nearest-upsampled to w*h f32 (nullable, win_mem); the winner
 * map at depth_level, dw*dh u16 in host memory (nullable).  The fuser can be
 * finished once. */
char *lsh_fuser_finish(lsh_fuser *f, float *img3, lsh_mem img_mem, float *winner_full, lsh_mem win_mem,
                       uint16_t *winner_level);
void lsh_fuser_free(lsh_fuser *f);

/* ---- a whole stack, streamed ---------------------------------------------- */

/* stack.rs:740-826 fuse_range over n frames (frames[i]: interleaved RGB u16 in
 * frames_mem) into a fresh frames handle and a fresh fuser, with no host round
 * trip per frame after the first: frame 0 takes lsh_frames_process (identity,
 * brightness anchor), every later frame registers against the previous aligned
 * frame from its Sim (the reference's guess), and its map and brightness gains
 * are derived on the device, then warp, focus slice (measure non-NULL: n slices
 * of dw*dh into slices, slices_mem) and the fold.  graph != 0: the per-frame step
 * is captured once as a hipGraph and replayed for every frame (the frame staged
 * into a fixed buffer first); graph == 0: the same kernels launched per frame.
 * Per frame (nullable): the Sim (n*8), the gains (n*3), the search's
 * evaluations (n; 0 for frame 0).  Finish the fuser with lsh_fuser_finish. */
char *lsh_stack_stream(lsh_frames *f, lsh_fuser *fu, const uint16_t *const *frames, uint32_t n, lsh_mem frames_mem,
                       const uint8_t *free8, uint32_t coarsen, uint32_t brightness, const lsh_depth_params *measure,
                       float *slices, lsh_mem slices_mem, uint32_t graph, double *sims_out, float *gains_out,
                       uint64_t *evals_out);

/* stack.rs:277-360 with io.rs:62-97: lsh_stack_stream over n image files
 * (paths[i], 16- or 8-bit PNG or TIFF, each the handles' size), decoded by
 * `threads` host threads (0: 8) ahead of their use into a ring of pinned host
 * buffers and uploaded from there on the library's copy stream, beside the
 * device's work on earlier frames.  Same outputs as lsh_stack_stream; a file
 * that cannot be decoded, or of another size, ends the call with its error. */
char *lsh_stack_stream_files(lsh_frames *f, lsh_fuser *fu, const char *const *paths, uint32_t n, uint32_t threads,
                             const uint8_t *free8, uint32_t coarsen, uint32_t brightness,
                             const lsh_depth_params *measure, float *slices, lsh_mem slices_mem, uint32_t graph,
                             double *sims_out, float *gains_out, uint64_t *evals_out);

/* ---- image files (io.rs) -------------------------------------------------- */

/* io.rs:62-97 load_rgb for PNG (colour types 0/2/4/6, 8 or 16 bits, not
 * interlaced) and baseline TIFF (strips, chunky or planar, uncompressed, LZW or
 * Deflate, predictor 1 or 2): the frame as interleaved RGB u16 (h*w*3, into rgb
 * with room for cap samples; NULL: size only), 8-bit samples v given as v*257
 * (the same fraction); *bits the file's bits per sample.  lsh_image_info: the
 * size and bits only. */
char *lsh_image_info(const char *path, uint32_t *w, uint32_t *h, uint32_t *bits);
char *lsh_image_load(const char *path, uint16_t *rgb, size_t cap, uint32_t *w, uint32_t *h, uint32_t *bits);
/* io.rs:273-315 save_rgb: three planes (3*w*h f32 in mem) as a 16- or 8-bit
 * TIFF (uncompressed, strips of about 8 MB) or PNG, by the path's extension;
 * samples (clamp(v, 0, 1) * 65535 + 0.5) or (... * 255 + 0.5), truncated.
 * lsh_image_save_u16: interleaved RGB u16 samples (lsh_to_u16's) as they are. */
char *lsh_image_save(const char *path, const float *img3, lsh_mem mem, uint32_t w, uint32_t h, uint32_t bits);
char *lsh_image_save_u16(const char *path, const uint16_t *rgb, lsh_mem mem, uint32_t w, uint32_t h);
/* io.rs:325-335 save_gray16: a plane in [0, scale] as a 16-bit gray PNG,
 * (clamp(v, 0, scale) * 65535 / scale + 0.5) truncated (depth maps: scale =
 * frames - 1; confidence: 1).  io.rs:337-346 save_gray: min-max normalised
 * 8-bit gray PNG. */
char *lsh_save_gray16(const char *path, const float *plane, lsh_mem mem, uint32_t w, uint32_t h, float scale);
char *lsh_save_gray(const char *path, const float *plane, lsh_mem mem, uint32_t w, uint32_t h);

/* ---- many stacks in one launch -------------------------------------------- */

/* One stack of a batch: n frames of w*h (interleaved RGB u16, device memory). */
typedef struct lsh_batch_stack {
    uint32_t w, h, n;
    const uint16_t *const *frames; /* host array of n device pointers */
} lsh_batch_stack;

/* How a batch is scheduled.  tune: 0 the configuration given here (0 fields:
 * defaults 16 rows, as many blocks per CU as fit, 2 frames per round); 1 the
 * autotuner, its choice cached per shape class (depth, energy window, log2 of
 * the pixel-frames, log2 of the stack count); 2 the autotuner again. */
typedef struct lsh_batch_options {
    uint32_t tune;
    uint32_t tile_h;           /* fold tile rows: 8 or 16 (0: default) */
    uint32_t blocks_per_cu;    /* resident blocks per CU (0: as many as fit) */
    uint32_t frames_per_round; /* frames of each stack per round, 1..4 (0: default) */
} lsh_batch_options;

typedef struct lsh_batch_report {
    double ms;                 /* device time of the batch launch */
    double mpix_frames;        /* sum over stacks of w*h*n / 1e6 */
    uint32_t rounds;           /* ceil(largest frame count / frames_per_round) */
    uint32_t max_levels;
    uint32_t tile_h;           /* the configuration run */
    uint32_t blocks_per_cu;    /* launched per CU (the grid is blocks_per_cu * CUs) */
    uint32_t resident_blocks;  /* of the grid, the blocks resident at the start, which do the work */
    uint32_t frames_per_round;
    uint32_t from_cache;       /* the configuration came from the shape-class cache */
    uint32_t ntried;           /* autotuner candidates timed (round 0 each) */
    uint32_t tried_cfg[8];     /* tile_h | blocks_per_cu << 8 | frames_per_round << 16 */
    float tried_us_per_mpf[8]; /* their device time per megapixel-frame */
} lsh_batch_report;

/* No counterpart in the reference, whose engine runs one stack at a time
 * (stack.rs:740-826 fuse_range): count stacks of uneven sizes and frame
 * counts through the per-frame fuse path (the frame as decoded, without
 * brightness matching; the REDUCE chain; the fold of every band-pass level; the
 * residual kept) in one persistent launch.  fusers[i] (fresh, from
 * lsh_fuser_new with the stack's size and levels; one energy window and chroma
 * setting for all, halo control off) receive stacks[i]; each is then finished
 * with lsh_fuser_finish as if its frames had been pushed one by one, with the
 * same result.  opt: nullable (autotuner with its cache); rep: nullable. */
char *lsh_batch_fuse(lsh_fuser **fusers, const lsh_batch_stack *stacks, uint32_t count, const lsh_batch_options *opt,
                     lsh_batch_report *rep);

/* ---- depth ------------------------------------------------------------ */

/* depth.rs:1119-1200 depth_from_slices: n slices of dw*dh (contiguous, frame
 * order) in slices_mem; the fused image (3*w*h) in fused_mem; depth and conf
 * (w*h) into out_mem; the noise-floor plane (dw*dh, nullable) into out_mem. */
char *lsh_depth_from_slices(const float *slices, lsh_mem slices_mem, uint32_t n, const float *fused3,
                            lsh_mem fused_mem, uint32_t w, uint32_t h, const lsh_depth_params *p, float *depth,
                            float *conf, lsh_mem out_mem, float *floor_plane, lsh_depth_info *info);

/* depth.rs:1120 the slice iterator (Iterator<Item = Result<Vec<f32>, String>>)
 * as a pull callback, called for m = 0, 1, ... n-1 in order.  It either sets
 * *slice to slice m (*cells floats in *mem) and returns NULL; or leaves *slice
 * NULL and returns NULL to say the source has ended (the pass then fails with
 * "depth: slice m of n missing", depth.rs:1138); or returns an error message,
 * which becomes the pass's error unchanged (the iterator's Err; the library
 * copies it as the callback returns).  A slice has been read before the next
 * call, so one buffer may serve every slice. */
typedef const char *(*lsh_slice_fn)(void *user, uint32_t m, const float **slice, uint64_t *cells, lsh_mem *mem);

/* depth.rs:1124 log: one progress line per call, NUL-terminated, valid during the call */
typedef void (*lsh_log_fn)(void *user, const char *line);

/* depth.rs:1119-1200 depth_from_slices with its slice iterator and log: as
 * lsh_depth_from_slices, the slices pulled one at a time from next (a slice of
 * the wrong cell count fails with "depth: slice m has c cells, the working grid
 * dwxdh", depth.rs:1139-1141), and the reference's progress lines (guide and
 * grid, slices aggregated, peaks and confidence, robust reweighting, WLS,
 * upsampling) passed to log when it is not NULL. */
char *lsh_depth_from_source(lsh_slice_fn next, void *next_user, uint32_t n, const float *fused3, lsh_mem fused_mem,
                            uint32_t w, uint32_t h, const lsh_depth_params *p, float *depth, float *conf,
                            lsh_mem out_mem, float *floor_plane, lsh_depth_info *info, lsh_log_fn log, void *log_user);

/* depth.rs:1125-1159, the depth pass up to the WLS solve: the working grid's
 * guide luma, the peak depth (after the median when p->median) and the
 * p90-normalised confidence, each dw*dh into out_mem; info gets dw, dh, p90
 * and floor.  These are the WLS system's inputs (lsh_wls_solve). */
char *lsh_depth_grid(const float *slices, lsh_mem slices_mem, uint32_t n, const float *fused3, lsh_mem fused_mem,
                     uint32_t w, uint32_t h, const lsh_depth_params *p, float *guide, float *depth, float *conf,
                     lsh_mem out_mem, lsh_depth_info *info);

/* ---- the weighted average, slabs -------------------------------------- */

/* wav.rs:60-80 WavParams: the contrast above the floor raised to power; the
 * box radius (grid cells) the contrast and then the weights are smoothed by,
 * 0 none; the margin over the noise floor; the weights' smoothing follows the
 * fused image's edges (the guided filter) when edge != 0 and a guide is given. */
typedef struct lsh_wav_params {
    float power;
    uint32_t smooth;
    float gate;
    uint32_t edge;
} lsh_wav_params;

typedef struct lsh_wav lsh_wav;

/* wav.rs:97-161 weighted_average as an accumulator over the aligned frames:
 * w x h frames, the depth pass's measure and grid (dp), floor the depth pass's
 * per-cell noise floor (dw*dh, lsh_depth_from_slices' floor_plane) in
 * floor_mem, fused3 the fused image (3*w*h, the guide; nullable) in fused_mem.
 * Push every frame (3*w*h planes, aligned and equalized) with its focus slice
 * (dw*dh, as the depth pass took it; nullable: measured here), or the frame a
 * frames handle holds; finish writes the average (3*w*h) into mem. */
char *lsh_wav_new(uint32_t w, uint32_t h, const lsh_depth_params *dp, const lsh_wav_params *wp, const float *floor,
                  lsh_mem floor_mem, const float *fused3, lsh_mem fused_mem, lsh_wav **out);
char *lsh_wav_push(lsh_wav *a, const float *img3, lsh_mem img_mem, const float *slice, lsh_mem slice_mem);
char *lsh_wav_push_frames(lsh_wav *a, lsh_frames *src);
char *lsh_wav_finish(lsh_wav *a, float *img3, lsh_mem mem);
void lsh_wav_free(lsh_wav *a);

/* stack.rs:76-92 slab_ranges: the [lo, hi] frame ranges of slabs of size frames
 * overlapping by overlap; returns their count, writes up to cap pairs into lohi
 * (nullable).  Each slab is fused on its own (a fuser per slab, stack.rs:873-893). */
uint32_t lsh_slab_ranges(uint32_t count, uint32_t size, uint32_t overlap, uint32_t *lohi, uint32_t cap);

/* align.rs:741-754 gauss_levels: the registration pyramid's levels above the
 * plane y (w x h) -- each half the size (rounded up) of the one below while
 * h > 64 and w > 8, by the border-renormalised Burt REDUCE (align.rs:650-734)
 * -- packed level after level into packed (out_mem); *nlevels gets their
 * count (packed NULL: the count only).  The frames handle builds the same
 * pyramid for every frame's search. */
char *lsh_gauss_levels(const float *y, lsh_mem y_mem, uint32_t w, uint32_t h, float *packed, lsh_mem out_mem,
                       uint32_t *nlevels);

/* ---- the output side --------------------------------------------------- */

/* align.rs:344-403 common_area: the largest pixel-aligned rectangle (x, y, w,
 * h into rect4) that every frame covers with its kernel's support inside the
 * frame, from the n Sims (n*8) of a w x h stack; the whole frame when nothing
 * is cut or nothing is common.  Host arithmetic only. */
char *lsh_common_area(const double *sims, uint32_t n, uint32_t w, uint32_t h, uint32_t interp, uint32_t *rect4);
/* Img3::crop: the rect4 (x, y, w, h) of each of planes w x h planes, packed. */
char *lsh_crop(const float *src, lsh_mem src_mem, uint32_t planes, uint32_t w, uint32_t h, const uint32_t *rect4,
               float *dst, lsh_mem dst_mem);
/* prep.rs:46-71 resize / resize_plane (the restretch of stack.rs:951-974):
 * planes pw x ph planes resampled to w x h by the warp with the Sim that maps
 * pixel centers onto pixel centers, in interpolation interp. */
char *lsh_resize(const float *src, lsh_mem src_mem, uint32_t planes, uint32_t pw, uint32_t ph, uint32_t w, uint32_t h,
                 uint32_t interp, float *dst, lsh_mem dst_mem);
/* io.rs:283-288: 3 planes w x h to interleaved RGB u16, (clamp(v, 0, 1) * 65535 + 0.5) truncated. */
char *lsh_to_u16(const float *img3, lsh_mem img_mem, uint32_t w, uint32_t h, uint16_t *rgb, lsh_mem rgb_mem);

/* ---- the WLS solve alone, in three precisions ------------------------- */

typedef enum lsh_wls_mode {
    LSH_WLS_REFERENCE = 0, /* depth.rs:1006-1077: FP32 MG-PCG, the CG recurrence's residual */
    LSH_WLS_REFINED = 1,   /* FP32 MG-PCG inside FP64 iterative refinement (FP64 iterate and residual);
                              exact FP64 coarsest solve */
    LSH_WLS_FP64 = 2       /* the FP64 oracle: FP64 vectors and V-cycle over the same operator */
} lsh_wls_mode;

typedef struct lsh_wls_report {
    uint32_t iters;  /* CG iterations (summed over the refinement's inner solves) */
    uint32_t outer;  /* refinement steps */
    double rel;      /* the solver's final relative residual: the CG recurrence's (REFERENCE), the
                        true one of the FP64 iterate (REFINED), the FP64 recurrence's (FP64) */
    double true_rel; /* ||W d - A u|| / ||W d|| of the returned FP32 u, in FP64 */
    double err_fp64; /* max |u - u*| against an FP64 solve to 1e-12 (want_err), else -1 */
    double ms;       /* device time of the solve: the FGS guess (or warm copy) and the PCG */
} lsh_wls_report;

/* depth.rs:1081 wls_solve (the one-call WlsSolver: new, set_weights, solve),
 * with the tolerance, the precision mode and an optional warm start (w*h, in
 * in_mem; NULL for the reference's FGS guess).  guide, conf, d in in_mem; u
 * (w*h f32) and u64 (w*h f64, nullable; the FP64 iterate of the REFINED and
 * FP64 modes) in out_mem.  lambda
 * must be positive (lambda = 0 is the identity the caller skips, depth.rs:1161). */
char *lsh_wls_solve(const float *guide, const float *conf, const float *d, const float *warm, lsh_mem in_mem,
                    uint32_t w, uint32_t h, float lambda, float sigma_c, uint32_t mode, double tol,
                    uint32_t max_iters, uint32_t want_err, float *u, double *u64, lsh_mem out_mem,
                    lsh_wls_report *rep);

/* No counterpart in the reference; for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): lsh_wls_solve in LSH_WLS_REFERENCE mode with every buffer
 * of the solver (the multigrid hierarchy, the CG vectors, the reduction
 * partials) in one device array the library holds per GPU, of
 * LSH_WLS_NOALLOC_BYTES (20 MiB: a 512 x 384 grid and larger ones up to about
 * 230 000 cells; a grid that needs more is refused with an error string).  It
 * calls no hipMalloc, hipFree or other device or pinned allocation.  guide,
 * conf, d, warm (nullable: the FGS guess) and u are device memory.  Every call
 * on one GPU shares that GPU's array and overwrites it: calls that overlap on
 * one GPU (other threads, other streams) race. */
#define LSH_WLS_NOALLOC_BYTES (20u << 20)
char *lsh_wls_solve_noalloc(const float *guide, const float *conf, const float *d, const float *warm, uint32_t w,
                            uint32_t h, float lambda, float sigma_c, double tol, uint32_t max_iters, float *u,
                            lsh_wls_report *rep);

/* ---- the learned warm start of the WLS solve ---------------------------- */

/* No counterpart in the reference, whose PCG starts from its FGS sweeps
 * (depth.rs:1019-1034): a bf16 convolutional encoder-decoder (channels
 * 16/32/64/64, 3x3 convolutions, ReLU, 255,489 parameters) on the grid of 8 x 8
 * blocks of the WLS grid.  From the system's planes (peak depth, p90-normalised
 * confidence, guide luma; lsh_depth_grid) averaged over the blocks -- the plain
 * mean depth and the weighted means over 4 x 4, 16 x 16 and 64 x 64 blocks of
 * blocks, each as its difference from the block's mean depth weighted by the
 * data weight max(conf, 0) + 1e-4, that weight's mean and log, the mean edge
 * conductances exp(-|dg| / sigma_c) -- it predicts the solution's block means
 * as that weighted mean plus a correction; their bilinear prolongation is the
 * start for lsh_wls_solve, in place of the FGS sweeps.  Inference and training run on the device (bf16 MFMA, f32
 * accumulation, f32 master weights, Adam).  Depth enters divided by dscale (the
 * frame-index range, n - 1) and leaves multiplied by it; the WLS solution
 * scales the same way.
 *
 * Parameters travel as one flat f32 array, layer by layer in network order
 * (e0a, e0b, e1a, e1b, e2a, e2b, e3a, e3b, d2, d1, d0, out): the weights
 * W[out][in][ky][kx], then the biases. */
char *lsh_warm_new(uint64_t seed, lsh_warm **out);
void lsh_warm_free(lsh_warm *net);
char *lsh_warm_nparams(lsh_warm *net, uint32_t *count);
char *lsh_warm_get(lsh_warm *net, float *params, uint32_t count);
/* Load parameters (resets the optimiser state). */
char *lsh_warm_set(lsh_warm *net, const float *params, uint32_t count);
/* The guide scale sigma_c of the WLS systems served (default 0.04): the edge
 * conductances among the inputs are exp(-|dg| / sigma_c). */
char *lsh_warm_set_sigma(lsh_warm *net, float sigma_c);
/* nimg systems of w*h planes each (image-major) in in_mem; the start u
 * (nimg*w*h) in out_mem. */
char *lsh_warm_predict(lsh_warm *net, const float *depth, const float *conf, const float *guide, lsh_mem in_mem,
                       uint32_t nimg, uint32_t w, uint32_t h, float dscale, float *u, lsh_mem out_mem);
/* The batch's loss against target (the WLS solution, nimg*w*h): per system
 * the mean squared error of the predicted block means against the target's
 * (depth / dscale, in units of 1/32), divided by the mean square of the
 * target's own deviation from the weighted mean depth (at least 1e-4), then
 * averaged over the systems; and its gradient (host, count floats, nullable),
 * from the device's backward pass. */
char *lsh_warm_grads(lsh_warm *net, const float *depth, const float *conf, const float *guide, const float *target,
                     lsh_mem in_mem, uint32_t nimg, uint32_t w, uint32_t h, float dscale, double *loss, float *grads,
                     uint32_t count);
/* One Adam step (beta 0.9 / 0.999) on the batch; loss before the step.  A
 * non-finite loss is reported and leaves the weights as they were. */
char *lsh_warm_train_step(lsh_warm *net, const float *depth, const float *conf, const float *guide,
                          const float *target, lsh_mem in_mem, uint32_t nimg, uint32_t w, uint32_t h, float dscale,
                          float lr, double *loss);

#ifdef __cplusplus
}
#endif
#endif
