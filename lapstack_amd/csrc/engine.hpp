// engine.hpp -- the device state behind the C ABI's opaque handles.
#pragma once
#include "common.hpp"

// A raw frame (interleaved RGB u16) as a kernel finds it: the pointer itself,
// or, inside a replayed graph, entry cur[0] + off of a device table of frame
// pointers (the table and the counter stay put while the frame changes).
struct FrameRef {
    const uint16_t *p = nullptr;
    const uint16_t *const *tab = nullptr;
    const unsigned *cur = nullptr;
    int off = 0;
    FrameRef(const uint16_t *q = nullptr) : p(q) {}
    int mod = 0;  // > 0: the entry is (cur[0] + off) % mod (a ring of staging buffers)
    FrameRef(const uint16_t *const *t, const unsigned *c, int o, int m = 0) : tab(t), cur(c), off(o), mod(m) {}
    __device__ __forceinline__ const uint16_t *get() const {
        if (!tab) return p;
        const unsigned i = *cur + off;
        return tab[mod ? i % mod : i];
    }
};

// The frames handle (the reference's GpuFrames, gpu.rs:636-900): the current
// frame's planes, its registration luma and pyramid, the previous aligned
// frame's luma pyramid (the next frame's search reference), and frame 0's
// brightness samples.
struct lsh_frames {
    size_t w = 0, h = 0;
    uint32_t interp = LSH_SPLINE4X4;
    lsh::DBuf<uint16_t> staging;  // a host frame uploaded here
    lsh::DBuf<float> planes;      // 3*w*h: the frame after warp and gains
    lsh::DBuf<float> yt;          // target luma pyramid: level 0 (full res) then the registration levels
    lsh::DBuf<float> yr;          // reference (previous aligned frame's luma) pyramid, same layout
    lsh::DBuf<float> yts[2];      // a pipelined stream's target pyramids (frame k in slot k % 2)
    lsh::DBuf<float> planes_b;    // a pipelined stream's second planes buffer
    std::vector<size_t> lw, lh, loff;  // registration level dims and offsets (level 0 = full res)
    bool have_ref = false;        // yr holds a reference
    bool have_grid = false;       // frame 0's brightness grid is set
    lsh::DBuf<float> grid;        // 3*gw*gh frame-0 samples on the step-4 grid
    size_t gw = 0, gh = 0;
    double ref_means[3] = {0, 0, 0};  // frame 0's channel means (registered stacks)
    bool have_means = false;
    lsh::DBuf<double> red;        // reduction partials
    lsh::DBuf<double> red_out;    // final sums
    lsh::DBuf<unsigned> ticket;   // last-block counter
    lsh::DBuf<int32_t> taps_dy, taps_dx;
    lsh::DBuf<float> taps_wt;
    uint32_t ntaps = 0, taps_key = 0xffffffffu;
    size_t processed = 0;
    uint64_t evals = 0;           // cost evaluations of the last search
    lsh::DBuf<uint8_t> nm_const;  // the search's level table
    lsh::DBuf<double> nm_state;   // guess in / Sim out, evaluation counts
    lsh::DBuf<unsigned> nm_flags; // grid-barrier counter, error flag
    lsh::DBuf<double> nm_part;    // per-block partial sums, two slots
    unsigned nm_grid = 0;
    unsigned nm_bpc = 1;  // the search's blocks per CU
    lsh::DBuf<unsigned long long> nm_digest;  // per participant of the last search: its digest of the folded sums
    lsh::DBuf<double> nm_trace;               // the counted costs of the last search (tracing on)
    size_t trace_cap = 0, trace_n = 0;        // tracing's capacity (0: off); costs the last search counted
    unsigned search_part = 0;                 // participants of the last host-steered search
    uint64_t lev_evals[16] = {0}, lev_ticks[16] = {0};  // the last search, per registration level
    uint64_t lev_passes[16] = {0};                       // its collective passes (several evaluations each)
    std::vector<uint8_t> nm_host;                        // the search constants last uploaded
    lsh::DBuf<uint8_t> fdev;                             // the streaming path's per-frame map and gains
    lsh::DBuf<unsigned> cursor;                          // [0] the frame index, [1] a sticky search error
    lsh::DBuf<double> rec_sims, rec_evals;               // per-frame Sim and evaluations of a stream
    lsh::DBuf<float> rec_gains;                          // per-frame gains of a stream
    uint64_t phase_ticks[4] = {0};
    uint64_t lev_phase_ticks[16][8] = {};
    size_t coarsen = 0;                                  // a stream's finest searched level
};

struct lsh_fuser {
    size_t w = 0, h = 0, levels = 0, depth_level = 0, count = 0;
    lsh_fuse_params p{};
    bool halo_on = false, finished = false;
    size_t guide = 0;
    float hard = 0.f;
    std::vector<size_t> lw, lh, off;  // level dims, per-plane offsets into a pyramid of S cells
    size_t S = 0;
    lsh::DBuf<float> G;        // 3*S Gaussian levels (level 0 used for host-pushed frames)
    lsh::DBuf<float> acc;      // 3*S fused band-pass levels (and the weighted residual with halo control)
    lsh::DBuf<float> best;     // S winning energies (weight sums with halo control)
    lsh::DBuf<uint16_t> winner;
    lsh::DBuf<float> tops;     // count * 3 * tsz residuals
    size_t tops_cap = 0;
    lsh::DBuf<float> en_guide, wgt_a, wgt_b, lap;  // halo control scratch
    lsh::DBuf<float> ent_tab;  // p*ln(p) per window count (residual rule)
    lsh::DBuf<unsigned> cnt;   // shares
};

namespace lsh {
// The device-only per-frame path of a stack stream (align.hip): the search
// constants for the stack, then per frame the registration, map, gains, warp
// and measurement without a host round trip, and the per-frame record.
void frames_stream_setup(lsh_frames *f, const uint8_t *free8, uint32_t coarsen, size_t nframes);
// The parts of a pipelined stream, slot = the frame's parity: its target luma
// and pyramid into yts[slot]; its serial chain -- the searched levels of that
// slot into the search's buffer, the search from the previous Sim, the map
// (kept with the Sim in the slot's FrameDev), the warped registration luma
// (the next frame's reference) and its pyramid; and its side warp (below).
// The focus slice and the fold of the planes are left to the caller, who may
// run them and the side warp beside the next frame's chain.  started (if
// given) is recorded between the search's reset and the search itself: work
// another stream orders after it is dispatched after the search, so it cannot
// take the CUs the search's grid needs at its start.
void frames_stream_slots(lsh_frames *f);
// out_idx (if given): the slice goes to out + out_idx[0] * (its cell count);
// luma: planes is the frame's luma plane (one plane), not its three planes;
// out_base (if given): out is read from device memory there, so a captured
// graph's launch holds no pointer to the caller's buffer (the runtime keeps
// a buffer that a graph's kernel arguments named from going back to the device)
// the learned start for one w x h WLS system, on the library's stream (warm.hip)
void warm_predict_dev(lsh_warm *n, const float *d, const float *conf, const float *g, int w, int h, float dscale,
                      float sigma_c, float *u);
void focus_slice(lsh_frames *f, const float *planes, const lsh_depth_params &p, float *out,
                 const unsigned *out_idx = nullptr, bool luma = false, float *const *out_base = nullptr);  // depth.hip
void frames_target(lsh_frames *f, const FrameRef &src, int slot);
void frames_chain(lsh_frames *f, int slot, bool plain_search, hipEvent_t started = nullptr);
// The rest of frame k's step, off the chain: the brightness sums and gains,
// the RGB warp into planes_out (the map of the FrameDev at slot) and, when
// fy_out is given, the warped frame's luma into it; the record.
void frames_side_warp(lsh_frames *f, const FrameRef &src, int slot, float *planes_out, bool brightness,
                      float *fy_out = nullptr);
void frames_stream_results(lsh_frames *f, size_t n, double *sims, float *gains, uint64_t *evals);
// fuse.hip: a push whose frame index comes from device memory (a replayed graph)
void fuser_reserve(lsh_fuser *f, size_t nframes);
void fuser_push_dev(lsh_fuser *f, const float *g0, const unsigned *idx_dev);
}  // namespace lsh
