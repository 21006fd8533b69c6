// stream.hip -- a whole stack through the frames handle and the fuser with no
// host round trip per frame (stack.rs:740-826 fuse_range, the per-frame
// loop).  Frame 0 takes the ordinary path (it is the reference and the
// brightness anchor); every later frame runs the device-only step
// (registration, map, gains, warp, focus slice), the fold into the fuser and a
// per-frame record.  With graph on, that step is captured once as a hipGraph
// and replayed per frame: the frame is staged into a fixed buffer before the
// replay, its focus slice leaves a fixed buffer after it, and the frame index
// the fold and the record need lives in device memory.
#include <vector>

#include "api.hpp"
#include "common.hpp"
#include "engine.hpp"

using namespace lsh;

extern "C" {

char *lsh_stack_stream(lsh_frames *f, lsh_fuser *fu, const uint16_t *const *frames, uint32_t n, lsh_mem frames_mem,
                       const uint8_t *free8, uint32_t coarsen, uint32_t brightness, const lsh_depth_params *measure,
                       float *slices, lsh_mem slices_mem, uint32_t graph, double *sims_out, float *gains_out,
                       uint64_t *evals_out) {
    LSH_API_BEGIN
    LSH_REQUIRE(f && fu && frames && free8, "lsh_stack_stream: null argument");
    LSH_REQUIRE(n > 0, "stream: no frames");
    LSH_REQUIRE(f->processed == 0 && fu->count == 0, "stream: the handles must be fresh");
    LSH_REQUIRE(f->w == fu->w && f->h == fu->h, "stream: frames and fuser sizes differ");
    LSH_REQUIRE(!measure || slices, "stream: measure needs slices");
    const size_t w = f->w, h = f->h, npx = w * h;
    size_t M = 0;
    if (measure) {
        LSH_REQUIRE(measure->scale <= 2, "focus slice: scale above 2 is not supported");
        const size_t k = (size_t)1 << measure->scale;
        M = ((w + k - 1) / k) * ((h + k - 1) / k);
    }
    // frame 0: the ordinary path
    double sim0[8];
    float g0[3];
    static const double ident[8] = {0, 0, 1, 0, 1, 0, 0, 0};
    char *e = lsh_frames_process(f, frames[0], frames_mem, ident, free8, coarsen, nullptr, nullptr, brightness, measure,
                                 sim0, g0, measure ? slices : nullptr, slices_mem);
    if (e) return e;
    e = lsh_fuser_push_frames(fu, f);
    if (e) return e;
    std::vector<double> sims((size_t)n * 8);
    std::vector<float> gains((size_t)n * 3);
    std::vector<uint64_t> evals(n, 0);
    std::copy(sim0, sim0 + 8, sims.begin());
    std::copy(g0, g0 + 3, gains.begin());
    if (n > 1) {
        frames_stream_setup(f, free8, coarsen, n);
        fuser_reserve(fu, n);
        f->staging.ensure(3 * npx);
        DBuf<float> sbuf(measure ? M : 0);
        hipGraph_t g = nullptr;
        hipGraphExec_t ge = nullptr;
        auto step = [&](bool plain) {
            frames_device_step(f, f->staging.p, brightness != 0, measure, sbuf.p, plain);
            fuser_push_dev(fu, f->planes.p, f->cursor.p);
            frames_record(f);
        };
        try {
            for (uint32_t i = 1; i < n; i++) {
                copy_mem(f->staging.p, LSH_DEVICE, frames[i], frames_mem, 3 * npx * sizeof(uint16_t));
                if (graph) {
                    if (!ge) {
                        prof().capturing = true;
                        LSH_CHECK(hipStreamBeginCapture(stream(), hipStreamCaptureModeThreadLocal));
                        try {
                            step(true);
                        } catch (...) {
                            hipGraph_t junk = nullptr;
                            (void)hipStreamEndCapture(stream(), &junk);
                            if (junk) (void)hipGraphDestroy(junk);
                            prof().capturing = false;
                            throw;
                        }
                        LSH_CHECK(hipStreamEndCapture(stream(), &g));
                        prof().capturing = false;
                        LSH_CHECK(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
                    } else {
                        // the host-side counts the captured step advanced once
                        f->processed++;
                        fu->count++;
                    }
                    LSH_CHECK(hipGraphLaunch(ge, stream()));
                } else {
                    step(false);
                }
                if (measure)
                    copy_mem(slices + (size_t)i * M, slices_mem, sbuf.p, LSH_DEVICE, M * sizeof(float));
            }
        } catch (...) {
            if (ge) (void)hipGraphExecDestroy(ge);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        sync();
        if (ge) LSH_CHECK(hipGraphExecDestroy(ge));
        if (g) LSH_CHECK(hipGraphDestroy(g));
        std::vector<double> ds((size_t)n * 8);
        std::vector<float> dg((size_t)n * 3);
        frames_stream_results(f, n, ds.data(), dg.data(), evals.data());
        evals[0] = 0;  // frame 0 is not searched
        std::copy(ds.begin() + 8, ds.end(), sims.begin() + 8);
        std::copy(dg.begin() + 3, dg.end(), gains.begin() + 3);
    }
    if (sims_out) std::copy(sims.begin(), sims.end(), sims_out);
    if (gains_out) std::copy(gains.begin(), gains.end(), gains_out);
    if (evals_out) std::copy(evals.begin(), evals.end(), evals_out);
    LSH_API_END
}

}  // extern "C"
