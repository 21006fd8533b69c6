// stream.hip -- a whole stack through the frames handle and the fuser with no
// host round trip per frame (stack.rs:740-826 fuse_range, the per-frame
// loop).  Frame 0 takes the ordinary path (it is the reference and the
// brightness anchor); every later frame k runs on the device alone:
//
//   the chain (main stream): the searched levels of frame k's target pyramid
//     into the search's buffer, the search from the previous Sim, the map
//     derived on the device, the warped registration luma (frame k+1's
//     reference) and its pyramid -- nothing else the next search waits for;
//   the side work: frame k's brightness gains, its RGB warp into planes[k % 2]
//     and its record, its focus slice and fold, and the target luma and
//     pyramid of frame k+1 into slot (k+1) % 2.
//
// Only the chain is serial across frames (frame k+1 registers against frame
// k's warped luma), so frame k's side work runs beside frame k+1's chain: the
// search is bound by the latency of its grid-wide reductions, and its grid
// here covers three quarters of the CUs, which leaves the rest to the side
// work.  Every value is the one the per-frame path computes; only the overlap
// differs.
//
// graph == 0: two streams and events.  graph != 0: frame k >= 2 (with a frame
// after it) is one replay of a captured hipGraph with two branches -- frame
// k's chain beside frame k-1's side work -- one graph per parity of k (the
// buffers alternate).  A replay finds its frames in a device table of the
// caller's frame pointers at a device counter the graph advances once per
// replay (host frames are staged into two fixed buffers instead), the fold
// takes its frame index from a device counter only the side work advances,
// and the focus slice writes the caller's device slices at that index (host
// slices leave through two fixed buffers).
#include <atomic>
#include <condition_variable>
#include <cstdlib>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "api.hpp"
#include "common.hpp"
#include "engine.hpp"
#include "imgio.hpp"

using namespace lsh;

namespace {

__global__ void k_counter_set(unsigned *c, unsigned v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) c[0] = v;
}
__global__ void k_counter_next(unsigned *c) {
    if (threadIdx.x == 0 && blockIdx.x == 0) c[0] = c[0] + 1;
}

struct Events {
    std::vector<hipEvent_t> ev;
    hipEvent_t make() {
        hipEvent_t e = nullptr;
        LSH_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        ev.push_back(e);
        return e;
    }
    ~Events() {
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    }
};

struct Graphs {
    hipGraph_t g[2] = {nullptr, nullptr};
    hipGraphExec_t x[2] = {nullptr, nullptr};
    ~Graphs() {
        for (int p = 0; p < 2; p++) {
            if (x[p]) (void)hipGraphExecDestroy(x[p]);
            if (g[p]) (void)hipGraphDestroy(g[p]);
        }
    }
};

// Frames decoded from files ahead of their use (stack.rs:277-360, the
// reference's read-ahead): worker threads decode frame j into pinned slot
// j % S once frame j - S has left it (its upload to the device, recorded as an
// event on the copy stream, has completed), so decoding, the upload and the
// device's work on earlier frames run at once.
struct FileRing {
    std::vector<std::string> paths;
    size_t w, h, nvals;
    int S = 0, dev = 0;
    std::vector<uint16_t *> buf;
    std::vector<hipEvent_t> ev;
    std::vector<long> ready, free_from;
    std::vector<char> pending;
    std::mutex mu;
    std::condition_variable cv;
    std::atomic<long> next{0};
    std::vector<std::thread> th;
    std::string err;
    bool stop = false;

    FileRing(const char *const *p, uint32_t n, size_t w_, size_t h_, unsigned threads) : w(w_), h(h_), nvals(w_ * h_ * 3) {
        for (uint32_t i = 0; i < n; i++) paths.emplace_back(p[i]);
        LSH_CHECK(hipGetDevice(&dev));
        threads = std::max(1u, std::min(threads, n));
        S = (int)std::min<uint32_t>(threads + 3, n);
        buf.assign(S, nullptr);
        ev.assign(S, nullptr);
        ready.assign(S, -1);
        pending.assign(S, 0);
        free_from.resize(S);
        for (int i = 0; i < S; i++) {
            free_from[i] = i;
            LSH_CHECK(hipHostMalloc((void **)&buf[i], nvals * sizeof(uint16_t), hipHostMallocDefault));
            LSH_CHECK(hipEventCreateWithFlags(&ev[i], hipEventDisableTiming));
        }
        for (unsigned t = 0; t < threads; t++) th.emplace_back([this] { work(); });
    }
    ~FileRing() {
        {
            std::lock_guard<std::mutex> lk(mu);
            stop = true;
        }
        cv.notify_all();
        for (auto &t : th) t.join();
        for (int i = 0; i < S; i++) {
            if (ev[i]) (void)hipEventSynchronize(ev[i]);
            if (ev[i]) (void)hipEventDestroy(ev[i]);
            if (buf[i]) (void)hipHostFree(buf[i]);
        }
    }
    void work() {
        (void)hipSetDevice(dev);
        for (;;) {
            const long j = next++;
            if (j >= (long)paths.size()) return;
            const int sl = (int)(j % S);
            bool sync_ev = false;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] { return stop || !err.empty() || free_from[sl] >= j; });
                if (stop || !err.empty()) return;
                sync_ev = pending[sl];
            }
            std::string e;
            try {
                if (sync_ev) LSH_CHECK(hipEventSynchronize(ev[sl]));
                io::Image im;
                im.ext = buf[sl];
                im.ext_cap = nvals;
                io::decode(paths[j], im);
                if (im.w != w || im.h != h)
                    e = "stream: " + paths[j] + " is " + std::to_string(im.w) + "x" + std::to_string(im.h) +
                        ", the stack is " + std::to_string(w) + "x" + std::to_string(h);
            } catch (const std::exception &x) {
                e = x.what();
            }
            {
                std::lock_guard<std::mutex> lk(mu);
                if (!e.empty() && err.empty()) err = e;
                if (e.empty()) ready[sl] = j;
            }
            cv.notify_all();
        }
    }
    // frame k, decoded, in pinned memory
    const uint16_t *get(uint32_t k) {
        const int sl = (int)(k % S);
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return !err.empty() || ready[sl] == (long)k; });
        if (!err.empty()) throw Error(err);
        return buf[sl];
    }
    // frame k's upload is queued on st (or done, st null): its slot is the next frame's once that has run
    void consumed(uint32_t k, hipStream_t st) {
        const int sl = (int)(k % S);
        if (st) LSH_CHECK(hipEventRecord(ev[sl], st));
        {
            std::lock_guard<std::mutex> lk(mu);
            pending[sl] = st != nullptr;
            ready[sl] = -1;
            free_from[sl] = (long)k + S;
        }
        cv.notify_all();
    }
};

// Where a stream's frames come from: device frames (a table of pointers), host
// frames (pointers to host memory), or files decoded ahead (a FileRing).
struct Source {
    const uint16_t *const *frames = nullptr;
    lsh_mem mem = LSH_DEVICE;
    FileRing *ring = nullptr;
    const uint16_t *host(uint32_t k) const { return ring ? ring->get(k) : frames[k]; }
    void done(uint32_t k, hipStream_t st) const {
        if (ring) ring->consumed(k, st);
    }
};

char *stream_impl(lsh_frames *f, lsh_fuser *fu, const Source &src, uint32_t n, const uint8_t *free8, uint32_t coarsen,
                  uint32_t brightness, const lsh_depth_params *measure, float *slices, lsh_mem slices_mem,
                  uint32_t graph, double *sims_out, float *gains_out, uint64_t *evals_out) {
    const lsh_mem frames_mem = src.mem;
    const uint16_t *const *frames = src.frames;
    LSH_REQUIRE(n > 0, "stream: no frames");
    LSH_REQUIRE(f->processed == 0 && fu->count == 0, "stream: the handles must be fresh");
    LSH_REQUIRE(f->w == fu->w && f->h == fu->h, "stream: frames and fuser sizes differ");
    LSH_REQUIRE(!measure || slices, "stream: measure needs slices");
    LSH_REQUIRE(!fu->halo_on, "stream: halo control is not supported");
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
    char *e = lsh_frames_process(f, src.host(0), frames_mem, ident, free8, coarsen, nullptr, nullptr, brightness,
                                 measure, sim0, g0, measure ? slices : nullptr, slices_mem);
    if (e) return e;
    src.done(0, nullptr);  // the frame was copied before the call returned
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
        frames_stream_slots(f);
        // The search's grid covers three quarters of the CUs here, so the side
        // work has the rest while it runs (a search block fills its CU, so a
        // full grid leaves the side work nothing but the gaps between
        // searches).  On the 100 x 45 MP stack this took the launched step
        // from 953 to 906 ms: the smallest grid keeping the finest level's
        // tile rounds (4) that frees eight CUs per XCD; 176 and 208 CUs were
        // both slower.  LSH_STREAM_SEARCH_CUS overrides it.
        int dev = 0, ncu = 0;
        LSH_CHECK(hipGetDevice(&dev));
        LSH_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        unsigned search_cus = (unsigned)ncu * 3 / 4;
        if (const char *env = std::getenv("LSH_STREAM_SEARCH_CUS"))
            if (std::atoi(env) > 0) search_cus = (unsigned)std::atoi(env);
        OnStream on_chain(stream(), search_cus);
        const hipStream_t Ms = stream(), Ss = side_stream();
        Events E;
        float *planes[2] = {f->planes.p, f->planes_b.p};
        // with a focus measure, the side warp also writes each frame's luma,
        // from which the focus slice reads one plane instead of three
        DBuf<float> fluma[2];
        if (measure)
            for (auto &b : fluma) b.alloc(npx);
        const size_t fbytes = 3 * npx * sizeof(uint16_t);
        // raw frames: the caller's device pointers (read through a device table
        // of them inside a replayed graph), or host frames uploaded into a ring
        // of NRING staging buffers (frame k in stage[k % NRING]) on the copy
        // stream, ahead of their first reader: frame k+1's upload waits only for
        // the side work of frame k+1-NRING (the last reader of its slot), so it
        // can run several cycles ahead.  Host frames in pinned memory upload
        // without a host copy.
        const bool staged = frames_mem != LSH_DEVICE;
        // eight slots (2.2 GB at 45 MP): against four, the uploads of the next
        // frames can run further ahead while the device works, which raised the
        // upload-inclusive rate 131.6 -> 134.0 frames/s (twelve: 135.7);
        // LSH_STREAM_RING overrides it
        int NRING = 8;
        if (const char *env = std::getenv("LSH_STREAM_RING"))
            if (std::atoi(env) >= 2 && std::atoi(env) <= 16) NRING = std::atoi(env);
        std::vector<DBuf<uint16_t>> stage(NRING);
        std::vector<hipEvent_t> e_up(NRING, nullptr), e_free(NRING, nullptr), e_half(NRING, nullptr);
        std::vector<char> slot_used(NRING, 0);
        if (staged)
            for (int i = 0; i < NRING; i++) {
                stage[i].alloc(3 * npx);
                e_up[i] = E.make();
                e_free[i] = E.make();
                e_half[i] = E.make();
            }
        const hipStream_t Cs = staged ? copy_stream() : nullptr, Cs2 = staged ? copy_stream2() : nullptr;
        DBuf<const uint16_t *> ftab;
        DBuf<unsigned> fcur(1);  // a replay's frame index (the graph's chain frame)
        if (graph) {
            std::vector<const uint16_t *> ring(NRING);
            for (int i = 0; i < NRING; i++) ring[i] = stage[i].p;
            ftab.alloc(staged ? NRING : n);
            copy_mem(ftab.p, LSH_DEVICE, staged ? ring.data() : frames, LSH_HOST, ftab.n * sizeof(const uint16_t *));
        }
        auto raw = [&](uint32_t k) -> const uint16_t * { return staged ? stage[k % NRING].p : frames[k]; };
        // frame k as a replayed cycle reads it: off = k - the cycle's frame index
        auto cyc = [&](int off) -> FrameRef { return FrameRef(ftab.p, fcur.p, off, staged ? NRING : 0); };
        // upload frame k into its slot once the slot's previous frame is done
        // with, in two halves on two copy streams (beside the device's work one
        // copy engine moved about 32 GB/s of the link's 57); the first stream
        // then waits for the second, so its event covers the whole frame
        auto put = [&](uint32_t k) {
            if (!staged) return;
            const int sl = (int)(k % NRING);
            if (slot_used[sl]) {
                LSH_CHECK(hipStreamWaitEvent(Cs, e_free[sl], 0));
                LSH_CHECK(hipStreamWaitEvent(Cs2, e_free[sl], 0));
            }
            slot_used[sl] = true;
            const size_t half = (fbytes / 2) & ~(size_t)255;
            const char *hs = (const char *)src.host(k);
            LSH_CHECK(hipMemcpyAsync(stage[sl].p, hs, half, hipMemcpyHostToDevice, Cs));
            LSH_CHECK(hipMemcpyAsync((char *)stage[sl].p + half, hs + half, fbytes - half, hipMemcpyHostToDevice, Cs2));
            LSH_CHECK(hipEventRecord(e_half[sl], Cs2));
            LSH_CHECK(hipStreamWaitEvent(Cs, e_half[sl], 0));
            src.done(k, Cs);
            LSH_CHECK(hipEventRecord(e_up[sl], Cs));
        };
        // stream st reads frame k only after its upload; frame k's slot is free
        // once what st has queued so far has run
        auto ready = [&](uint32_t k, hipStream_t st) {
            if (staged) LSH_CHECK(hipStreamWaitEvent(st, e_up[k % NRING], 0));
        };
        auto release = [&](uint32_t k, hipStream_t st) {
            if (staged) LSH_CHECK(hipEventRecord(e_free[k % NRING], st));
        };
        // slices: in place in the caller's device buffer (launches: at frame k's
        // offset; a graph: at the fold index's), else through device buffers
        // copied out
        const bool direct = measure && slices_mem == LSH_DEVICE;
        DBuf<float> sall, sbuf[2];
        if (measure && !direct) {
            if (graph) {
                sbuf[0].alloc(M);
                sbuf[1].alloc(M);
            } else {
                sall.alloc((size_t)n * M);
            }
        }
        auto slice_at = [&](uint32_t k) -> float * {
            if (!measure) return nullptr;
            if (direct) return graph ? slices : slices + (size_t)k * M;
            return graph ? sbuf[k & 1].p : sall.p + (size_t)k * M;
        };
        DBuf<unsigned> fidx(1);  // the fold's frame index (side work only)
        // a graph writes the slices through a device copy of their address:
        // the runtime keeps any buffer a captured launch's arguments name from
        // going back to the device when the caller frees it, even after the
        // graph is destroyed (on the 45 MP grid: the whole slice stack per call)
        DBuf<float *> sbase(1);
        if (graph && direct) copy_mem(sbase.p, LSH_DEVICE, &slices, LSH_HOST, sizeof(float *));
        launch("stream_fold_index", k_counter_set, dim3(1), dim3(64), 0, fidx.p, 1u);
        const hipEvent_t e_main = E.make();
        hipEvent_t e_target[2] = {E.make(), E.make()}, e_warp[2] = {E.make(), E.make()},
                   e_side[2] = {E.make(), E.make()};
        const bool bright = brightness != 0;
        // the pieces; each goes to the stream current when it is called
        auto chain = [&](uint32_t k, bool plain, hipEvent_t started) { frames_chain(f, (int)(k & 1), plain, started); };
        // frame k's side work: its gains, RGB warp and record, its focus slice
        // and its fold (the fold index is k)
        auto side = [&](uint32_t k, const FrameRef &src) {
            frames_side_warp(f, src, (int)(k & 1), planes[k & 1], bright, measure ? fluma[k & 1].p : nullptr);
            if (measure)
                focus_slice(f, fluma[k & 1].p, *measure, graph && direct ? nullptr : slice_at(k),
                            graph && direct ? fidx.p : nullptr, true, graph && direct ? sbase.p : nullptr);
            fuser_push_dev(fu, planes[k & 1], fidx.p);
            launch("stream_fold_next", k_counter_next, dim3(1), dim3(64), 0, fidx.p);
        };
        auto target = [&](uint32_t k) {  // frame k's target pyramid into slot k % 2 (on the current stream)
            put(k);
            ready(k, stream());
            frames_target(f, raw(k), (int)(k & 1));
        };
        // the side stream starts after frame 0 (the fuser's state, the frames handle's)
        LSH_CHECK(hipEventRecord(e_main, Ms));
        LSH_CHECK(hipStreamWaitEvent(Ss, e_main, 0));
        {
            OnStream on(Ss);
            target(1);
            LSH_CHECK(hipEventRecord(e_target[1], Ss));
            if (n > 2) {
                target(2);
                LSH_CHECK(hipEventRecord(e_target[0], Ss));
            }
        }
        if (!graph) {
            // frame k-1's side work and frame k+1's target follow frame k's
            // search reset on the side stream (dispatched behind the search, and
            // after frame k-1's chain, the last reader of slot (k+1) % 2)
            const hipEvent_t e_started = E.make();
            for (uint32_t k = 1; k < n; k++) {
                LSH_CHECK(hipStreamWaitEvent(Ms, e_target[k & 1], 0));
                // FrameDev k % 2: frame k-2's side work reads it
                if (k >= 3) LSH_CHECK(hipStreamWaitEvent(Ms, e_side[k & 1], 0));
                chain(k, true, e_started);
                if (k < 2) continue;
                OnStream on(Ss);
                LSH_CHECK(hipStreamWaitEvent(Ss, e_started, 0));
                side(k - 1, raw(k - 1));
                release(k - 1, Ss);
                LSH_CHECK(hipEventRecord(e_side[(k - 1) & 1], Ss));
                if (k + 1 < n) {
                    target(k + 1);
                    LSH_CHECK(hipEventRecord(e_target[(k + 1) & 1], Ss));
                }
            }
            LSH_CHECK(hipEventRecord(e_warp[0], Ms));  // the last frame's side work
            OnStream on(Ss);
            LSH_CHECK(hipStreamWaitEvent(Ss, e_warp[0], 0));
            side(n - 1, raw(n - 1));
        } else {
            // frame 1 and the last frame run eagerly (their cycles differ); frames 2 .. n-2 replay a graph
            // holding frame k's chain beside frame k-1's side work and frame k+1's target pyramid
            LSH_CHECK(hipStreamWaitEvent(Ms, e_target[1], 0));
            chain(1, true, nullptr);
            Graphs G;
            // device frames: a replay reads frame k + off as entry fcur + off of the table
            launch("stream_frame_index", k_counter_set, dim3(1), dim3(64), 0, fcur.p, 2u);
            // the host-side counts the captured cycle advances once per replay
            auto cycle = [&](uint32_t k) {
                chain(k, true, e_main);  // fork after search k's reset
                {
                    OnStream on(Ss);
                    LSH_CHECK(hipStreamWaitEvent(Ss, e_main, 0));
                    side(k - 1, cyc(-1));
                    if (k + 1 < n) frames_target(f, cyc(1), (int)((k + 1) & 1));
                    LSH_CHECK(hipEventRecord(e_side[0], Ss));  // join
                }
                LSH_CHECK(hipStreamWaitEvent(Ms, e_side[0], 0));
                launch("stream_frame_next", k_counter_next, dim3(1), dim3(64), 0, fcur.p);
            };
            LSH_CHECK(hipStreamWaitEvent(Ms, e_target[0], 0));  // the prologue's side work is done
            for (uint32_t k = 2; k < n; k++) {
                if (k + 1 < n) {
                    // host frames: frame k+1 uploaded for the replay's target
                    put(k + 1);
                    ready(k + 1, Ms);
                    const int p = (int)(k & 1);
                    if (!G.x[p]) {
                        prof().capturing = true;
                        LSH_CHECK(hipStreamBeginCapture(Ms, hipStreamCaptureModeThreadLocal));
                        try {
                            cycle(k);
                        } catch (...) {
                            hipGraph_t junk = nullptr;
                            (void)hipStreamEndCapture(Ms, &junk);
                            if (junk) (void)hipGraphDestroy(junk);
                            prof().capturing = false;
                            throw;
                        }
                        LSH_CHECK(hipStreamEndCapture(Ms, &G.g[p]));
                        prof().capturing = false;
                        LSH_CHECK(hipGraphInstantiate(&G.x[p], G.g[p], nullptr, nullptr, 0));
                    } else {
                        f->processed++;
                        fu->count++;
                    }
                    LSH_CHECK(hipGraphLaunch(G.x[p], Ms));
                    release(k - 1, Ms);  // the replay's side work was frame k-1's last reader
                } else {
                    cycle(k);  // the last frame: no target after it
                }
                if (measure && !direct)
                    copy_mem(slices + (size_t)(k - 1) * M, slices_mem, sbuf[(k - 1) & 1].p, LSH_DEVICE,
                             M * sizeof(float));
            }
            // the last frame's side work
            LSH_CHECK(hipEventRecord(e_main, Ms));
            {
                OnStream on(Ss);
                LSH_CHECK(hipStreamWaitEvent(Ss, e_main, 0));
                side(n - 1, raw(n - 1));
                LSH_CHECK(hipEventRecord(e_side[1], Ss));
            }
            LSH_CHECK(hipStreamWaitEvent(Ms, e_side[1], 0));
            if (measure && !direct)
                copy_mem(slices + (size_t)(n - 1) * M, slices_mem, sbuf[(n - 1) & 1].p, LSH_DEVICE,
                         M * sizeof(float));
        }
        sync();
        if (measure && !direct && !graph)
            copy_mem(slices + M, slices_mem, sall.p + M, LSH_DEVICE, (size_t)(n - 1) * M * sizeof(float));
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
    return nullptr;
}

}  // namespace

extern "C" {

char *lsh_stack_stream(lsh_frames *f, lsh_fuser *fu, const uint16_t *const *frames, uint32_t n, lsh_mem frames_mem,
                       const uint8_t *free8, uint32_t coarsen, uint32_t brightness, const lsh_depth_params *measure,
                       float *slices, lsh_mem slices_mem, uint32_t graph, double *sims_out, float *gains_out,
                       uint64_t *evals_out) {
    LSH_API_BEGIN
    LSH_REQUIRE(f && fu && frames && free8, "lsh_stack_stream: null argument");
    Source src;
    src.frames = frames;
    src.mem = frames_mem;
    return stream_impl(f, fu, src, n, free8, coarsen, brightness, measure, slices, slices_mem, graph, sims_out,
                       gains_out, evals_out);
    LSH_API_END
}

char *lsh_stack_stream_files(lsh_frames *f, lsh_fuser *fu, const char *const *paths, uint32_t n, uint32_t threads,
                             const uint8_t *free8, uint32_t coarsen, uint32_t brightness,
                             const lsh_depth_params *measure, float *slices, lsh_mem slices_mem, uint32_t graph,
                             double *sims_out, float *gains_out, uint64_t *evals_out) {
    LSH_API_BEGIN
    LSH_REQUIRE(f && fu && paths && free8, "lsh_stack_stream_files: null argument");
    LSH_REQUIRE(n > 0, "stream: no frames");
    for (uint32_t i = 0; i < n; i++) LSH_REQUIRE(paths[i], "lsh_stack_stream_files: null path");
    FileRing ring(paths, n, f->w, f->h, threads ? threads : 8);
    Source src;
    src.mem = LSH_HOST;
    src.ring = &ring;
    char *e = stream_impl(f, fu, src, n, free8, coarsen, brightness, measure, slices, slices_mem, graph, sims_out,
                          gains_out, evals_out);
    sync();  // every upload out of the ring has run before its buffers go
    return e;
    LSH_API_END
}

}  // extern "C"
