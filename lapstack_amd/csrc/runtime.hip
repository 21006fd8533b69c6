// runtime.hip -- the library's stream, event timing, copies, homography math
// and the library-level C ABI (probe, version, errors, stream, timing).
#include <cstdio>
#include <map>
#include <vector>
#include <cstdlib>

#include "api.hpp"
#include "common.hpp"

namespace lsh {

namespace {
std::mutex g_mu;
struct DevState {
    hipStream_t s = nullptr;
    hipStream_t side = nullptr;  // a second stream, for work that runs beside the main chain
    hipStream_t copy = nullptr;  // a third, for host-to-device uploads beside both
    hipStream_t copy2 = nullptr;  // a fourth: uploads split over two copy engines
    void *pinned = nullptr;
    size_t pinned_n = 0;
    Prof prof;
};
std::map<int, DevState> g_dev;

DevState &dev_state() {
    int d = 0;
    LSH_CHECK(hipGetDevice(&d));
    std::lock_guard<std::mutex> lk(g_mu);
    DevState &st = g_dev[d];
    if (!st.s) LSH_CHECK(hipStreamCreateWithFlags(&st.s, hipStreamNonBlocking));
    return st;
}
}  // namespace

thread_local hipStream_t t_override = nullptr;  // OnStream: this thread's launches go there
thread_local unsigned t_cus = 0;                 // and the CUs that stream may use (0: all)

hipStream_t stream() { return t_override ? t_override : dev_state().s; }

unsigned stream_cus() { return t_override ? t_cus : 0; }

hipStream_t side_stream() {
    DevState &st = dev_state();
    if (!st.side) LSH_CHECK(hipStreamCreateWithFlags(&st.side, hipStreamNonBlocking));
    return st.side;
}

hipStream_t copy_stream2() {
    DevState &st = dev_state();
    if (!st.copy2) LSH_CHECK(hipStreamCreateWithFlags(&st.copy2, hipStreamNonBlocking));
    return st.copy2;
}

hipStream_t copy_stream() {
    DevState &st = dev_state();
    if (!st.copy) LSH_CHECK(hipStreamCreateWithFlags(&st.copy, hipStreamNonBlocking));
    return st.copy;
}

OnStream::OnStream(hipStream_t s, unsigned cus) : prev(t_override), prev_cus(t_cus) {
    t_override = s;
    t_cus = cus;
}
OnStream::~OnStream() {
    t_override = prev;
    t_cus = prev_cus;
}

void sync() {
    DevState &st = dev_state();
    if (t_override && t_override != st.s && t_override != st.side) LSH_CHECK(hipStreamSynchronize(t_override));
    LSH_CHECK(hipStreamSynchronize(st.s));
    if (st.side) LSH_CHECK(hipStreamSynchronize(st.side));
    if (st.copy) LSH_CHECK(hipStreamSynchronize(st.copy));
    if (st.copy2) LSH_CHECK(hipStreamSynchronize(st.copy2));
}

void *pinned_scratch(size_t n) {
    DevState &st = dev_state();
    if (st.pinned_n < n) {
        sync();
        if (st.pinned) (void)hipHostFree(st.pinned);
        st.pinned = nullptr;
        st.pinned_n = 0;
        LSH_CHECK(hipHostMalloc(&st.pinned, n, hipHostMallocDefault));
        st.pinned_n = n;
    }
    return st.pinned;
}

Prof &prof() { return dev_state().prof; }

// ------------------------------------------------------------ device memory
// The library's device buffers come and go with every call (a depth pass
// takes and returns a few GB at the 45 MP working grid, in a few dozen
// blocks).  hipFree waits for the whole device and unmaps the block, so
// freeing them between the steps of a call leaves the device idle for
// milliseconds each time.  Released blocks are kept instead, per device and
// size class (powers of two up to 1 MiB, then multiples of 2 MiB), each with
// an event recorded on each of the library's streams at its release; a block
// is handed out again only once all of its events have completed, so no work
// queued before the release can still be using it; a request may take a
// cached block of up to 5/4 of its class.  The cache holds at most an eighth
// of the device's memory (at least 4 GiB): past that, released blocks whose
// events have completed go back to the device, largest first, so buffers whose
// sizes change from call to call (a crop, a frame count) do not pile up.  When
// the device has no room for a new block, every cached block is returned to it
// and the allocation tried again.  LSH_POOL=0 turns the cache off.
namespace {
struct PoolBlock {
    void *p;
    hipEvent_t ev[4];
    int nev;
};
struct DevPool {
    std::multimap<size_t, PoolBlock> freed;  // by the block's own size
    std::map<void *, size_t> live;           // blocks handed out: their own sizes
    std::vector<hipEvent_t> spare;
    size_t cached = 0, cap = 0;
};
struct Pools {
    std::mutex mu;
    std::map<int, DevPool> dev;
    int on = -1;
};
Pools &pools() {
    static Pools *p = new Pools();  // never destroyed: buffers may be released during exit
    return *p;
}
size_t size_class(size_t b) {
    if (b <= ((size_t)1 << 20)) {
        size_t c = 256;
        while (c < b) c <<= 1;
        return c;
    }
    const size_t g = (size_t)2 << 20;
    return (b + g - 1) / g * g;
}
bool pool_on() {
    Pools &P = pools();
    if (P.on < 0) {
        const char *e = std::getenv("LSH_POOL");
        P.on = !(e && std::atoi(e) == 0);
    }
    return P.on != 0;
}
bool block_ready(const PoolBlock &b) {
    for (int i = 0; i < b.nev; i++)
        if (hipEventQuery(b.ev[i]) != hipSuccess) return false;
    return true;
}
// returns every cached block of the current device to it (the caller holds the lock)
void pool_trim(DevPool &D) {
    for (auto &kv : D.freed) {
        for (int i = 0; i < kv.second.nev; i++) {
            (void)hipEventSynchronize(kv.second.ev[i]);
            D.spare.push_back(kv.second.ev[i]);
        }
        (void)hipFree(kv.second.p);
    }
    D.freed.clear();
    D.cached = 0;
}
// returns completed cached blocks, largest first, until the cache is within its cap
void pool_fit(DevPool &D) {
    for (auto it = D.freed.end(); D.cached > D.cap && it != D.freed.begin();) {
        --it;
        if (!block_ready(it->second)) continue;
        for (int i = 0; i < it->second.nev; i++) D.spare.push_back(it->second.ev[i]);
        (void)hipFree(it->second.p);
        D.cached -= it->first;
        it = D.freed.erase(it);
    }
}
}  // namespace

void *dev_alloc(size_t bytes, int *dev) {
    LSH_CHECK(hipGetDevice(dev));
    if (!pool_on()) {
        void *p = nullptr;
        LSH_CHECK(hipMalloc(&p, bytes));
        return p;
    }
    const size_t cls = size_class(bytes);
    Pools &P = pools();
    std::lock_guard<std::mutex> lk(P.mu);
    DevPool &D = P.dev[*dev];
    if (!D.cap) {
        size_t fr = 0, total = 0;
        LSH_CHECK(hipMemGetInfo(&fr, &total));
        D.cap = std::max(total / 8, (size_t)4 << 30);
    }
    for (auto it = D.freed.lower_bound(cls); it != D.freed.end() && it->first <= cls + cls / 4; ++it)
        if (block_ready(it->second)) {
            void *p = it->second.p;
            for (int i = 0; i < it->second.nev; i++) D.spare.push_back(it->second.ev[i]);
            D.cached -= it->first;
            D.live[p] = it->first;
            D.freed.erase(it);
            return p;
        }
    void *p = nullptr;
    if (hipMalloc(&p, cls) != hipSuccess) {
        (void)hipGetLastError();
        pool_trim(D);
        LSH_CHECK(hipMalloc(&p, cls));
    }
    D.live[p] = cls;
    return p;
}

void dev_free(void *p, size_t bytes, int dev) {
    if (!p) return;
    if (!pool_on()) {
        (void)hipFree(p);
        return;
    }
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess) return;
    if (cur != dev && hipSetDevice(dev) != hipSuccess) return;
    {
        DevState &st = dev_state();
        Pools &P = pools();
        std::lock_guard<std::mutex> lk(P.mu);
        DevPool &D = P.dev[dev];
        auto lv = D.live.find(p);
        const size_t size = lv != D.live.end() ? lv->second : size_class(bytes);
        if (lv != D.live.end()) D.live.erase(lv);
        PoolBlock b{p, {nullptr, nullptr, nullptr, nullptr}, 0};
        const hipStream_t ss[4] = {st.s, st.side, st.copy, st.copy2};
        bool ok = true;
        for (hipStream_t s : ss) {
            if (!s) continue;
            hipEvent_t e = nullptr;
            if (!D.spare.empty()) {
                e = D.spare.back();
                D.spare.pop_back();
            } else if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
                ok = false;
                break;
            }
            if (hipEventRecord(e, s) != hipSuccess) {
                D.spare.push_back(e);
                ok = false;
                break;
            }
            b.ev[b.nev++] = e;
        }
        if (ok) {
            D.freed.emplace(size, b);
            D.cached += size;
            if (D.cached > D.cap) pool_fit(D);
        } else {  // no event to guard the block with: give it back to the device
            for (int i = 0; i < b.nev; i++) D.spare.push_back(b.ev[i]);
            (void)hipGetLastError();
            (void)hipFree(p);
        }
    }
    if (cur != dev) (void)hipSetDevice(cur);
}

hipEvent_t Prof::get() {
    if (pool.empty()) {
        hipEvent_t e;
        LSH_CHECK(hipEventCreate(&e));
        return e;
    }
    hipEvent_t e = pool.back();
    pool.pop_back();
    return e;
}

void Prof::flush() {
    if (pending.empty()) return;
    LSH_CHECK(hipEventSynchronize(pending.back().b));
    for (auto &r : pending) {
        float ms = 0.f;
        LSH_CHECK(hipEventElapsedTime(&ms, r.a, r.b));
        auto &t = totals[r.name];
        t.ms += ms;
        t.bytes += r.bytes;
        t.n += 1;
        pool.push_back(r.a);
        pool.push_back(r.b);
    }
    pending.clear();
}

void copy_mem(void *dst, lsh_mem dmem, const void *src, lsh_mem smem, size_t bytes) {
    if (!bytes) return;
    hipMemcpyKind kind = dmem == LSH_DEVICE ? (smem == LSH_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice)
                                            : (smem == LSH_DEVICE ? hipMemcpyDeviceToHost : hipMemcpyHostToHost);
    LSH_CHECK(hipMemcpyAsync(dst, src, bytes, kind, stream()));
    if (dmem == LSH_HOST || smem == LSH_HOST) sync();
}

// align.rs:241-252
void sim_matrix(const Sim &sim, size_t w, size_t h, double *m) {
    const double *s = sim.v;
    double cx = (double)w / 2.0, cy = (double)h / 2.0;
    double c, sn;
    sincos_cr(s[3], &sn, &c);  // correctly rounded: the device's maps take the same bits
    double sc = s[2];
    double a = sc * c, b = sc * (c * s[5] - sn * s[4]), d = sc * sn, e = sc * (sn * s[5] + c * s[4]);
    double gx = s[6] / (double)w, gy = s[7] / (double)h;
    double a2 = a + cx * gx, b2 = b + cx * gy, d2 = d + cy * gx, e2 = e + cy * gy;
    double tx = cx + s[0] * (double)w - (a2 * cx + b2 * cy);
    double ty = cy + s[1] * (double)h - (d2 * cx + e2 * cy);
    m[0] = a2;
    m[1] = b2;
    m[2] = tx;
    m[3] = d2;
    m[4] = e2;
    m[5] = ty;
    m[6] = gx;
    m[7] = gy;
    m[8] = 1.0 - gx * cx - gy * cy;
}

// align.rs:258-286
void inverse9(const double *m, double *inv) {
    if (m[6] == 0.0 && m[7] == 0.0 && m[8] == 1.0) {
        double a = m[0], b = m[1], tx = m[2], d = m[3], e = m[4], ty = m[5];
        double det = a * e - b * d;
        double ia = e / det, ib = -b / det, id = -d / det, ie = a / det;
        inv[0] = ia;
        inv[1] = ib;
        inv[2] = -(ia * tx + ib * ty);
        inv[3] = id;
        inv[4] = ie;
        inv[5] = -(id * tx + ie * ty);
        inv[6] = 0.0;
        inv[7] = 0.0;
        inv[8] = 1.0;
        return;
    }
    double t[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            int r0 = (j + 1) % 3, r1 = (j + 2) % 3, c0 = (i + 1) % 3, c1 = (i + 2) % 3;
            t[i * 3 + j] = m[r0 * 3 + c0] * m[r1 * 3 + c1] - m[r0 * 3 + c1] * m[r1 * 3 + c0];
        }
    double k = t[8];
    for (int i = 0; i < 9; i++) inv[i] = t[i] / k;
}

char *make_err(const std::string &msg) {
    char *s = (char *)std::malloc(msg.size() + 1);
    if (s) std::memcpy(s, msg.c_str(), msg.size() + 1);
    return s;
}

}  // namespace lsh

using namespace lsh;

extern "C" {

uint32_t lsh_abi(void) { return LSH_ABI_VERSION; }

void lsh_free_str(char *s) { std::free(s); }

char *lsh_available(void) {
    LSH_API_BEGIN
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) return make_err("no HIP device (is the amdgpu driver loaded?)");
    int d = 0;
    LSH_CHECK(hipGetDevice(&d));
    hipDeviceProp_t prop;
    LSH_CHECK(hipGetDeviceProperties(&prop, d));
    std::string arch = prop.gcnArchName;
    if (arch.rfind("gfx950", 0) != 0) return make_err("device " + std::to_string(d) + " is " + arch + ", not gfx950");
    (void)stream();
    LSH_API_END
}

char *lsh_stream(void **out) {
    LSH_API_BEGIN
    LSH_REQUIRE(out, "lsh_stream: null output");
    *out = (void *)stream();
    LSH_API_END
}

char *lsh_sync(void) {
    LSH_API_BEGIN
    sync();
    LSH_API_END
}

char *lsh_mem_stats(uint64_t *live, uint64_t *cached) {
    LSH_API_BEGIN
    LSH_REQUIRE(live && cached, "lsh_mem_stats: null output");
    int dev = 0;
    LSH_CHECK(hipGetDevice(&dev));
    Pools &P = pools();
    std::lock_guard<std::mutex> lk(P.mu);
    DevPool &D = P.dev[dev];
    size_t l = 0;
    for (auto &kv : D.live) l += kv.second;
    *live = l;
    *cached = D.cached;
    LSH_API_END
}

char *lsh_mem_release(void) {
    LSH_API_BEGIN
    int dev = 0;
    LSH_CHECK(hipGetDevice(&dev));
    Pools &P = pools();
    std::lock_guard<std::mutex> lk(P.mu);
    pool_trim(P.dev[dev]);
    LSH_API_END
}

char *lsh_prof_enable(uint32_t on) {
    LSH_API_BEGIN
    Prof &p = prof();
    if (!on) p.flush();
    p.on = on != 0;
    LSH_API_END
}

char *lsh_prof_reset(void) {
    LSH_API_BEGIN
    Prof &p = prof();
    p.flush();
    p.totals.clear();
    LSH_API_END
}

char *lsh_prof_read(char *names, size_t names_cap, double *total_ms, double *total_bytes, uint64_t *counts, uint32_t cap,
                    uint32_t *n) {
    LSH_API_BEGIN
    Prof &p = prof();
    p.flush();
    std::string all;
    uint32_t k = 0;
    for (auto &kv : p.totals) {
        if (k < cap) {
            if (total_ms) total_ms[k] = kv.second.ms;
            if (total_bytes) total_bytes[k] = kv.second.bytes;
            if (counts) counts[k] = kv.second.n;
            all += kv.first;
            all += '\n';
        }
        k++;
    }
    if (names && names_cap) {
        size_t m = all.size() < names_cap - 1 ? all.size() : names_cap - 1;
        std::memcpy(names, all.data(), m);
        names[m] = 0;
    }
    if (n) *n = k < cap ? k : cap;
    LSH_API_END
}

}  // extern "C"
