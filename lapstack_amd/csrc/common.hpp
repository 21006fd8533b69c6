// common.hpp -- runtime plumbing shared by the gfx950 kernels: errors, device
// buffers, the library's stream, launches with optional event timing, and the
// arithmetic helpers every kernel must share with the CPU path bit for bit.
//
// All sources are compiled with -ffp-contract=off: the CPU path sums taps
// left to right without fused multiply-adds, and so do these kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/lsh.h"
#include "sincos.hpp"

namespace lsh {

struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define LSH_CHECK(expr)                                                                                    \
    do {                                                                                                   \
        hipError_t e_ = (expr);                                                                            \
        if (e_ != hipSuccess) throw ::lsh::Error(std::string(#expr) + ": " + hipGetErrorString(e_));       \
    } while (0)

#define LSH_REQUIRE(cond, msg)                 \
    do {                                       \
        if (!(cond)) throw ::lsh::Error(msg);  \
    } while (0)

// ---------------------------------------------------------------- runtime
hipStream_t stream();          // the library's stream on the current device (or the OnStream one)
hipStream_t side_stream();     // a second stream on the current device (default priority), made on first use
hipStream_t copy_stream();     // a third, for uploads (default priority), made on first use
hipStream_t copy_stream2();    // a fourth, for the other half of split uploads (default priority)
void sync();                   // wait for both (and for the OnStream stream, if another)
// While one lives, this thread's launches and copies go to stream s; cus: the
// number of CUs s may use when it is restricted to a subset (0: all of them).
struct OnStream {
    hipStream_t prev;
    unsigned prev_cus;
    explicit OnStream(hipStream_t s, unsigned cus = 0);
    ~OnStream();
    OnStream(const OnStream &) = delete;
    OnStream &operator=(const OnStream &) = delete;
};
unsigned stream_cus();  // the CU budget of the current stream (0: the whole device)
void *pinned_scratch(size_t);  // a per-device pinned host block (reused; at least the size asked)
// Device memory for the library's buffers, on the current device, from a
// per-device cache of released blocks (dev_free); dev is the device it was
// taken on.  A released block is handed out again only once the work the
// library's streams had queued when it was released has run.
void *dev_alloc(size_t bytes, int *dev);
void dev_free(void *p, size_t bytes, int dev);

// ------------------------------------------------------------- profiling
struct Prof {
    bool on = false;
    bool capturing = false;  // a graph is being captured on the stream: no timing events
    struct Rec {
        const char *name;
        double bytes;
        hipEvent_t a, b;
    };
    struct Tot {
        double ms = 0.0, bytes = 0.0;
        uint64_t n = 0;
    };
    std::vector<Rec> pending;
    std::map<std::string, Tot> totals;
    std::vector<hipEvent_t> pool;
    hipEvent_t get();
    void flush();  // synchronizes and folds pending records into totals
};
Prof &prof();

// Launch on the library's stream.  `bytes` is the launch's algorithmic HBM
// traffic (each array it must read or write, once, in its stored format),
// recorded with the event timing so that achieved GB/s per kernel comes out.
template <typename K, typename... A>
inline void launch_b(const char *name, double bytes, K kernel, dim3 grid, dim3 block, size_t shmem, A... args) {
    if (grid.x == 0 || grid.y == 0 || grid.z == 0) return;
    Prof &p = prof();
    hipStream_t s = stream();
    if (p.on && !p.capturing) {
        hipEvent_t a = p.get(), b = p.get();
        LSH_CHECK(hipEventRecord(a, s));
        hipLaunchKernelGGL(kernel, grid, block, shmem, s, args...);
        LSH_CHECK(hipGetLastError());
        LSH_CHECK(hipEventRecord(b, s));
        p.pending.push_back({name, bytes, a, b});
        if (p.pending.size() > 4096) p.flush();
    } else {
        hipLaunchKernelGGL(kernel, grid, block, shmem, s, args...);
        LSH_CHECK(hipGetLastError());
    }
}

template <typename K, typename... A>
inline void launch(const char *name, K kernel, dim3 grid, dim3 block, size_t shmem, A... args) {
    launch_b(name, 0.0, kernel, grid, block, shmem, args...);
}

// A cooperative launch (every block resident at once, as grid barriers need).
inline void launch_coop(const char *name, double bytes, const void *kernel, dim3 grid, dim3 block, void **args,
                        size_t shmem) {
    Prof &p = prof();
    hipStream_t s = stream();
    hipEvent_t a = nullptr, b = nullptr;
    if (p.on) {
        a = p.get();
        b = p.get();
        LSH_CHECK(hipEventRecord(a, s));
    }
    LSH_CHECK(hipLaunchCooperativeKernel(kernel, grid, block, args, (unsigned)shmem, s));
    LSH_CHECK(hipGetLastError());
    if (p.on) {
        LSH_CHECK(hipEventRecord(b, s));
        p.pending.push_back({name, bytes, a, b});
    }
}

// Set the algorithmic bytes of the last timed launch (known only once it has run).
inline void prof_set_last_bytes(double bytes) {
    Prof &p = prof();
    if (p.on && !p.pending.empty()) p.pending.back().bytes = bytes;
}

inline unsigned cdiv(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

// --------------------------------------------------------------- buffers
// Device memory the buffer owns (alloc), or a view of memory it does not own
// (view: released without a free).
template <typename T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    bool owned = false;
    int dev = -1;  // the device an owned block was taken on
    DBuf() = default;
    explicit DBuf(size_t count) { alloc(count); }
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    DBuf(DBuf &&o) noexcept : p(o.p), n(o.n), owned(o.owned), dev(o.dev) {
        o.p = nullptr;
        o.n = 0;
        o.owned = false;
    }
    DBuf &operator=(DBuf &&o) noexcept {
        if (this != &o) {
            release();
            p = o.p;
            n = o.n;
            owned = o.owned;
            dev = o.dev;
            o.p = nullptr;
            o.n = 0;
            o.owned = false;
        }
        return *this;
    }
    ~DBuf() { release(); }
    void alloc(size_t count) {
        release();
        if (count) p = (T *)dev_alloc(count * sizeof(T), &dev);
        n = count;
        owned = true;
    }
    void view(T *q, size_t count) {
        release();
        p = q;
        n = count;
        owned = false;
    }
    void ensure(size_t count) {
        if (count > n) alloc(count);
    }
    void release() {
        if (p && owned) dev_free(p, n * sizeof(T), dev);
        p = nullptr;
        n = 0;
        owned = false;
    }
    void zero() {
        if (n) LSH_CHECK(hipMemsetAsync(p, 0, n * sizeof(T), stream()));
    }
    size_t bytes() const { return n * sizeof(T); }
};

// copy count elements between any two lsh_mem locations on the library's stream
void copy_mem(void *dst, lsh_mem dmem, const void *src, lsh_mem smem, size_t bytes);

// ----------------------------------------------------- shared arithmetic
// Reflect-101 (pyramid.rs:51-66): -1 -> 1, n -> n-2; loops for tiny n.
__host__ __device__ __forceinline__ int rf101(int i, int n) {
    if (n == 1) return 0;
    for (;;) {
        if (i < 0)
            i = -i;
        else if (i >= n)
            i = 2 * (n - 1) - i;
        else
            return i;
    }
}

// Stages COUNT items into LDS with NT threads: each thread issues all of its
// loads before it stores any of them, so a tile waits on memory once rather
// than once per item (a plain strided loop waits after every load).  ld(k, v)
// loads item k into v and returns whether the item exists; st(k, v) stores it.
template <int COUNT, int NT, class T, class Ld, class St>
__device__ __forceinline__ void lds_stage(const Ld &ld, const St &st) {
    constexpr int IT = (COUNT + NT - 1) / NT;
    T v[IT];
    bool ok[IT];
#pragma unroll
    for (int i = 0; i < IT; i++) {
        const int k = (int)threadIdx.x + i * NT;
        ok[i] = k < COUNT && ld(k, v[i]);
    }
#pragma unroll
    for (int i = 0; i < IT; i++)
        if (ok[i]) st((int)threadIdx.x + i * NT, v[i]);
}

// Rust f32::clamp
__host__ __device__ __forceinline__ float rclampf(float v, float lo, float hi) {
    if (v < lo) v = lo;
    if (v > hi) v = hi;
    return v;
}

// f32::total_cmp key
__host__ __device__ __forceinline__ int32_t total_key(float v) {
    int32_t b = __builtin_bit_cast(int32_t, v);
    return b ^ (int32_t)(((uint32_t)(b >> 31)) >> 1);
}

// A Sim (align.rs:199-253) as 8 doubles; the homography math runs on the host.
struct Sim {
    double v[8] = {0, 0, 1, 0, 1, 0, 0, 0};
    bool is_id() const {
        static const double id[8] = {0, 0, 1, 0, 1, 0, 0, 0};
        for (int k = 0; k < 8; k++)
            if (v[k] != id[k]) return false;
        return true;
    }
};
void sim_matrix(const Sim &s, size_t w, size_t h, double *m9);  // align.rs:241-252
void inverse9(const double *m, double *inv);                     // align.rs:258-286

// A 3x3 inverse homography passed by value to kernels.
struct Inv9 {
    double m[9];
};

}  // namespace lsh
