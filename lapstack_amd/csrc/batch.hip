// batch.hip -- many independent stacks of uneven sizes and frame counts
// through the fuse path (the frame as decoded, REDUCE chain, region-energy
// fold, residual store; fuse.rs:401-484 per frame) in one persistent
// cooperative launch.
//
// The launch walks rounds; round r takes frames [rF, rF + F) of every stack
// that has them (F, the frames per round, is tuned).  Each stack owns F
// pyramid slots (levels 1..L; level 0 is never written: its readers take the
// u16 frame itself).  A round is L_max + 1 phases; phase p holds, per stack:
//   - REDUCE of level p of every frame of the round (p <= L; level 1 straight
//     from the u16 frame), 64x4 output tiles, the three planes per item;
//   - the fold of level p - 2 (2 <= p <= L + 1; its inputs G_{p-2}, G_{p-1}
//     are complete since phase p - 1), one 64 x TY tile per item running all
//     the round's frames in order with the best energy, the accumulator and
//     the winner of the tile's pixels held in registers, so best is read and
//     the results written once per F frames;
//   - the residual level of each frame into the store (p = L + 1).
// A phase's work items are numbered through per-(round, phase) prefix sums
// over the stacks (sorted by frame count, so the active ones are a prefix);
// blocks take items by a grid-stride walk and a grid barrier closes the phase.
// Every sample is computed by the tile bodies of the per-frame kernels
// (fuse_tiles.hpp) in the per-frame order, so each stack's result is the
// per-stack result bit for bit.  The fold tile height, the blocks per CU and F
// are chosen by an autotuner that times round 0 (which only writes, so it can
// be replayed) per candidate, cached per shape class.
#include <algorithm>
#include <map>
#include <numeric>
#include <tuple>
#include <vector>

#include "api.hpp"
#include "common.hpp"
#include "engine.hpp"
#include "fuse_tiles.hpp"

namespace lsh {

#define BMAXL 14
#define BMAXF 4

struct BStack {
    const uint16_t *const *frames;  // n device pointers (in device memory)
    int n, L, depth, w, h;
    int lw[BMAXL + 1], lh[BMAXL + 1];
    long long off[BMAXL + 1];  // the fuser's per-plane cell offsets of the levels
    long long slot;            // floats per pyramid slot: 3 * (cells of levels 1..L)
    float *Gg;                 // BMAXF slots; level l (>= 1) of slot q at q * slot + 3 * (off[l] - off[1])
    float *acc, *best, *tops;
    uint16_t *winner;
    long long tsz;
};

struct BPlan {
    int nstacks, rounds, Lmax, chroma, F;
    const int *active;     // [rounds] stacks with a frame in round r (a prefix of the sorted order)
    const long long *pre;  // [rounds][Lmax + 1][nstacks + 1] items before each stack
    Wt9 wt;
};

__device__ __forceinline__ int cdivd(int a, int b) { return (a + b - 1) / b; }

// Control words (device memory, zeroed before each launch).
#define B_BAR 0     // grid barrier arrivals (monotonic)
#define B_ERR 1     // sticky error: a bounded wait ran out
#define B_JOIN 2    // tickets taken at start; bit 31 set once the set is closed
#define B_NPART 3   // the number of participating blocks, published at close
#define B_XCNT 4    // per XCD: the participants' tickets taken there (B_NXCD words)
#define B_XPUB 12   // per XCD: its participants, published at close (B_NXCD words)
#define B_WORDS 20
#define B_NXCD 8
#define B_RUN 256   // items per run dealt to an XCD
#define B_CLOSED 0x80000000u

// The XCD (L2 and its CUs) the block runs on: the XCC_ID hardware register.
__device__ __forceinline__ unsigned b_xcd() {
    return __builtin_amdgcn_s_getreg(20 | (0 << 6) | (3 << 11)) & (B_NXCD - 1);  // hwreg(HW_REG_XCC_ID, 0, 4)
}

// Which blocks take part.  A cooperative launch is accepted up to the occupancy
// the runtime computes from LDS, but the SGPRs of this kernel can allow fewer
// resident blocks per CU, and a block that is not resident cannot start until
// a resident one ends -- a grid barrier would then wait for it forever.  So
// every block takes a ticket as it starts; the first one closes the set once
// every block has joined or none has joined for 20 us (resident blocks start
// within a few us of each other) and publishes the count.  Each participant
// also takes a ticket on its XCD (its rank there); the closer waits for those
// (taken right after the first, by running blocks) and publishes the count per
// XCD with the total.  Blocks starting after the close leave at once.  Every
// participant was running when it joined, so the barriers below only wait for
// running blocks.  sh: the ticket, the count, the XCD, the rank on it, then
// the participants per XCD.
__device__ void b_join(unsigned *ctl, unsigned *sh) {
    if (threadIdx.x == 0) {
        const unsigned t = atomicAdd(&ctl[B_JOIN], 1u);
        unsigned n = 0, x = 0, rx = 0, cnt[B_NXCD] = {};
        if (!(t & B_CLOSED)) {
            x = b_xcd();
            rx = atomicAdd(&ctl[B_XCNT + x], 1u);
            if (t == 0) {
                unsigned seen = 1, spins = 0;
                unsigned long long last = __builtin_amdgcn_s_memrealtime();  // 100 MHz
                for (;;) {
                    const unsigned j = __hip_atomic_load(&ctl[B_JOIN], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    const unsigned long long now = __builtin_amdgcn_s_memrealtime();
                    if (j >= gridDim.x) break;
                    if (j != seen) {
                        seen = j;
                        last = now;
                    } else if (now - last > 2000ull) {
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                    if (++spins > (1u << 22)) break;
                }
                n = atomicOr(&ctl[B_JOIN], B_CLOSED) & ~B_CLOSED;
                // the n participants' XCD tickets
                spins = 0;
                for (;;) {
                    unsigned sum = 0;
                    for (int y = 0; y < B_NXCD; y++)
                        sum += cnt[y] = __hip_atomic_load(&ctl[B_XCNT + y], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if (sum >= n) break;
                    __builtin_amdgcn_s_sleep(1);
                    if (++spins > (1u << 24)) {
                        atomicOr(&ctl[B_ERR], 1u);
                        break;
                    }
                }
                for (int y = 0; y < B_NXCD; y++)
                    __hip_atomic_store(&ctl[B_XPUB + y], cnt[y], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(&ctl[B_NPART], n, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            } else {
                unsigned spins = 0;
                while ((n = __hip_atomic_load(&ctl[B_NPART], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT)) == 0) {
                    __builtin_amdgcn_s_sleep(1);
                    if (++spins > (1u << 24)) {
                        atomicOr(&ctl[B_ERR], 1u);
                        break;
                    }
                }
                for (int y = 0; y < B_NXCD; y++)
                    cnt[y] = __hip_atomic_load(&ctl[B_XPUB + y], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        // a failed join (a wait that ran out) sends every block away
        if (__hip_atomic_load(&ctl[B_ERR], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) n = 0;
        sh[0] = t;
        sh[1] = n;
        sh[2] = x;
        sh[3] = rx;
        for (int y = 0; y < B_NXCD; y++) sh[4 + y] = cnt[y];
    }
    __syncthreads();
}

// a grid barrier over the nb participants: one monotonic counter (the block's
// writes released by its increment, everyone's acquired after the wait),
// bounded waits that raise the error
__device__ void b_barrier(unsigned *ctl, unsigned nb, unsigned &gen) {
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned target = (gen + 1) * nb;
        __hip_atomic_fetch_add(&ctl[B_BAR], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        unsigned spins = 0;
        while (__hip_atomic_load(&ctl[B_BAR], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (__hip_atomic_load(&ctl[B_ERR], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
            __builtin_amdgcn_s_sleep(1);
            if (++spins > (1u << 24)) {
                atomicOr(&ctl[B_ERR], 1u);
                break;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    gen++;
    __syncthreads();
}

// The fold of level l on the 64 x TY tile at (bx, by) for frames rF .. rF+nf-1
// of stack B, in frame order: per frame the stages of the per-frame fold, then
// fuse.rs:429-442's select against the running best in registers (loaded once,
// unless the round holds frame 0, whose energies and L samples are taken as
// they are and whose winner index 0 clears the map).  Ends with a barrier.
template <int R, int TY>
__device__ void fold_group(const BStack &B, int l, int r, int F, int nf, const Wt9 &wt, int chroma, int bx, int by,
                           FoldLds<R, TY> &S) {
    constexpr int NQ = TY / 4;
    const int w = B.lw[l], h = B.lh[l], cw = B.lw[l + 1], ch = B.lh[l + 1];
    const size_t n = (size_t)w * h;
    const int tid = threadIdx.x, tx = tid & 63, x = bx + tx;
    float bst[NQ], a0[NQ], a1[NQ], a2[NQ];
    unsigned wi[NQ];
    unsigned changed = 0, won = 0;  // a bit per row of the thread
    const long long gl = l > 0 ? 3 * (B.off[l] - B.off[1]) : 0, gn = 3 * (B.off[l + 1] - B.off[1]);
    for (int q = 0; q < nf; q++) {
        const unsigned widx = (unsigned)(r * F + q);
        const float *Gq = B.Gg + (size_t)q * B.slot;
        if (l == 0)
            fold_stage<R, TY, false>(RgbSrc{B.frames[widx]}, Gq + gn, w, h, cw, ch, wt, chroma, bx, by, S);
        else
            fold_stage<R, TY, false>(PlaneSrc{Gq + gl, n}, Gq + gn, w, h, cw, ch, wt, chroma, bx, by, S);
        if (x < w) {
#pragma unroll
            for (int k = 0; k < NQ; k++) {
                const int ty = (tid >> 6) + 4 * k, y = by + ty;
                if (y < h) {
                    const float en = fold_energy<R, TY>(wt, tx, ty, S);
                    if (widx == 0) {
                        bst[k] = en;
                        a0[k] = S.Ls[0][ty][tx];
                        a1[k] = S.Ls[1][ty][tx];
                        a2[k] = S.Ls[2][ty][tx];
                        wi[k] = 0;
                        changed |= 1u << k;
                        won |= 1u << k;
                    } else {
                        if (q == 0) bst[k] = B.best[B.off[l] + (size_t)y * w + x];
                        if (en > bst[k]) {  // strict: ties keep the earlier frame (fuse.rs:429, 442)
                            bst[k] = en;
                            a0[k] = S.Ls[0][ty][tx];
                            a1[k] = S.Ls[1][ty][tx];
                            a2[k] = S.Ls[2][ty][tx];
                            wi[k] = widx;
                            changed |= 1u << k;
                            won |= 1u << k;
                        }
                    }
                }
            }
        }
        __syncthreads();  // the LDS tile is free for the next frame / item
    }
    if (!changed) return;
    float *best = B.best + B.off[l], *acc = B.acc + 3 * B.off[l];
    uint16_t *winner = l == B.depth ? B.winner : nullptr;
#pragma unroll
    for (int k = 0; k < NQ; k++) {
        if (!(changed >> k & 1u)) continue;
        const size_t o = (size_t)(by + (tid >> 6) + 4 * k) * w + x;
        best[o] = bst[k];
        acc[o] = a0[k];
        acc[n + o] = a1[k];
        acc[2 * n + o] = a2[k];
        if (winner && (won >> k & 1u)) winner[o] = (uint16_t)wi[k];
    }
}

template <int R, int TY>
__global__ __launch_bounds__(256) void k_batch(const BStack *__restrict__ st, BPlan P, unsigned *ctl) {
    constexpr size_t LDS = sizeof(ReduceLds) > sizeof(FoldLds<R, TY>) ? sizeof(ReduceLds) : sizeof(FoldLds<R, TY>);
    __shared__ __attribute__((aligned(16))) unsigned char lds[LDS];
    __shared__ unsigned ticket[4 + B_NXCD];
    ReduceLds &SR = *reinterpret_cast<ReduceLds *>(lds);
    FoldLds<R, TY> &SF = *reinterpret_cast<FoldLds<R, TY> *>(lds);
    b_join(ctl, ticket);
    const unsigned vb = ticket[0], nb = ticket[1];
    if ((vb & B_CLOSED) || nb == 0 || vb >= nb) return;  // started after the close (or the join failed)
    // A phase's items are cut into runs of B_RUN (a few rows of tiles), dealt
    // to the XCDs in turn; an XCD's participants stride over its runs.  So
    // neighbouring tiles (items in row order), which read the same halo rows,
    // are worked on together by one XCD and meet those rows in its L2, while
    // every XCD gets a like share of every kind of item.
    const unsigned xcd = ticket[2], rx = ticket[3], nx = ticket[4 + xcd];
    unsigned before = 0;  // the participants on the XCDs numbered below this one
    for (unsigned y = 0; y < xcd; y++) before += ticket[4 + y];
    if (nx == 0 || rx >= nx || before + nx > nb) {  // counts that do not add up: the launch fails
        if (threadIdx.x == 0) atomicOr(&ctl[B_ERR], 1u);
        return;
    }
    unsigned gen = 0;
    const int nph = P.Lmax + 1;
    for (int r = 0; r < P.rounds; r++) {
        const int A = P.active[r];
        for (int p = 1; p <= nph; p++) {
            const long long *pre = P.pre + ((size_t)r * nph + (p - 1)) * (P.nstacks + 1);
            const long long total = pre[A];
            if (total == 0) continue;
            int s = 0;
            for (long long u = rx;; u += nx) {
                const long long it = ((u / B_RUN) * B_NXCD + xcd) * B_RUN + u % B_RUN;
                if (it >= total) break;
                while (pre[s + 1] <= it) s++;  // items only grow along the walk
                long long j = it - pre[s];
                const BStack &B = st[s];
                const int nf = min(P.F, B.n - r * P.F), L = B.L;
                if (p <= L) {  // REDUCE: level p from level p - 1 (pyramid.rs:86-119)
                    const int ow = B.lw[p], oh = B.lh[p], tx = cdivd(ow, RED_TX), tr = tx * cdivd(oh, RED_TY);
                    if (j < (long long)nf * tr) {
                        const int q = (int)(j / tr), t = (int)(j - (long long)q * tr);
                        const int w = B.lw[p - 1], h = B.lh[p - 1];
                        float *Gq = B.Gg + (size_t)q * B.slot;
                        float *dst = Gq + 3 * (B.off[p] - B.off[1]);
                        const int ox0 = (t % tx) * RED_TX, oy0 = (t / tx) * RED_TY;
                        if (p == 1) {
                            const RgbSrc src{B.frames[r * P.F + q]};
                            for (int c = 0; c < 3; c++)
                                reduce_tile<false>(src, c, w, h, dst + (size_t)c * ow * oh, ow, oh, ox0, oy0, SR);
                        } else {
                            const PlaneSrc src{Gq + 3 * (B.off[p - 1] - B.off[1]), (size_t)w * h};
                            for (int c = 0; c < 3; c++)
                                reduce_tile<false>(src, c, w, h, dst + (size_t)c * ow * oh, ow, oh, ox0, oy0, SR);
                        }
                        __syncthreads();
                        continue;
                    }
                    j -= (long long)nf * tr;
                }
                if (p >= 2) {  // the fold of level p - 2
                    const int l = p - 2, tx = cdivd(B.lw[l], 64), tf = tx * cdivd(B.lh[l], TY);
                    if (j < tf) {
                        fold_group<R, TY>(B, l, r, P.F, nf, P.wt, P.chroma, (int)(j % tx) * 64, (int)(j / tx) * TY,
                                          SF);
                        continue;
                    }
                    j -= tf;
                }
                // the residual level of frame rF + j into the store (fuse.rs:470-484)
                const float *top = B.Gg + (size_t)j * B.slot + 3 * (B.off[L] - B.off[1]);
                float *dst = B.tops + (size_t)(r * P.F + j) * 3 * B.tsz;
                for (long long i = threadIdx.x; i < 3 * B.tsz; i += 256) dst[i] = top[i];
            }
            b_barrier(ctl, nb, gen);
            if (__hip_atomic_load(&ctl[B_ERR], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
        }
    }
}

__global__ void k_batch_reset(unsigned *ctl) {
    if (threadIdx.x < B_WORDS) ctl[threadIdx.x] = 0;
}

static void binomial_host(size_t r, float *out) {  // fuse.rs:139-149
    size_t n = 2 * r + 1;
    std::vector<uint64_t> row(n, 1);
    for (size_t i = 1; i < n; i++)
        for (size_t j = i - 1; j >= 1; j--) row[j] += row[j - 1];
    uint64_t s = 0;
    for (auto v : row) s += v;
    float sf = (float)s;
    for (size_t i = 0; i < n; i++) out[i] = (float)row[i] / sf;
}

using KFn = void (*)(const BStack *, BPlan, unsigned *);

template <int R>
static KFn pick(int ty) {
    return ty == 8 ? (KFn)k_batch<R, 8> : (KFn)k_batch<R, 16>;
}
static KFn kernel_for(int r, int ty) {
    switch (r) {
    case 0: return pick<0>(ty);
    case 1: return pick<1>(ty);
    case 2: return pick<2>(ty);
    case 3: return pick<3>(ty);
    default: return pick<4>(ty);
    }
}

struct BatchCfg {
    int ty = 16, per_cu = 0, F = 2;  // per_cu 0: as many as fit
};
static std::map<std::tuple<int, int, int, int>, BatchCfg> g_tuned;  // per shape class
static std::mutex g_tune_mu;

}  // namespace lsh

using namespace lsh;

extern "C" {

char *lsh_batch_fuse(lsh_fuser **fusers, const lsh_batch_stack *stacks, uint32_t count, const lsh_batch_options *opt,
                     lsh_batch_report *rep) {
    LSH_API_BEGIN
    LSH_REQUIRE(fusers && stacks && count > 0, "lsh_batch_fuse: null argument");
    lsh_batch_options o{1, 0, 0, 0};
    if (opt) o = *opt;
    LSH_REQUIRE(o.tune <= 2, "batch: tune is 0, 1 or 2");
    LSH_REQUIRE(o.tile_h == 0 || o.tile_h == 8 || o.tile_h == 16, "batch: tile_h is 8 or 16");
    LSH_REQUIRE(o.frames_per_round <= BMAXF, "batch: frames_per_round above 4");
    const lsh_fuser &f0 = *fusers[0];
    const int R = (int)f0.p.energy_radius, chroma = f0.p.use_chroma ? 1 : 0;
    LSH_REQUIRE(R <= 4, "energy_radius above 4 is not supported");
    // the stacks, sorted by frame count (descending): the active set of a round is a prefix
    std::vector<uint32_t> order(count);
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return stacks[a].n > stacks[b].n; });
    std::vector<BStack> hs(count);
    std::vector<DBuf<const uint16_t *>> fptr(count);
    std::vector<DBuf<float>> slots(count);
    int Lmax = 0, nmax = 0;
    double px_frames = 0.0;
    for (uint32_t k = 0; k < count; k++) {
        const lsh_batch_stack &S = stacks[order[k]];
        LSH_REQUIRE(fusers[order[k]], "batch: null fuser");
        lsh_fuser &F = *fusers[order[k]];
        LSH_REQUIRE(S.frames && S.n > 0, "batch: a stack without frames");
        LSH_REQUIRE(S.n <= 65535, "batch: more than 65535 frames in a stack");
        LSH_REQUIRE(F.count == 0 && !F.finished, "batch: the fusers must be fresh");
        LSH_REQUIRE(F.w == S.w && F.h == S.h, "batch: a fuser's size differs from its stack's");
        LSH_REQUIRE(!F.halo_on, "batch: halo control is not supported");
        LSH_REQUIRE((int)F.p.energy_radius == R && (F.p.use_chroma ? 1 : 0) == chroma,
                    "batch: the stacks must share the energy window and the chroma setting");
        LSH_REQUIRE(F.levels <= BMAXL, "batch: pyramid deeper than supported");
        BStack &B = hs[k];
        std::memset(&B, 0, sizeof(B));
        fptr[k].alloc(S.n);
        copy_mem(fptr[k].p, LSH_DEVICE, S.frames, LSH_HOST, S.n * sizeof(const uint16_t *));
        B.frames = fptr[k].p;
        B.n = (int)S.n;
        B.L = (int)F.levels;
        B.depth = (int)F.depth_level;
        B.w = (int)F.w;
        B.h = (int)F.h;
        for (int l = 0; l <= B.L; l++) {
            B.lw[l] = (int)F.lw[l];
            B.lh[l] = (int)F.lh[l];
            B.off[l] = (long long)F.off[l];
        }
        B.slot = 3 * (long long)(F.S - F.off[1]);
        slots[k].alloc((size_t)BMAXF * B.slot);
        B.Gg = slots[k].p;
        B.acc = F.acc.p;
        B.best = F.best.p;
        B.winner = F.winner.p;
        B.tsz = (long long)F.lw[F.levels] * F.lh[F.levels];
        fuser_reserve(&F, S.n);
        B.tops = F.tops.p;
        Lmax = std::max(Lmax, B.L);
        nmax = std::max(nmax, B.n);
        px_frames += (double)F.w * F.h * S.n;
    }
    const int nph = Lmax + 1;
    // per round: the active prefix and the per-phase item prefixes (they depend on F and the tile height)
    auto plan = [&](int ty, int F, int nrounds, std::vector<int> &active, std::vector<long long> &pre) {
        active.assign(nrounds, 0);
        pre.assign((size_t)nrounds * nph * (count + 1), 0);
        for (int r = 0; r < nrounds; r++) {
            while (active[r] < (int)count && hs[active[r]].n > r * F) active[r]++;
            for (int p = 1; p <= nph; p++) {
                long long *pp = pre.data() + ((size_t)r * nph + (p - 1)) * (count + 1);
                for (uint32_t k = 0; k < count; k++) {
                    const BStack &B = hs[k];
                    const long long nf = std::max(0, std::min(F, B.n - r * F));
                    long long items = 0;
                    if (nf > 0) {
                        if (p <= B.L)
                            items += nf * ((B.lw[p] + RED_TX - 1) / RED_TX) * ((B.lh[p] + RED_TY - 1) / RED_TY);
                        if (p >= 2 && p <= B.L + 1)
                            items += (long long)((B.lw[p - 2] + 63) / 64) * ((B.lh[p - 2] + ty - 1) / ty);
                        if (p == B.L + 1) items += nf;
                    }
                    pp[k + 1] = pp[k] + items;
                }
            }
        }
    };
    DBuf<uint8_t> dstacks(count * sizeof(BStack));
    copy_mem(dstacks.p, LSH_DEVICE, hs.data(), LSH_HOST, count * sizeof(BStack));
    DBuf<unsigned> ctl(B_WORDS);
    DBuf<int> dactive;
    DBuf<long long> dpre;
    BPlan P{};
    P.nstacks = (int)count;
    P.Lmax = Lmax;
    P.chroma = chroma;
    binomial_host(R, P.wt.k);
    int dev = 0, cus = 0;
    LSH_CHECK(hipGetDevice(&dev));
    LSH_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    auto max_per_cu = [&](int ty) {
        int occ = 0;
        LSH_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void *)kernel_for(R, ty), 256, 0));
        LSH_REQUIRE(occ > 0, "batch: the persistent kernel cannot be resident");
        return occ;
    };
    hipEvent_t e0, e1;
    LSH_CHECK(hipEventCreate(&e0));
    LSH_CHECK(hipEventCreate(&e1));
    // one launch over the first nrounds rounds with configuration c; returns the device ms
    unsigned resident = 0;  // the blocks that took part in the last launch
    auto run = [&](const BatchCfg &c, int nrounds) -> float {
        const int occ = max_per_cu(c.ty);
        const int per_cu = c.per_cu > 0 ? std::min(c.per_cu, occ) : occ;
        std::vector<int> active;
        std::vector<long long> pre;
        plan(c.ty, c.F, nrounds, active, pre);
        dactive.ensure(active.size());
        copy_mem(dactive.p, LSH_DEVICE, active.data(), LSH_HOST, active.size() * sizeof(int));
        dpre.ensure(pre.size());
        copy_mem(dpre.p, LSH_DEVICE, pre.data(), LSH_HOST, pre.size() * sizeof(long long));
        P.active = dactive.p;
        P.pre = dpre.p;
        P.rounds = nrounds;
        P.F = c.F;
        launch("batch_reset", k_batch_reset, dim3(1), dim3(64), 0, ctl.p);
        const BStack *sp = (const BStack *)dstacks.p;
        unsigned *cp = ctl.p;
        void *args[] = {(void *)&sp, (void *)&P, (void *)&cp};
        const unsigned grid = (unsigned)(per_cu * cus);
        LSH_CHECK(hipEventRecord(e0, stream()));
        launch_coop("batch_fuse", 0.0, (const void *)kernel_for(R, c.ty), dim3(grid), dim3(256), args, 0);
        LSH_CHECK(hipEventRecord(e1, stream()));
        LSH_CHECK(hipEventSynchronize(e1));
        unsigned words[4] = {0, 0, 0, 0};
        copy_mem(words, LSH_HOST, ctl.p, LSH_DEVICE, sizeof(words));
        LSH_REQUIRE(words[B_ERR] == 0, "batch: a grid barrier timed out");
        LSH_REQUIRE(words[B_NPART] > 0, "batch: no block took part");
        resident = words[B_NPART];
        float ms = 0.0f;
        LSH_CHECK(hipEventElapsedTime(&ms, e0, e1));
        return ms;
    };
    // the shape class: window radius, depth, log2 of the pixel-frames, log2 of the stack count
    const auto key = std::make_tuple(R, Lmax, (int)std::log2(std::max(px_frames, 1.0)), (int)std::log2((double)count));
    BatchCfg best;
    if (o.tile_h) best.ty = (int)o.tile_h;
    if (o.blocks_per_cu) best.per_cu = (int)o.blocks_per_cu;
    if (o.frames_per_round) best.F = (int)o.frames_per_round;
    bool cached = false;
    std::vector<std::pair<BatchCfg, float>> tried;
    try {
        if (o.tune == 1) {
            std::lock_guard<std::mutex> lk(g_tune_mu);
            auto it = g_tuned.find(key);
            if (it != g_tuned.end()) {
                best = it->second;
                cached = true;
            }
        }
        if (o.tune && !cached) {
            // round 0 only writes (its first frame sets every pixel), so each candidate can replay it;
            // compare device time per pixel-frame of round 0
            auto trial = [&](const BatchCfg &c) {
                double pf = 0.0;
                for (const BStack &B : hs) pf += (double)B.w * B.h * std::min(c.F, B.n);
                const float us = (float)(1e3 * run(c, 1) / (pf / 1e6));
                tried.push_back({c, us});
                return us;
            };
            float bus = 1e30f;
            for (int ty : {8, 16}) {  // tile height x blocks per CU, one frame per round
                const int occ = max_per_cu(ty);
                std::vector<int> pcs{2, occ};
                if (occ <= 2) pcs = {occ};
                for (int pc : pcs) {
                    BatchCfg c{ty, pc, 1};
                    const float us = trial(c);
                    if (us < bus) bus = us, best = c;
                }
            }
            const BatchCfg base = best;
            for (int F : {2, 4}) {  // then the frames per round
                BatchCfg c = base;
                c.F = F;
                const float us = trial(c);
                if (us < bus) bus = us, best = c;
            }
            std::lock_guard<std::mutex> lk(g_tune_mu);
            g_tuned[key] = best;
        }
    } catch (...) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    float ms = 0.0f;
    const int rounds = (nmax + best.F - 1) / best.F;
    try {
        ms = run(best, rounds);
    } catch (...) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    for (uint32_t k = 0; k < count; k++) fusers[order[k]]->count = stacks[order[k]].n;
    if (rep) {
        std::memset(rep, 0, sizeof(*rep));
        rep->ms = ms;
        rep->mpix_frames = px_frames / 1e6;
        rep->rounds = (uint32_t)rounds;
        rep->max_levels = (uint32_t)Lmax;
        rep->tile_h = (uint32_t)best.ty;
        rep->blocks_per_cu = (uint32_t)(best.per_cu > 0 ? std::min(best.per_cu, max_per_cu(best.ty))
                                                        : max_per_cu(best.ty));
        rep->resident_blocks = resident;
        rep->frames_per_round = (uint32_t)best.F;
        rep->from_cache = cached ? 1u : 0u;
        rep->ntried = (uint32_t)std::min<size_t>(tried.size(), 8);
        for (uint32_t q = 0; q < rep->ntried; q++) {
            const BatchCfg &c = tried[q].first;
            rep->tried_cfg[q] = (uint32_t)c.ty | (uint32_t)c.per_cu << 8 | (uint32_t)c.F << 16;
            rep->tried_us_per_mpf[q] = tried[q].second;
        }
    }
    LSH_API_END
}

}  // extern "C"
