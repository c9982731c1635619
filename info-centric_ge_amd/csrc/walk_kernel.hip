// walk_kernel.hip - the persistent walk kernel and the corpus compaction.
//
// Replaces the per-walker step loop of WalkEngine::internal_walk_epoch
// (include/walk.hpp:848-1122) for one GPU holding the whole graph:
//   - one walker per lane; a wavefront refills finished lanes from a global
//     work cursor with one atomic per refill (ballot + popcount + mbcnt), so
//     walks of uneven length run in one launch;
//   - each step draws Philox(seed; step, attempt, walker, round) and takes an
//     alias step: column j = hi32(r0 * deg), keep column j when r1 < thresh_j
//     else take alias(j) - one 16-byte gather per step after the vertex record;
//   - the information-centric stop rule (include/walk.hpp:1045-1062) runs in
//     FP64 on the lane: prior visits f of the entered vertex come from the
//     walker's history (first 32 vertices in LDS, the rest from its path in
//     global memory);
//   - a walker whose start vertex carries a learned length budget runs
//     exactly that many vertices and keeps no entropy state;
//   - reference special cases kept: degree 0 ends a walk (:964-969), a
//     degree-2 vertex whose first neighbour is itself ends it (:977-979), a
//     self loop drawn at step 0 is redrawn (:981-983, at most 64 times).
// Walk i of the round writes its path into a fixed slot of `slot` vertices
// (walker-major), continuing in an overflow pool if it grows past the slot;
// the first 32 vertices are written when the walk ends, from LDS.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include "felog_dev.h"
#include "felog_internal.h"

namespace felog {

namespace {

struct WalkArgs {
    const VertexRec *__restrict__ vwalk;  // start vertices' walk records
    const EdgeRec *__restrict__ erec;
    const uint32_t *__restrict__ active;
    const RoundState *rs;  // round and walkers started (device)
    uint32_t V;
    uint32_t *__restrict__ slots;
    uint32_t slot;
    uint32_t *__restrict__ lens;
    uint64_t *__restrict__ ovf_base;
    uint32_t *__restrict__ ovf;
    uint64_t ovf_cap;
    unsigned long long *counters;
    const double *__restrict__ T;
    const double *__restrict__ dT;
    const uint16_t *__restrict__ fixed_len;  // per start vertex, 0: stop rule (null: none)
    const float *__restrict__ cn_data;       // edge data: the reference's transition (null: alias)
    uint32_t min_L, max_len;
    uint64_t seed;
};

constexpr int kWalkBlock = 256;
constexpr uint64_t kNoOverflow = ~0ull;

__device__ __forceinline__ uint32_t path_at(const WalkArgs &a, uint32_t w, uint64_t ob, uint32_t k) {
    return k < a.slot ? a.slots[(uint64_t)w * a.slot + k] : a.ovf[ob + (k - a.slot)];
}

// Writes the first min(len, kHistLds) vertices of a finished walk from the
// lane's LDS history to its slot (16-byte stores when the slot is a multiple
// of four vertices) and, past the slot, to its overflow run.  A lane writes
// its whole prefix at once, so the slot's lines are filled in one burst
// instead of one 4-byte store per step into a different line each time.
__device__ __forceinline__ void flush_path(const WalkArgs &a, const uint32_t *H, uint32_t w, uint64_t ob,
                                           uint32_t len) {
    const uint32_t nh = len < (uint32_t)kHistLds ? len : (uint32_t)kHistLds;
    const uint32_t ns = nh < a.slot ? nh : a.slot;
    uint32_t *dst = a.slots + (uint64_t)w * a.slot;
    if ((a.slot & 3u) == 0) {
        // k + 3 < slot (slot a multiple of 4) and k + 3 < kHistLds (k <= 28)
        for (uint32_t k = 0; k < ns; k += 4) {
            uint4 q;
            q.x = H[k * 64];
            q.y = H[(k + 1) * 64];
            q.z = H[(k + 2) * 64];
            q.w = H[(k + 3) * 64];
            *reinterpret_cast<uint4 *>(dst + k) = q;
        }
    } else {
        for (uint32_t k = 0; k < ns; ++k) dst[k] = H[k * 64];
    }
    for (uint32_t k = a.slot; k < nh; ++k) a.ovf[ob + (k - a.slot)] = H[k * 64];
}

__global__ __launch_bounds__(kWalkBlock) void k_walk(WalkArgs a) {
    __shared__ uint32_t hist[(kWalkBlock / 64) * kHistLds * 64];
    const uint32_t lane = threadIdx.x & 63;
    uint32_t *H = hist + (threadIdx.x >> 6) * (kHistLds * 64) + lane;
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));

    bool active = false;
    bool more = true;  // wave-uniform: the work cursor may still hold walkers
    uint32_t w = 0, v = 0, s = 0, attempt = 0, lim = 0;
    VertexRec cur{0, 0};  // walk record of v (degree with kSelfFirst)
    uint64_t ob = kNoOverflow;
    WalkStats st;
    walk_stats_init(st);
    unsigned long long steps = 0;
    uint32_t maxlen = 0;
    const uint32_t n_active = a.rs->n_active, round = a.rs->round;

    while (true) {
        const uint64_t need = __ballot(!active);
        if (more && need) {
            const uint32_t cnt = (uint32_t)__popcll(need);
            const uint32_t leader = (uint32_t)__ffsll((long long)need) - 1;
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd((unsigned int *)&a.counters[C_WORK], cnt);
            base = __shfl(base, (int)leader);
            if ((uint64_t)base + cnt >= n_active) more = false;
            if (!active) {
                const uint32_t idx = base + (uint32_t)__popcll(need & lt_mask);
                if (idx < n_active) {
                    w = a.active[idx];
                    v = w % a.V;
                    cur = a.vwalk[v];
                    lim = a.fixed_len ? a.fixed_len[v] : 0u;
                    s = 0;
                    attempt = 0;
                    ob = kNoOverflow;
                    walk_stats_init(st);
                    H[0] = v;
                    active = true;
                }
            }
        }
        if (__ballot(active) == 0) break;
        if (active) {
        bool stop = false;
        if (s + 1 >= a.max_len || (lim && s + 1 >= lim)) {
            stop = true;
        } else {
            const uint32_t deg = cur.deg & ~kSelfFirst;
            if (deg == 0 || (cur.deg & kSelfFirst)) {
                stop = true;  // degree 0, or degree 2 with itself first
            } else {
                const u32x4 r = draw(a.seed, s, DOM_WALK | attempt, w, round);
                const uint32_t j = alias_bucket(r.x, deg);
                const EdgeRec e = a.erec[cur.begin + j];
                // alias step, or the reference's uniform neighbour (cn_data)
                const bool first = a.cn_data || r.y < e.thresh;
                const uint32_t x = first ? e.col : e.alias_col;
                if (s == 0 && x == v) {
                    if (++attempt >= kMaxStartAttempts) stop = true;
                } else {
                    attempt = 0;
                    // append vertex y at position s + 1: its prior visits on
                    // this walk (stop-rule walks only), storage, the stop rule
                    auto append = [&](uint32_t y) -> bool {
                        uint32_t f = 0;
                        if (!lim) {
                            const uint32_t nl = s + 1 < (uint32_t)kHistLds ? s + 1 : (uint32_t)kHistLds;
                            for (uint32_t k = 0; k < nl; ++k) f += (H[k * 64] == y) ? 1u : 0u;
                            for (uint32_t k = kHistLds; k <= s; ++k) f += (path_at(a, w, ob, k) == y) ? 1u : 0u;
                        }
                        const uint32_t ns = s + 1;
                        if (ns >= a.slot && ob == kNoOverflow) {
                            const uint64_t need_n = a.max_len - a.slot;
                            const uint64_t got = atomicAdd(&a.counters[C_OVF_NEXT], (unsigned long long)need_n);
                            if (got + need_n > a.ovf_cap) {
                                atomicAdd(&a.counters[C_OVF_FAIL], 1ull);
                                return true;  // truncated; the host reports FELOG_ERR_CAPACITY
                            }
                            ob = got;
                        }
                        // the first kHistLds vertices stay in LDS until the
                        // walk ends (flush_path); later ones go straight out
                        if (ns < (uint32_t)kHistLds) {
                            H[ns * 64] = y;
                        } else if (ns < a.slot) {
                            a.slots[(uint64_t)w * a.slot + ns] = y;
                        } else {
                            a.ovf[ob + (ns - a.slot)] = y;
                        }
                        s = ns;
                        return lim ? s + 1 >= lim : walk_stats_step(st, a.T, a.dT, f, s + 1, s, a.min_L);
                    };
                    bool trace = false;
                    if (a.cn_data)
                        trace = cn_trace(deg, e.col_vr.deg & ~kSelfFirst, a.cn_data[cur.begin + j],
                                         uniform53(r.y, r.z));
                    stop = append(x);
                    if (trace) {
                        // include/walk.hpp:1063-1094: back to v, which stays current
                        if (!stop) stop = s + 1 >= a.max_len || append(v);
                    } else {
                        v = x;
                        cur = first ? e.col_vr : e.alias_vr;
                    }
                }
            }
        }
        if (stop) {
            flush_path(a, H, w, ob, s + 1);
            a.lens[w] = s + 1;
            a.ovf_base[w] = ob;
            steps += s;
            maxlen = maxlen > s + 1 ? maxlen : s + 1;
            active = false;
        }
        }
    }
    // wave totals
    for (int o = 32; o > 0; o >>= 1) {
        steps += __shfl_xor(steps, o);
        const uint32_t m2 = __shfl_xor(maxlen, o);
        maxlen = maxlen > m2 ? maxlen : m2;
    }
    // one atomic per workgroup and counter: atomics on one address
    // serialise, and a wavefront's each would queue thousands of them at the
    // end of the launch
    // (the partials go to the start of each wavefront's own history, which
    // it no longer reads: more LDS would cost the launch a workgroup per CU)
    uint32_t *part = hist + (threadIdx.x >> 6) * (kHistLds * 64);
    if (lane == 0) {
        part[0] = (uint32_t)steps;
        part[1] = (uint32_t)(steps >> 32);
        part[2] = maxlen;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < kWalkBlock / 64; ++i) {
            const uint32_t *q = hist + i * (kHistLds * 64);
            steps += ((unsigned long long)q[1] << 32) | q[0];
            maxlen = maxlen > q[2] ? maxlen : q[2];
        }
        atomicAdd(&a.counters[C_STEPS], steps);
        atomicMax(&a.counters[C_MAXLEN], (unsigned long long)maxlen);
    }
}

__global__ void k_walk_ok(const uint8_t *__restrict__ flags, uint32_t walkers, uint32_t V,
                          uint32_t *__restrict__ ok) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < walkers; i += stride)
        ok[i] = flags[i % V] ? 0 : 1;
}

// After the inclusive sum of the start flags (sel[W-1] = walkers started):
// walker i starts when the sum steps at i, and goes to position sel[i] - 1
// of the active list, which keeps walker-id order.
__global__ void k_select(const uint32_t *__restrict__ sel, uint32_t W, uint32_t *__restrict__ active,
                         RoundState *__restrict__ rs) {
    const uint32_t stride = gridDim.x * blockDim.x;
    if (blockIdx.x == 0 && threadIdx.x == 0) rs->n_active = W ? sel[W - 1] : 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < W; i += stride) {
        const uint32_t p = sel[i], q = i ? sel[i - 1] : 0;
        if (p != q) active[q] = i;
    }
}

// Lengths of this round's walks into off[1 .. W] (0 past the walkers
// started) and off[0] = 0, so the offsets scan runs over a fixed W entries.
__global__ void k_gather_lens(const uint32_t *__restrict__ active, const RoundState *__restrict__ rs, uint32_t W,
                              const uint32_t *__restrict__ lens, uint64_t *__restrict__ off) {
    const uint32_t n = rs->n_active;
    const uint32_t stride = gridDim.x * blockDim.x;
    if (blockIdx.x == 0 && threadIdx.x == 0) off[0] = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < W; i += stride)
        off[1 + i] = i < n ? lens[active[i]] : 0;
}

// Sixteen lanes per walk (four walks per wavefront, so four walks' loads
// are in flight at once): lanes copy consecutive vertices of the path.
// Thread 0 also records the round's token count.
__global__ void k_compact(const uint32_t *__restrict__ active, RoundState *__restrict__ rs,
                          uint64_t *__restrict__ off, const uint32_t *__restrict__ slots,
                          uint32_t slot, const uint64_t *__restrict__ ovf_base,
                          const uint32_t *__restrict__ ovf, uint32_t *__restrict__ tok,
                          uint32_t *__restrict__ sent_id) {
    const uint32_t n = rs->n_active;
    if (blockIdx.x == 0 && threadIdx.x == 0) rs->n_tokens = off[n];
    const uint32_t l = threadIdx.x & 15;
    const uint32_t group = (blockIdx.x * blockDim.x + threadIdx.x) >> 4;
    const uint32_t ngroups = (gridDim.x * blockDim.x) >> 4;
    for (uint32_t i = group; i < n; i += ngroups) {
        const uint32_t w = active[i];
        const uint64_t b = off[i], len = off[i + 1] - b;
        if (l == 0) sent_id[i] = w;
        for (uint64_t k = l; k < len; k += 16)
            tok[b + k] = k < slot ? slots[(uint64_t)w * slot + k] : ovf[ovf_base[w] + (k - slot)];
    }
}

// Learning rate of round r: ExponentialDecayLR(base, decay) with one
// get_lr() call per round (include/lr_scheduler.hpp:23-31,
// src/word2vec.cu:1958,2001).  Eager and captured rounds both take it from
// here, so they train with the same bits.
__device__ __forceinline__ float scheduled_alpha(float base, float decay, uint32_t r) {
    return base * expf(-decay * (float)(r - 1));
}

// Eager calls: the round, its learning rate (alpha < 0: the schedule's) and,
// unless n_sent is ~0u, the sentence count of the corpus to train.
__global__ void k_round_set(RoundState *rs, uint32_t round, float alpha, float base, float decay, uint32_t n_sent) {
    rs->round = round;
    rs->alpha = alpha < 0.f ? scheduled_alpha(base, decay, round) : alpha;
    if (n_sent != ~0u) rs->n_active = n_sent;
}

// Captured rounds: advance to the next round and its learning rate, and
// clear the per-round counters (a kernel, not a memset node: such nodes are
// avoided in captured graphs).
__global__ void k_round_next(RoundState *rs, float base, float decay, unsigned long long *counters, int n) {
    if (threadIdx.x == 0) {
        const uint32_t r = rs->round + 1;
        rs->round = r;
        rs->alpha = scheduled_alpha(base, decay, r);
        rs->evaluated = 0;
    }
    for (int i = threadIdx.x; i < n; i += blockDim.x) counters[i] = 0;
}

// Captured rounds: the plain rule's stop decision after the evaluation
// (src/word2vec.cu:2073-2078: stop once the unconverged count shrinks by
// less than 30% in a round), in the host's float arithmetic.
__global__ void k_round_end(RoundState *rs, const unsigned long long *counters) {
    RoundState &s = *rs;
    if (!round_evaluates(s)) return;
    const uint32_t eva = (uint32_t)counters[C_EVALNUM];
    const float ratio = s.last_eva ? (float)eva / (float)s.last_eva : 0.f;
    if (s.last_eva != 0 && ratio > 0.7f) s.stop = 1;
    s.last_eva = eva;
    s.evaluated = 1;
}

inline unsigned grid_for(uint64_t n, unsigned block, unsigned cap) {
    uint64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (unsigned)g;
}

}  // namespace

#define HIP_TRY(x)                              \
    do {                                        \
        hipError_t e_ = (x);                    \
        if (e_ != hipSuccess) return e_;        \
    } while (0)

// Scratch of the per-round library primitives (walker select, offsets
// scan over W entries) and the fixed-capacity corpus buffers, sized once so
// a round allocates nothing and can be captured.
hipError_t round_buffers(felog_ctx *c) {
    const uint32_t W = c->walkers;
    size_t t1 = 0, t2 = 0;
    HIP_TRY(hipcub::DeviceScan::InclusiveSum(nullptr, t1, c->d_active, c->d_sel, W, c->stream));
    HIP_TRY(hipcub::DeviceScan::InclusiveSum(nullptr, t2, (uint64_t *)nullptr, (uint64_t *)nullptr, (uint64_t)W,
                                             c->stream));
    const size_t tmp = t1 > t2 ? t1 : t2;
    if (tmp > c->tmp_bytes) {
        dfree(c->d_tmp);
        HIP_TRY(dalloc(c, &c->d_tmp, tmp));
        c->tmp_bytes = tmp;
    }
    if ((uint64_t)W + 1 > c->off_cap) {
        dfree(c->d_off);
        dfree(c->d_sent_id);
        HIP_TRY(dalloc(c, &c->d_off, ((uint64_t)W + 1) * 8));
        HIP_TRY(dalloc(c, &c->d_sent_id, ((uint64_t)W + 1) * 4));
        c->off_cap = (uint64_t)W + 1;
    }
    // every token of a round sits in a walker's slot or the overflow pool
    const uint64_t tok = (uint64_t)W * c->slot + c->ovf_cap;
    if (tok > c->tok_cap) {
        dfree(c->d_tok);
        HIP_TRY(dalloc(c, &c->d_tok, tok * 4));
        c->tok_cap = tok;
    }
    return hipSuccess;
}

// Walker select (walkers whose start vertex is unconverged or, under a
// learned budget, whose count covers the round; walker-id order,
// include/walk.hpp:561-576) and the persistent walk, without a host wait.
// The select is a scan of the start flags (held in the active list's
// buffer) and a scatter: kernels only, so it can be captured (a library
// select clears its counter with a memset).
hipError_t walk_enqueue(felog_ctx *c) {
    const uint32_t W = c->walkers;
    if (c->budget_on) {
        HIP_TRY(budget_walk_ok(c, c->d_active));
    } else {
        k_walk_ok<<<grid_for(W, 256, 65535), 256, 0, c->stream>>>(c->d_flags, W, c->V, c->d_active);
        HIP_TRY(hipGetLastError());
    }
    size_t tmp = c->tmp_bytes;
    HIP_TRY(hipcub::DeviceScan::InclusiveSum(c->d_tmp, tmp, c->d_active, c->d_sel, W, c->stream));
    k_select<<<grid_for(W, 256, 65535), 256, 0, c->stream>>>(c->d_sel, W, c->d_active, c->d_rs);
    HIP_TRY(hipGetLastError());
    WalkArgs a;
    a.vwalk = c->d_vwalk;
    a.erec = c->d_erec;
    a.active = c->d_active;
    a.rs = c->d_rs;
    a.V = c->V;
    a.slots = c->d_slots;
    a.slot = c->slot;
    a.lens = c->d_lens;
    a.ovf_base = c->d_ovf_base;
    a.ovf = c->d_ovf;
    a.ovf_cap = c->ovf_cap;
    a.counters = (unsigned long long *)c->d_counters;
    a.T = c->d_T;
    a.dT = c->d_dT;
    a.fixed_len = c->budget_on ? c->d_budget_len : nullptr;
    // the reference's transition reads common-neighbour counts from the edge
    // data, or from the counts felog_use_common_neighbours computed
    a.cn_data = c->opt.compat_walk ? (c->d_cn ? c->d_cn : c->d_w) : nullptr;
    a.min_L = c->opt.min_L;
    a.max_len = c->opt.max_walk_len;
    a.seed = c->opt.seed;
    int per_cu = 0;
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_walk, kWalkBlock, 0));
    if (per_cu < 1) per_cu = 1;
    uint64_t blocks = (uint64_t)per_cu * c->cu_count;
    const uint64_t need_blocks = ((uint64_t)W + kWalkBlock - 1) / kWalkBlock;
    if (blocks > need_blocks) blocks = need_blocks;
    k_walk<<<(unsigned)blocks, kWalkBlock, 0, c->stream>>>(a);
    return hipGetLastError();
}

// Offsets (scan over W lengths) and the flat corpus, without a host wait.
hipError_t compact_enqueue(felog_ctx *c) {
    const uint32_t W = c->walkers;
    k_gather_lens<<<grid_for(W, 256, 65535), 256, 0, c->stream>>>(c->d_active, c->d_rs, W, c->d_lens, c->d_off);
    HIP_TRY(hipGetLastError());
    size_t tmp = c->tmp_bytes;
    HIP_TRY(hipcub::DeviceScan::InclusiveSum(c->d_tmp, tmp, c->d_off + 1, c->d_off + 1, (uint64_t)W, c->stream));
    k_compact<<<grid_for((uint64_t)W * 16, 256, 1u << 16), 256, 0, c->stream>>>(
        c->d_active, c->d_rs, c->d_off, c->d_slots, c->slot, c->d_ovf_base, c->d_ovf, c->d_tok, c->d_sent_id);
    return hipGetLastError();
}

hipError_t round_set(felog_ctx *c, uint32_t round, float alpha, uint32_t n_sent) {
    k_round_set<<<1, 1, 0, c->stream>>>(c->d_rs, round, alpha, alpha_base(c->opt), kAlphaDecay, n_sent);
    return hipGetLastError();
}

hipError_t round_next(felog_ctx *c) {
    k_round_next<<<1, 64, 0, c->stream>>>(c->d_rs, alpha_base(c->opt), kAlphaDecay,
                                          (unsigned long long *)c->d_counters, kNumCounters);
    return hipGetLastError();
}

hipError_t round_end(felog_ctx *c) {
    k_round_end<<<1, 1, 0, c->stream>>>(c->d_rs, (const unsigned long long *)c->d_counters);
    return hipGetLastError();
}

// Eager round: walk, then read the walker count and the counters.
hipError_t run_walks(felog_ctx *c, float *walk_ms) {
    HIP_TRY(hipMemsetAsync(c->d_counters, 0, kNumCounters * 8, c->stream));
    HIP_TRY(hipEventRecord(c->ev0, c->stream));
    HIP_TRY(walk_enqueue(c));
    HIP_TRY(hipEventRecord(c->ev1, c->stream));
    HIP_TRY(hipEventSynchronize(c->ev1));
    HIP_TRY(hipEventElapsedTime(walk_ms, c->ev0, c->ev1));
    HIP_TRY(hipMemcpy(c->h_counters, c->d_counters, kNumCounters * 8, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(c->h_rs, c->d_rs, sizeof(RoundState), hipMemcpyDeviceToHost));
    c->corpus.n_walks = c->h_rs->n_active;
    return hipSuccess;
}

hipError_t compact_corpus(felog_ctx *c, float *ms) {
    HIP_TRY(hipEventRecord(c->ev0, c->stream));
    HIP_TRY(compact_enqueue(c));
    HIP_TRY(hipEventRecord(c->ev1, c->stream));
    HIP_TRY(hipEventSynchronize(c->ev1));
    HIP_TRY(hipEventElapsedTime(ms, c->ev0, c->ev1));
    HIP_TRY(hipMemcpy(c->h_rs, c->d_rs, sizeof(RoundState), hipMemcpyDeviceToHost));
    c->corpus.n_tokens = c->h_rs->n_tokens;
    return hipSuccess;
}

}  // namespace felog
