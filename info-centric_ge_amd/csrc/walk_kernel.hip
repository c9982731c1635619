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
// (walker-major), continuing in an overflow pool if it grows past the slot.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include "felog_dev.h"
#include "felog_internal.h"

namespace felog {

namespace {

struct WalkArgs {
    const VertexRec *__restrict__ vrec;
    const EdgeRec *__restrict__ erec;
    const uint32_t *__restrict__ active;
    uint32_t n_active;
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
    uint32_t min_L, max_len, round;
    uint64_t seed;
};

constexpr int kWalkBlock = 256;
constexpr uint64_t kNoOverflow = ~0ull;

__device__ __forceinline__ uint32_t path_at(const WalkArgs &a, uint32_t w, uint64_t ob, uint32_t k) {
    return k < a.slot ? a.slots[(uint64_t)w * a.slot + k] : a.ovf[ob + (k - a.slot)];
}

__global__ __launch_bounds__(kWalkBlock) void k_walk(WalkArgs a) {
    __shared__ uint32_t hist[(kWalkBlock / 64) * kHistLds * 64];
    const uint32_t lane = threadIdx.x & 63;
    uint32_t *H = hist + (threadIdx.x >> 6) * (kHistLds * 64) + lane;
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));

    bool active = false;
    bool more = true;  // wave-uniform: the work cursor may still hold walkers
    uint32_t w = 0, v = 0, s = 0, attempt = 0, lim = 0;
    uint64_t ob = kNoOverflow;
    WalkStats st;
    walk_stats_init(st);
    unsigned long long steps = 0;
    uint32_t maxlen = 0;

    while (true) {
        const uint64_t need = __ballot(!active);
        if (more && need) {
            const uint32_t cnt = (uint32_t)__popcll(need);
            const uint32_t leader = (uint32_t)__ffsll((long long)need) - 1;
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd((unsigned int *)&a.counters[C_WORK], cnt);
            base = __shfl(base, (int)leader);
            if ((uint64_t)base + cnt >= a.n_active) more = false;
            if (!active) {
                const uint32_t idx = base + (uint32_t)__popcll(need & lt_mask);
                if (idx < a.n_active) {
                    w = a.active[idx];
                    v = w % a.V;
                    lim = a.fixed_len ? a.fixed_len[v] : 0u;
                    s = 0;
                    attempt = 0;
                    ob = kNoOverflow;
                    walk_stats_init(st);
                    H[0] = v;
                    a.slots[(uint64_t)w * a.slot] = v;
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
            const VertexRec vr = a.vrec[v];
            if (vr.deg == 0) {
                stop = true;
            } else if (vr.deg == 2 && a.erec[vr.begin].col == v) {
                stop = true;
            } else {
                const u32x4 r = draw(a.seed, s, DOM_WALK | attempt, w, a.round);
                const uint32_t j = alias_bucket(r.x, vr.deg);
                const EdgeRec e = a.erec[vr.begin + j];
                const uint32_t x = (r.y < e.thresh) ? e.col : e.alias_col;
                if (s == 0 && x == v) {
                    if (++attempt >= kMaxStartAttempts) stop = true;
                } else {
                    attempt = 0;
                    // prior visits of x on this walk (stop-rule walks only)
                    uint32_t f = 0;
                    if (!lim) {
                        const uint32_t nl = s + 1 < (uint32_t)kHistLds ? s + 1 : (uint32_t)kHistLds;
                        for (uint32_t k = 0; k < nl; ++k) f += (H[k * 64] == x) ? 1u : 0u;
                        for (uint32_t k = kHistLds; k <= s; ++k) f += (path_at(a, w, ob, k) == x) ? 1u : 0u;
                    }
                    const uint32_t ns = s + 1;
                    bool stored = true;
                    if (ns < a.slot) {
                        a.slots[(uint64_t)w * a.slot + ns] = x;
                    } else {
                        if (ob == kNoOverflow) {
                            const uint64_t need_n = a.max_len - a.slot;
                            const uint64_t got = atomicAdd(&a.counters[C_OVF_NEXT], (unsigned long long)need_n);
                            if (got + need_n <= a.ovf_cap) ob = got;
                            else { atomicAdd(&a.counters[C_OVF_FAIL], 1ull); stored = false; }
                        }
                        if (stored) a.ovf[ob + (ns - a.slot)] = x;
                    }
                    if (!stored) {
                        stop = true;  // truncated; the host reports FELOG_ERR_CAPACITY
                    } else {
                        if (ns < (uint32_t)kHistLds) H[ns * 64] = x;
                        s = ns;
                        stop = lim ? s + 1 >= lim : walk_stats_step(st, a.T, a.dT, f, s + 1, s, a.min_L);
                        v = x;
                    }
                }
            }
        }
        if (stop) {
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
    if (lane == 0) {
        atomicAdd(&a.counters[C_STEPS], steps);
        atomicMax(&a.counters[C_MAXLEN], (unsigned long long)maxlen);
    }
}

__global__ void k_walk_ok(const uint8_t *__restrict__ flags, uint32_t walkers, uint32_t V,
                          uint8_t *__restrict__ ok) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < walkers; i += stride)
        ok[i] = flags[i % V] ? 0 : 1;
}

__global__ void k_gather_lens(const uint32_t *__restrict__ active, uint32_t n,
                              const uint32_t *__restrict__ lens, uint64_t *__restrict__ out) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        out[i] = lens[active[i]];
}

// One wavefront per walk: lanes copy consecutive vertices of the path.
__global__ void k_compact(const uint32_t *__restrict__ active, uint32_t n,
                          const uint64_t *__restrict__ off, const uint32_t *__restrict__ slots,
                          uint32_t slot, const uint64_t *__restrict__ ovf_base,
                          const uint32_t *__restrict__ ovf, uint32_t *__restrict__ tok,
                          uint32_t *__restrict__ sent_id) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
    for (uint32_t i = wave; i < n; i += nwaves) {
        const uint32_t w = active[i];
        const uint64_t b = off[i], len = off[i + 1] - b;
        if (lane == 0) sent_id[i] = w;
        const uint64_t ob = ovf_base[w];
        for (uint64_t k = lane; k < len; k += 64)
            tok[b + k] = k < slot ? slots[(uint64_t)w * slot + k] : ovf[ob + (k - slot)];
    }
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

// Builds the active list (walkers whose start vertex is unconverged, in
// walker-id order, include/walk.hpp:561-576) and runs the persistent walk.
hipError_t run_walks(felog_ctx *c, uint32_t round, uint32_t n_active_unused, float *walk_ms) {
    (void)n_active_unused;
    const uint32_t W = c->walkers;
    if (c->budget_on) {
        HIP_TRY(budget_walk_ok(c, round));
    } else {
        k_walk_ok<<<grid_for(W, 256, 65535), 256, 0, c->stream>>>(c->d_flags, W, c->V, c->d_walk_ok);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipMemsetAsync(c->d_counters, 0, kNumCounters * 8, c->stream));
    size_t tmp = 0;
    hipcub::CountingInputIterator<uint32_t> ids(0);
    unsigned long long *d_nsel = (unsigned long long *)(c->d_counters + C_EVALCONV + 2);
    HIP_TRY(hipcub::DeviceSelect::Flagged(nullptr, tmp, ids, c->d_walk_ok, c->d_active, d_nsel, W,
                                          c->stream));
    if (tmp > c->tmp_bytes) {
        dfree(c->d_tmp);
        HIP_TRY(hipMalloc(&c->d_tmp, tmp));
        c->tmp_bytes = tmp;
    }
    HIP_TRY(hipcub::DeviceSelect::Flagged(c->d_tmp, tmp, ids, c->d_walk_ok, c->d_active, d_nsel, W,
                                          c->stream));
    HIP_TRY(hipMemcpyAsync(c->h_counters, c->d_counters, kNumCounters * 8, hipMemcpyDeviceToHost,
                           c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    const uint32_t n_active = (uint32_t)c->h_counters[C_EVALCONV + 2];
    c->corpus.n_walks = n_active;
    if (n_active == 0) {
        *walk_ms = 0.f;
        return hipSuccess;
    }

    WalkArgs a;
    a.vrec = c->d_vrec;
    a.erec = c->d_erec;
    a.active = c->d_active;
    a.n_active = n_active;
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
    a.min_L = c->opt.min_L;
    a.max_len = c->opt.max_walk_len;
    a.round = round;
    a.seed = c->opt.seed;

    int per_cu = 0;
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_walk, kWalkBlock, 0));
    if (per_cu < 1) per_cu = 1;
    uint64_t blocks = (uint64_t)per_cu * c->cu_count;
    const uint64_t need_blocks = ((uint64_t)n_active + kWalkBlock - 1) / kWalkBlock;
    if (blocks > need_blocks) blocks = need_blocks;
    HIP_TRY(hipEventRecord(c->ev0, c->stream));
    k_walk<<<(unsigned)blocks, kWalkBlock, 0, c->stream>>>(a);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(c->ev1, c->stream));
    HIP_TRY(hipEventSynchronize(c->ev1));
    HIP_TRY(hipEventElapsedTime(walk_ms, c->ev0, c->ev1));
    HIP_TRY(hipMemcpy(c->h_counters, c->d_counters, kNumCounters * 8, hipMemcpyDeviceToHost));
    return hipSuccess;
}

hipError_t compact_corpus(felog_ctx *c, uint32_t n, float *ms) {
    if (n + 1 > c->off_cap) {
        dfree(c->d_off);
        dfree(c->d_sent_id);
        HIP_TRY(hipMalloc(&c->d_off, ((uint64_t)n + 1) * 8));
        HIP_TRY(hipMalloc(&c->d_sent_id, ((uint64_t)n + 1) * 4));
        c->off_cap = (uint64_t)n + 1;
    }
    HIP_TRY(hipEventRecord(c->ev0, c->stream));
    k_gather_lens<<<grid_for(n, 256, 65535), 256, 0, c->stream>>>(c->d_active, n, c->d_lens, c->d_off + 1);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemsetAsync(c->d_off, 0, 8, c->stream));
    size_t tmp = 0;
    HIP_TRY(hipcub::DeviceScan::InclusiveSum(nullptr, tmp, c->d_off + 1, c->d_off + 1, (uint64_t)n, c->stream));
    if (tmp > c->tmp_bytes) {
        dfree(c->d_tmp);
        HIP_TRY(hipMalloc(&c->d_tmp, tmp));
        c->tmp_bytes = tmp;
    }
    HIP_TRY(hipcub::DeviceScan::InclusiveSum(c->d_tmp, tmp, c->d_off + 1, c->d_off + 1, (uint64_t)n, c->stream));
    uint64_t total = 0;
    HIP_TRY(hipMemcpyAsync(&total, c->d_off + n, 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (total > c->tok_cap) {
        dfree(c->d_tok);
        const uint64_t cap = total + total / 8 + 1024;
        HIP_TRY(hipMalloc(&c->d_tok, cap * 4));
        c->tok_cap = cap;
    }
    k_compact<<<grid_for((uint64_t)n * 64, 256, 1u << 16), 256, 0, c->stream>>>(
        c->d_active, n, c->d_off, c->d_slots, c->slot, c->d_ovf_base, c->d_ovf, c->d_tok, c->d_sent_id);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(c->ev1, c->stream));
    HIP_TRY(hipEventSynchronize(c->ev1));
    HIP_TRY(hipEventElapsedTime(ms, c->ev0, c->ev1));
    c->corpus.n_tokens = total;
    return hipSuccess;
}

}  // namespace felog
