// eval_kernel.hip - convergence evaluation of the walk rounds.
//
// Restates src/word2vec.cu:2016-2080 with the pair kernel of :1448-1506:
// every vocabulary vertex that is not yet converged is scored by the mean
// cosine similarity (f32) of its syn0 row to at most 30 neighbours
// (EVALUATION_NEIGHBOUR_NUM, :45) - all of them when the degree is <= 30,
// otherwise 30 distinct neighbours drawn with Floyd's algorithm from Philox
// (the reference shuffles with a std::random_device seed, :1787-1792).
// Scores above 0.65 (:46) set the vertex's flag, after which no walker
// starts there.  The budget recheck runs the same test once, at min_R, on
// the budgeted vertices whose count ends there: each that fails gets one
// more round; flags and the stop rule's counts are left alone.
#include <hip/hip_runtime.h>
#include "felog_dev.h"
#include "felog_internal.h"

namespace felog {

namespace {

constexpr uint32_t kEvalNbrs = 30;
constexpr float kConverge = 0.65f;

__device__ __forceinline__ uint64_t readlane64(uint64_t x, uint32_t lane) {
    const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)x, lane);
    const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(x >> 32), lane);
    return ((uint64_t)hi << 32) | lo;
}

// sum over the 16 lanes of this lane's group
__device__ __forceinline__ float group_sum16(float x) {
    for (int o = 8; o > 0; o >>= 1) x += __shfl_xor(x, o);
    return x;
}

// Columns e0 .. e0+7 of a row, 0 from column d on.  Row pitches are
// multiples of 32 elements, e0 one of 8 below d, so the group is always one
// aligned 16- or 32-byte load inside the row.
template <typename T>
__device__ __forceinline__ void load8(const T *row, uint32_t e0, uint32_t d, float (&o)[8]) {
    if constexpr (sizeof(T) == 4) {
        const float4 a = *(const float4 *)(row + e0), b = *(const float4 *)(row + e0 + 4);
        o[0] = a.x; o[1] = a.y; o[2] = a.z; o[3] = a.w;
        o[4] = b.x; o[5] = b.y; o[6] = b.z; o[7] = b.w;
    } else {
        const uint4 a = *(const uint4 *)(row + e0);
        const uint32_t w[4] = {a.x, a.y, a.z, a.w};
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            o[2 * j] = bf16_to_f32((uint16_t)(w[j] & 0xFFFFu));
            o[2 * j + 1] = bf16_to_f32((uint16_t)(w[j] >> 16));
        }
    }
#pragma unroll
    for (int j = 0; j < 8; ++j)
        if (e0 + j >= d) o[j] = 0.f;
}

template <typename T>
__global__ __launch_bounds__(256) void k_eval(const uint64_t *__restrict__ row_ptr,
                                              const uint32_t *__restrict__ col,
                                              const uint8_t *__restrict__ in_vocab,
                                              uint8_t *__restrict__ flags, float *__restrict__ scores,
                                              const T *__restrict__ syn0, uint32_t V, uint32_t dp,
                                              uint32_t d, const RoundState *__restrict__ rs, int gated,
                                              int recheck, uint64_t seed, uint16_t *__restrict__ budget_cnt,
                                              unsigned long long *counters) {
    const RoundState s = *rs;
    if (gated) {
        // the recheck belongs to round min_R, when that round walked
        const bool due = recheck ? s.round == s.eval_from && !(s.budget && s.n_active == 0) : round_evaluates(s);
        if (!due) return;
    }
    const uint32_t round = s.round;
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
    const uint32_t grp = lane >> 4, l = lane & 15;
    unsigned long long n_eval = 0, n_conv = 0;
    // Each wavefront takes 64 consecutive vertices at a time: every lane
    // tests one for the evaluation (in vocabulary, not converged, budget
    // state) and reads its CSR range, then the wavefront scores the chosen
    // ones one after the other.
    for (uint32_t base = wave * 64; base < V; base += nwaves * 64) {
        const uint32_t vl = base + lane;
        bool pick = false;
        uint64_t lb = 0, le = 0;
        if (vl < V && in_vocab[vl] && !flags[vl]) {
            // vertices with a learned walk budget take no convergence test,
            // except the recheck's when their count ends this round
            const uint32_t bc = budget_cnt ? budget_cnt[vl] : 0u;
            pick = recheck ? bc == round : bc == 0u;
            if (pick) {
                lb = row_ptr[vl];
                le = row_ptr[vl + 1];
            }
        }
        for (uint64_t m = __ballot(pick); m; m &= m - 1) {
            const uint32_t src = (uint32_t)__builtin_ctzll(m);
            const uint32_t v = base + src;
            // read from lane src into scalar registers: everything that
            // follows on v's degree is wave-uniform control flow
            const uint64_t b = readlane64(lb, src);
            const uint32_t deg = (uint32_t)(readlane64(le, src) - b);
            float score = 0.f;
            if (deg > 0) {
                const uint32_t cnt = deg < kEvalNbrs ? deg : kEvalNbrs;
                // chosen neighbour index held by lane i (i < cnt)
                uint32_t mine = lane;
                if (deg > kEvalNbrs) {
                    // Floyd's subset: draw i (lane i, all at once) picks
                    // t_i in [0, j_i], j_i = deg - 30 + i; taken already -> j_i
                    const uint32_t jl = deg - kEvalNbrs + lane;
                    uint32_t t = 0;
                    if (lane < kEvalNbrs) t = draw(seed, v, DOM_EVAL | lane, 0xFFFFFFFFu, round).x % (jl + 1);
                    mine = 0xFFFFFFFFu;
                    for (uint32_t i = 0; i < kEvalNbrs; ++i) {
                        const uint32_t ti = __builtin_amdgcn_readlane(t, i);
                        const bool dup = __ballot(lane < i && mine == ti) != 0;
                        if (lane == i) mine = dup ? jl : t;
                    }
                }
                // neighbour vertex of lane i (i < cnt), one gather for all
                const uint32_t nbr = lane < cnt ? col[b + mine] : v;
                // Four neighbours at a time, one per 16-lane group: lane l of a
                // group holds columns 8l .. 8l+7 of every 128-column block and
                // the group reduces the dot product and the norm over 16 lanes.
                // Neighbours past cnt point at v's own row (nbr = v there),
                // so the loads of a round of four need no test; four rounds
                // (sixteen neighbours) are loaded before any is reduced.
                const T *rv = syn0 + (uint64_t)v * dp;
                uint32_t ru[8];  // neighbour of round k in this group
    #pragma unroll
                for (int k = 0; k < 8; ++k) ru[k] = __shfl(nbr, 4 * k + (int)grp);
                float nn = 0.f, dot[8], mm[8];
    #pragma unroll
                for (int k = 0; k < 8; ++k) dot[k] = mm[k] = 0.f;
                for (uint32_t e0 = 8 * l; e0 < d; e0 += 128) {
                    float x[8];
                    load8(rv, e0, d, x);
    #pragma unroll
                    for (int j = 0; j < 8; ++j) nn += x[j] * x[j];
    #pragma unroll
                    for (int h = 0; h < 8; h += 4) {
                        if (4 * h >= cnt) break;
                        float y[4][8];
    #pragma unroll
                        for (int k = 0; k < 4; ++k) load8(syn0 + (uint64_t)ru[h + k] * dp, e0, d, y[k]);
    #pragma unroll
                        for (int k = 0; k < 4; ++k) {
    #pragma unroll
                            for (int j = 0; j < 8; ++j) {
                                dot[h + k] += x[j] * y[k][j];
                                mm[h + k] += y[k][j] * y[k][j];
                            }
                        }
                    }
                }
                nn = group_sum16(nn);
                const float nv = sqrtf(nn);
                float sum = 0.f;  // this group's neighbours
    #pragma unroll
                for (int k = 0; k < 8; ++k) {
                    if (4 * k >= cnt) break;
                    const float dk = group_sum16(dot[k]);
                    const float nu = sqrtf(group_sum16(mm[k]));
                    if (4 * k + grp < cnt) sum += (nv == 0.f || nu == 0.f) ? 0.f : dk / (nv * nu);
                }
                sum += __shfl_xor(sum, 16);
                sum += __shfl_xor(sum, 32);
                score = sum / (float)cnt;
            }
            if (lane == 0) {
                if (recheck) {
                    if (!(score > kConverge)) budget_cnt[v] = (uint16_t)(round + 1);
                } else {
                    scores[v] = score;
                    if (score > kConverge) {
                        flags[v] = 1;
                        ++n_conv;
                    } else {
                        ++n_eval;
                    }
                }
            }
        }
    }
    // one atomic per workgroup and counter (atomics on one address serialise)
    __shared__ unsigned long long part[2][4];
    if (lane == 0) {
        part[0][threadIdx.x >> 6] = n_eval;
        part[1][threadIdx.x >> 6] = n_conv;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        n_eval = part[0][0] + part[0][1] + part[0][2] + part[0][3];
        n_conv = part[1][0] + part[1][1] + part[1][2] + part[1][3];
        if (n_eval) atomicAdd(&counters[C_EVALNUM], n_eval);
        if (n_conv) atomicAdd(&counters[C_EVALCONV], n_conv);
    }
}

}  // namespace

#define HIP_TRY(x)                              \
    do {                                        \
        hipError_t e_ = (x);                    \
        if (e_ != hipSuccess) return e_;        \
    } while (0)

hipError_t eval_launch(felog_ctx *c, bool gated, bool recheck) {
    const unsigned blocks = c->cu_count * 8;
    uint16_t *bcnt = c->budget_on ? c->d_budget_cnt : nullptr;
    if (c->opt.precision == FELOG_BF16)
        k_eval<uint16_t><<<blocks, 256, 0, c->stream>>>(c->d_row_ptr, c->d_col, c->d_in_vocab, c->d_flags,
                                                       c->d_scores, (const uint16_t *)c->d_syn0, c->V, c->dp,
                                                       c->d, c->d_rs, gated ? 1 : 0, recheck ? 1 : 0,
                                                       c->opt.seed, bcnt, (unsigned long long *)c->d_counters);
    else
        k_eval<float><<<blocks, 256, 0, c->stream>>>(c->d_row_ptr, c->d_col, c->d_in_vocab, c->d_flags,
                                                    c->d_scores, (const float *)c->d_syn0, c->V, c->dp, c->d,
                                                    c->d_rs, gated ? 1 : 0, recheck ? 1 : 0, c->opt.seed, bcnt,
                                                    (unsigned long long *)c->d_counters);
    return hipGetLastError();
}

hipError_t eval_enqueue(felog_ctx *c, bool gated) { return eval_launch(c, gated, false); }

hipError_t recheck_enqueue(felog_ctx *c, bool gated) {
    if (!c->budget_on) return hipSuccess;
    return eval_launch(c, gated, true);
}

// Eager evaluation of the round the round state holds.
hipError_t run_eval(felog_ctx *c, float *ms) {
    HIP_TRY(hipMemsetAsync(c->d_counters + C_EVALNUM, 0, 16, c->stream));
    HIP_TRY(hipEventRecord(c->ev0, c->stream));
    HIP_TRY(eval_enqueue(c, false));
    HIP_TRY(hipEventRecord(c->ev1, c->stream));
    HIP_TRY(hipEventSynchronize(c->ev1));
    HIP_TRY(hipEventElapsedTime(ms, c->ev0, c->ev1));
    HIP_TRY(hipMemcpy(c->h_counters, c->d_counters, kNumCounters * 8, hipMemcpyDeviceToHost));
    return hipSuccess;
}

}  // namespace felog
