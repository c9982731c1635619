// w2v_kernels.hip - the reference's per-pair word2vec trainers on the GPU.
//
// felog's defaults select word2vec kernels that the HogBatch tile trainer
// does not replace: -cbow 1 trains cbow_kernel (src/word2vec.cu:538-646),
// -cbow 0 -reuse-neg 1 trains __sgNegReuse (:368-440), and -cbow 0
// -reuse-neg 0 skip_gram_kernel (:442-536).  These kernels restate them
// with the reference's update order and random streams:
//   - one wavefront per sentence (lanes hold columns e, e + 64, ...; dot
//     products reduce over the wave); pairs, and for CBOW centres, are
//     applied one after another inside a sentence, and sentences run
//     concurrently with hardware float atomics on the tables (Hogwild, as
//     the reference's atomicAdd);
//   - skip-gram and CBOW draw the window shrink and the negatives from the
//     kernels' LCGs (multipliers 2514903917 and 25214903917, seeded with
//     the sentence's index in its training batch of at most 15000 sentences
//     or 299,980 tokens, src/word2vec.cu:1196-1234) through the 1e8-entry
//     unigram table (src/word2vec.cu:674-694), built here bit for bit;
//   - the sigmoid is read from the 1000-entry expTable at (int)((f+6)*83);
//   - the negative-reuse trainer takes a full window and one set of
//     negatives per sentence through the same table; the reference draws
//     them from a host LCG shared with the subsampling (:1251-1259), here
//     from Philox(seed; j, W2V, sentence, round).
// Sentences are filtered as for the tile trainer (vocabulary, the Philox
// subsample coin, the 1000-token cut) so both trainers see the same tokens.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <cmath>
#include <vector>
#include "felog_dev.h"
#include "felog_internal.h"

namespace felog {
namespace {

constexpr uint32_t kBatchSentences = 15000;            // src/word2vec.cu:1196 (MAX_SENTENCE)
constexpr uint64_t kBatchTokens = (15000ull - 1) * 20;  // src/word2vec.cu:1234
constexpr uint32_t kKeepMax = 1000;                     // MAX_SENTENCE_LENGTH
constexpr int kW2vBlock = 256;
constexpr uint32_t DOM_W2V = 0x70000000u;              // x = negative j, z = sentence, w = round

struct W2vArgs {
    const uint64_t *__restrict__ off;
    const uint32_t *__restrict__ tok;
    uint32_t *__restrict__ ftok;  // kept tokens of sentence s at off[s] + i
    uint32_t *__restrict__ kcnt;  // kept count
    uint32_t *__restrict__ blk;   // index of the sentence in its batch
    const RoundState *rs;
    float *syn0, *syn1;
    uint32_t dp, d;
    const uint8_t *__restrict__ in_vocab;
    const uint32_t *__restrict__ vocab_rank, *__restrict__ keep_thr, *__restrict__ vocab_order;
    const int32_t *__restrict__ table;
    uint32_t table_size;
    uint64_t vocab_size;
    const float *__restrict__ exp_table;
    uint32_t window, negative;
    int compat_break;
    uint64_t seed;
    unsigned long long *counters;
};

__device__ __forceinline__ float wave_sum(float x) {
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    return x;
}

// The sentence filter of the tile trainer's prep: vocabulary lookup, the
// subsample coin Philox(seed; position, SGNS, sentence, round) word 0
// against the keep threshold, 1000 kept tokens at most, optional break at
// vocabulary index 0 or a dropped vertex.
__global__ __launch_bounds__(kW2vBlock) void k_w2v_filter(W2vArgs a) {
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    const uint64_t n_sent = a.rs->n_active;
    const uint32_t round = a.rs->round;
    unsigned long long my_words = 0, my_kept = 0;
    for (uint64_t s = wave; s < n_sent; s += nwaves) {
        const uint64_t b0 = a.off[s];
        const uint32_t n = (uint32_t)(a.off[s + 1] - b0);
        uint32_t m = 0, words = 0;
        for (uint32_t chunk = 0; chunk < n; chunk += 64) {
            const uint32_t p = chunk + lane;
            int status = 0;  // 0 none, 1 drop, 2 keep, 3 break
            uint32_t t = 0;
            if (p < n) {
                t = a.tok[b0 + p];
                const uint32_t u = draw(a.seed, p, DOM_SGNS, (uint32_t)s, round).x;
                const uint32_t thr = a.keep_thr[t];
                if (a.compat_break && (a.in_vocab[t] == 0 || a.vocab_rank[t] == 0)) status = 3;
                else if (thr == 0u) status = 1;
                else status = (thr == 0xFFFFFFFFu || u < thr) ? 2 : 1;
            }
            const uint64_t brk_bal = __ballot(status == 3);
            const uint32_t brk = brk_bal ? chunk + (uint32_t)__ffsll((long long)brk_bal) - 1 : 0xFFFFFFFFu;
            const bool keep = status == 2 && p < brk;
            const uint64_t bal = __ballot(keep);
            const uint32_t idx = m + (uint32_t)__popcll(bal & lt_mask);
            if (keep && idx < kKeepMax) a.ftok[b0 + idx] = t;
            uint32_t read_end = n < chunk + 64 ? n : chunk + 64;
            if (brk != 0xFFFFFFFFu) read_end = brk + 1;
            const uint64_t fill = __ballot(keep && idx == kKeepMax - 1);
            if (fill) read_end = chunk + (uint32_t)__ffsll((long long)fill);
            words = read_end;
            m += (uint32_t)__popcll(bal);
            if (brk != 0xFFFFFFFFu || m >= kKeepMax) break;
        }
        if (m > kKeepMax) m = kKeepMax;
        if (lane == 0) a.kcnt[s] = m;
        my_words += words;
        my_kept += m;
    }
    if (lane == 0) {
        atomicAdd(&a.counters[C_WORDS], my_words);
        atomicAdd(&a.counters[C_KEPT], my_kept);
    }
}

// Index of every sentence in its training batch: one wavefront scans the
// kept counts 64 sentences at a time.  A batch closes after the sentence
// that makes it 15000 sentences or 299,980 kept tokens; a batch holds at
// least 300 sentences (1000 tokens each at most), so at most one batch
// closes inside a group of 64.
__global__ __launch_bounds__(64) void k_w2v_batches(W2vArgs a) {
    const uint32_t lane = threadIdx.x;
    const uint64_t n_sent = a.rs->n_active;
    uint64_t cnt = 0, tot = 0;
    for (uint64_t base = 0; base < n_sent; base += 64) {
        const uint64_t s = base + lane;
        const bool valid = s < n_sent;
        uint64_t P = valid ? a.kcnt[s] : 0;
        for (int o = 1; o < 64; o <<= 1) {
            const uint64_t y = __shfl_up(P, o);
            if ((int)lane >= o) P += y;
        }
        const uint32_t nvalid = (uint32_t)__popcll(__ballot(valid));
        const uint64_t last = __shfl(P, (int)nvalid - 1);
        const uint64_t close = __ballot(valid && (cnt + lane + 1 == kBatchSentences || tot + P >= kBatchTokens));
        if (!close) {
            if (valid) a.blk[s] = (uint32_t)(cnt + lane);
            cnt += nvalid;
            tot += last;
        } else {
            const uint32_t c = (uint32_t)__ffsll((long long)close) - 1;
            const uint64_t Pc = __shfl(P, (int)c);
            if (valid) a.blk[s] = lane <= c ? (uint32_t)(cnt + lane) : lane - c - 1;
            cnt = nvalid - c - 1;
            tot = last - Pc;
        }
    }
}

// Sigmoid step of the reference: the clamp at +-6, then expTable.
__device__ __forceinline__ float w2v_grad(float f, float label, float alpha, const float *__restrict__ et) {
    if (f > 6.f) return (label - 1.f) * alpha;
    if (f < -6.f) return label * alpha;
    return (label - et[(int)((f + 6.f) * 83.f)]) * alpha;
}

// One negative of the kernels' LCG through the unigram table
// (src/word2vec.cu:500-505); returns a vertex id.
__device__ __forceinline__ uint32_t lcg_negative(const W2vArgs &a, unsigned long long &nr) {
    nr = nr * 25214903917ull + 11ull;
    uint32_t tr = (uint32_t)a.table[(nr >> 16) % a.table_size];
    if (tr == 0) tr = (uint32_t)(nr % (a.vocab_size - 1) + 1);
    return a.vocab_order[tr];
}

// MODE 0 skip-gram, 1 negative reuse, 2 CBOW; NE columns per lane.
template <int MODE, int NE>
__global__ __launch_bounds__(kW2vBlock) void k_w2v_train(W2vArgs a) {
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    const uint64_t n_sent = a.rs->n_active;
    const uint32_t round = a.rs->round;
    const float alpha = a.rs->alpha;
    const int w = (int)a.window;
    const uint32_t k = a.negative, d = a.d, dp = a.dp;
    unsigned long long my_pairs = 0, my_sent = 0;
    for (uint64_t s = wave; s < n_sent; s += nwaves) {
        const uint64_t b0 = a.off[s];
        const int m = (int)a.kcnt[s];
        if (m < 2) continue;
        ++my_sent;
        unsigned long long nr = a.blk[s];
        uint32_t my_neg = 0;  // MODE 1: negative `lane` of the sentence
        if (MODE == 1 && lane < k) {
            const uint32_t r = draw(a.seed, lane, DOM_W2V, (uint32_t)s, round).x;
            uint32_t tr = (uint32_t)a.table[r % a.table_size];
            if (tr == 0) tr = (uint32_t)(r % (a.vocab_size - 1) + 1);
            my_neg = a.vocab_order[tr];
        }
        for (int p = 0; p < m; ++p) {
            const uint32_t word = a.ftok[b0 + p];
            int b = 0;
            if (MODE != 1) {
                nr = nr * 2514903917ull + 11ull;  // src/word2vec.cu:458 (sic)
                b = (int)(nr % (unsigned long long)w);
            }
            if (MODE == 2) {
                float h[NE];
#pragma unroll
                for (int t = 0; t < NE; ++t) h[t] = 0.f;
                int cw = 0;
                for (int q = b; q < 2 * w + 1 - b; ++q) {
                    const int c = p - w + q;
                    if (q == w || c < 0 || c >= m) continue;
                    const float *r0 = a.syn0 + (uint64_t)a.ftok[b0 + c] * dp;
#pragma unroll
                    for (int t = 0; t < NE; ++t) {
                        const uint32_t e = lane + 64 * t;
                        if (e < d) h[t] += r0[e];
                    }
                    ++cw;
                }
                if (!cw) continue;
#pragma unroll
                for (int t = 0; t < NE; ++t) h[t] /= (float)cw;
                float ev[NE];
#pragma unroll
                for (int t = 0; t < NE; ++t) ev[t] = 0.f;
                for (uint32_t dd = 0; dd <= k; ++dd) {
                    uint32_t target = word;
                    float label = 1.f;
                    if (dd) {
                        target = lcg_negative(a, nr);
                        if (target == word) continue;
                        label = 0.f;
                    }
                    float *r1 = a.syn1 + (uint64_t)target * dp;
                    float l2[NE], part = 0.f;
#pragma unroll
                    for (int t = 0; t < NE; ++t) {
                        const uint32_t e = lane + 64 * t;
                        l2[t] = e < d ? r1[e] : 0.f;
                        part += h[t] * l2[t];
                    }
                    const float g = w2v_grad(wave_sum(part), label, alpha, a.exp_table);
#pragma unroll
                    for (int t = 0; t < NE; ++t) {
                        const uint32_t e = lane + 64 * t;
                        ev[t] += g * l2[t];
                        if (e < d) atomicAdd(&r1[e], g * h[t]);
                    }
                }
                for (int q = b; q < 2 * w + 1 - b; ++q) {
                    const int c = p - w + q;
                    if (q == w || c < 0 || c >= m) continue;
                    float *r0 = a.syn0 + (uint64_t)a.ftok[b0 + c] * dp;
#pragma unroll
                    for (int t = 0; t < NE; ++t) {
                        const uint32_t e = lane + 64 * t;
                        if (e < d) atomicAdd(&r0[e], ev[t]);
                    }
                }
                my_pairs += cw;
                continue;
            }
            for (int q = b; q < 2 * w + 1 - b; ++q) {
                const int c = p - w + q;
                if (q == w || c < 0 || c >= m) continue;
                float *r0 = a.syn0 + (uint64_t)a.ftok[b0 + c] * dp;
                float l1[NE], ev[NE];
#pragma unroll
                for (int t = 0; t < NE; ++t) {
                    const uint32_t e = lane + 64 * t;
                    l1[t] = e < d ? r0[e] : 0.f;
                    ev[t] = 0.f;
                }
                ++my_pairs;
                for (uint32_t dd = 0; dd <= k; ++dd) {
                    uint32_t target = word;
                    float label = 1.f;
                    if (dd) {
                        target = MODE == 1 ? (uint32_t)__shfl((int)my_neg, (int)dd - 1) : lcg_negative(a, nr);
                        if (target == word) continue;
                        label = 0.f;
                    }
                    float *r1 = a.syn1 + (uint64_t)target * dp;
                    float l2[NE], part = 0.f;
#pragma unroll
                    for (int t = 0; t < NE; ++t) {
                        const uint32_t e = lane + 64 * t;
                        l2[t] = e < d ? r1[e] : 0.f;
                        part += l1[t] * l2[t];
                    }
                    const float g = w2v_grad(wave_sum(part), label, alpha, a.exp_table);
#pragma unroll
                    for (int t = 0; t < NE; ++t) {
                        const uint32_t e = lane + 64 * t;
                        ev[t] += g * l2[t];
                        if (e < d) atomicAdd(&r1[e], g * l1[t]);
                    }
                }
#pragma unroll
                for (int t = 0; t < NE; ++t) {
                    const uint32_t e = lane + 64 * t;
                    if (e < d) atomicAdd(&r0[e], ev[t]);
                }
            }
        }
    }
    if (lane == 0) {
        if (my_pairs) atomicAdd(&a.counters[C_PAIRS], my_pairs);
        if (my_sent) atomicAdd(&a.counters[C_GROUPS], my_sent);
    }
}

// Unigram table of src/word2vec.cu:674-694: slot a holds index i until
// a / T passes d1_i, the cumulative share of indices 0..i (a sequential
// double sum, made on the host like the reference), the slot that passes it
// included.  With f_i = the first slot a with a / T > d1_i, index i leaves
// at slot a_i = max(a_{i-1} + 1, f_i) = i + max_{j <= i}(f_j - j), and
// table[a] = #{i : a_i < a} (at most vocab_size - 1): a max-scan and a
// binary search per slot give the sequential loop's table.
__global__ void k_table_first(const double *__restrict__ d1, uint64_t n, uint32_t T, long long *__restrict__ fm) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const double x = d1[i];
        long long a0 = (long long)(x * (double)T) - 2;
        if (a0 < 0) a0 = 0;
        while (!((double)a0 / (double)T > x)) ++a0;
        fm[i] = a0 - (long long)i;
    }
}

__global__ void k_table_fill(const long long *__restrict__ M, uint64_t n, uint32_t T, int32_t *__restrict__ table) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t a = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; a < T; a += stride) {
        // first i with a_i = i + M[i] >= a
        uint64_t lo = 0, hi = n;
        while (lo < hi) {
            const uint64_t mid = lo + (hi - lo) / 2;
            if ((long long)mid + M[mid] >= (long long)a) hi = mid;
            else lo = mid + 1;
        }
        table[a] = (int32_t)(lo < n - 1 ? lo : n - 1);
    }
}

struct MaxLL {
    __host__ __device__ long long operator()(long long x, long long y) const { return x > y ? x : y; }
};

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

hipError_t w2v_prepare(felog_ctx *c) {
    if (c->d_w2v_table) return hipSuccess;
    const uint32_t V = c->V, T = c->opt.w2v_table_size ? c->opt.w2v_table_size : 100000000u;
    const uint64_t vs = c->vocab_size;
    // expTable (src/word2vec.cu:2292-2297)
    float et[1001];
    for (int i = 0; i < 1000; ++i) {
        et[i] = expf((i / (float)1000 * 2 - 1) * 6);
        et[i] = et[i] / (et[i] + 1);
    }
    et[1000] = et[999];
    HIP_TRY(hipMalloc(&c->d_w2v_exp, sizeof(et)));
    HIP_TRY(hipMemcpy(c->d_w2v_exp, et, sizeof(et), hipMemcpyHostToDevice));
    // cumulative shares in vocabulary order, summed in order in double
    std::vector<uint32_t> order(V), deg(V);
    HIP_TRY(hipMemcpy(order.data(), c->d_vocab_order, (uint64_t)V * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(deg.data(), c->d_deg, (uint64_t)V * 4, hipMemcpyDeviceToHost));
    double pw = 0.0;
    for (uint64_t i = 0; i < vs; ++i) pw += std::pow((double)deg[order[i]], 0.75);
    std::vector<double> d1(vs ? vs : 1);
    double acc = vs ? std::pow((double)deg[order[0]], 0.75) / pw : 0.0;
    for (uint64_t i = 0; i < vs; ++i) {
        if (i) acc += std::pow((double)deg[order[i]], 0.75) / pw;
        d1[i] = acc;
    }
    double *d_d1 = nullptr;
    long long *fm = nullptr, *M = nullptr;
    void *tmp = nullptr;
    size_t tb = 0;
    const uint64_t n = vs ? vs : 1;
    HIP_TRY(hipMalloc(&d_d1, n * 8));
    HIP_TRY(hipMalloc(&fm, n * 8));
    HIP_TRY(hipMalloc(&M, n * 8));
    HIP_TRY(hipMemcpy(d_d1, d1.data(), n * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMalloc(&c->d_w2v_table, (uint64_t)T * 4));
    k_table_first<<<grid_for(n, 256, 65535), 256, 0, c->stream>>>(d_d1, n, T, fm);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipcub::DeviceScan::InclusiveScan(nullptr, tb, fm, M, MaxLL{}, n, c->stream));
    HIP_TRY(hipMalloc(&tmp, tb + 16));
    HIP_TRY(hipcub::DeviceScan::InclusiveScan(tmp, tb, fm, M, MaxLL{}, n, c->stream));
    k_table_fill<<<grid_for(T, 256, 1u << 20), 256, 0, c->stream>>>(M, n, T, c->d_w2v_table);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(c->stream));
    dfree(d_d1);
    dfree(fm);
    dfree(M);
    dfree(tmp);
    c->w2v_table_size = T;
    return hipSuccess;
}

hipError_t w2v_reserve(felog_ctx *c, uint64_t n_sent, uint64_t n_tok) {
    HIP_TRY(w2v_prepare(c));
    if (n_sent > c->w2v_cap_sent) {
        drop_round_graph(c);
        dfree(c->d_w2v_kcnt);
        dfree(c->d_w2v_blk);
        c->d_w2v_kcnt = c->d_w2v_blk = nullptr;
        c->w2v_cap_sent = 0;
        HIP_TRY(hipMalloc(&c->d_w2v_kcnt, (n_sent + 1) * 4));
        HIP_TRY(hipMalloc(&c->d_w2v_blk, (n_sent + 1) * 4));
        c->w2v_cap_sent = n_sent;
    }
    if (n_tok > c->w2v_cap_tok) {
        drop_round_graph(c);
        dfree(c->d_w2v_ftok);
        c->d_w2v_ftok = nullptr;
        c->w2v_cap_tok = 0;
        HIP_TRY(hipMalloc(&c->d_w2v_ftok, (n_tok + 1) * 4));
        c->w2v_cap_tok = n_tok;
    }
    return hipSuccess;
}

hipError_t w2v_enqueue(felog_ctx *c, const uint64_t *d_off, const uint32_t *d_tok) {
    W2vArgs a;
    a.off = d_off;
    a.tok = d_tok;
    a.ftok = c->d_w2v_ftok;
    a.kcnt = c->d_w2v_kcnt;
    a.blk = c->d_w2v_blk;
    a.rs = c->d_rs;
    a.syn0 = (float *)c->d_syn0;
    a.syn1 = (float *)c->d_syn1;
    a.dp = c->dp;
    a.d = c->d;
    a.in_vocab = c->d_in_vocab;
    a.vocab_rank = c->d_vocab_rank;
    a.keep_thr = c->d_keep_thr;
    a.vocab_order = c->d_vocab_order;
    a.table = c->d_w2v_table;
    a.table_size = c->w2v_table_size;
    a.vocab_size = c->vocab_size;
    a.exp_table = c->d_w2v_exp;
    a.window = c->opt.window;
    a.negative = c->opt.negative;
    a.compat_break = c->opt.compat_vocab_break;
    a.seed = c->opt.seed;
    a.counters = (unsigned long long *)c->d_counters;
    const unsigned blocks = c->cu_count * 16;
    k_w2v_filter<<<blocks, kW2vBlock, 0, c->stream>>>(a);
    HIP_TRY(hipGetLastError());
    k_w2v_batches<<<1, 64, 0, c->stream>>>(a);
    HIP_TRY(hipGetLastError());
    const int mode = c->opt.trainer;
    const uint32_t ne = c->d <= 64 ? 1 : c->d <= 128 ? 2 : c->d <= 256 ? 4 : 8;
#define W2V_LAUNCH(M)                                                                             \
    do {                                                                                          \
        if (ne == 1) k_w2v_train<M, 1><<<blocks, kW2vBlock, 0, c->stream>>>(a);                   \
        else if (ne == 2) k_w2v_train<M, 2><<<blocks, kW2vBlock, 0, c->stream>>>(a);              \
        else if (ne == 4) k_w2v_train<M, 4><<<blocks, kW2vBlock, 0, c->stream>>>(a);              \
        else k_w2v_train<M, 8><<<blocks, kW2vBlock, 0, c->stream>>>(a);                           \
    } while (0)
    if (mode == FELOG_TRAIN_SKIPGRAM) W2V_LAUNCH(0);
    else if (mode == FELOG_TRAIN_SKIPGRAM_REUSE) W2V_LAUNCH(1);
    else W2V_LAUNCH(2);
#undef W2V_LAUNCH
    return hipGetLastError();
}

hipError_t run_w2v(felog_ctx *c, const uint64_t *d_off, const uint32_t *d_tok, uint64_t n_sent, uint64_t n_tok,
                   float *ms) {
    HIP_TRY(w2v_reserve(c, n_sent, n_tok));
    HIP_TRY(hipMemsetAsync(c->d_counters + C_WORDS, 0, 8 * (C_BIGTILES - C_WORDS + 1), c->stream));
    HIP_TRY(hipEventRecord(c->ev0, c->stream));
    HIP_TRY(w2v_enqueue(c, d_off, d_tok));
    HIP_TRY(hipEventRecord(c->ev1, c->stream));
    HIP_TRY(hipEventSynchronize(c->ev1));
    HIP_TRY(hipEventElapsedTime(&ms[0], c->ev0, c->ev1));
    ms[1] = 0.f;
    ms[2] = ms[0];
    HIP_TRY(hipMemcpy(c->h_counters, c->d_counters, kNumCounters * 8, hipMemcpyDeviceToHost));
    return hipSuccess;
}

}  // namespace felog

// ---------------------------------------------------------------------------
// -classes: k-means over the vocabulary's syn0 rows (src/word2vec.cu:2128-2175):
// row a (vocabulary order) starts in class a % K; each of `iters` rounds
// sums the rows of every class (FP64 here), divides by (members + 1) as the
// reference's centcn, which starts at 1, does, scales every centroid to unit
// length, and moves each row to the class of the largest dot product (the
// first on ties, starting from -10).
// ---------------------------------------------------------------------------
namespace felog {
namespace {

__global__ void k_km_init(int32_t *__restrict__ cl, uint64_t vs, uint32_t K) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t a = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; a < vs; a += stride) cl[a] = (int32_t)(a % K);
}

__global__ void k_km_clear(double *__restrict__ sum, unsigned long long *__restrict__ cnt, uint64_t n, uint32_t K) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        sum[i] = 0.0;
        if (i < K) cnt[i] = 0;
    }
}

// One wavefront per row: its columns into the class sums.
__global__ __launch_bounds__(256) void k_km_accum(const float *__restrict__ syn0, uint32_t dp, uint32_t d,
                                                  const uint32_t *__restrict__ order, uint64_t vs,
                                                  const int32_t *__restrict__ cl, double *__restrict__ sum,
                                                  unsigned long long *__restrict__ cnt) {
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    for (uint64_t a = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; a < vs; a += nwaves) {
        const float *row = syn0 + (uint64_t)order[a] * dp;
        const int32_t c = cl[a];
        for (uint32_t e = lane; e < d; e += 64) atomicAdd(&sum[(uint64_t)c * d + e], (double)row[e]);
        if (lane == 0) atomicAdd(&cnt[c], 1ull);
    }
}

// One wavefront per class: mean (members + 1 in the divisor), unit length.
__global__ __launch_bounds__(64) void k_km_norm(const double *__restrict__ sum, const unsigned long long *__restrict__ cnt,
                                                uint32_t K, uint32_t d, float *__restrict__ cent) {
    const uint32_t lane = threadIdx.x;
    for (uint32_t k = blockIdx.x; k < K; k += gridDim.x) {
        const double div = (double)(cnt[k] + 1);
        double ss = 0.0;
        for (uint32_t e = lane; e < d; e += 64) {
            const double x = sum[(uint64_t)k * d + e] / div;
            ss += x * x;
        }
        for (int o = 32; o > 0; o >>= 1) ss += __shfl_xor(ss, o);
        // an empty class divides 0 by 0, as the reference does: its NaN
        // centroid never wins an assignment
        const double nrm = sqrt(ss);
        for (uint32_t e = lane; e < d; e += 64) {
            const double x = sum[(uint64_t)k * d + e] / div;
            cent[(uint64_t)k * d + e] = (float)(x / nrm);
        }
    }
}

// One wavefront per row: the class of the largest dot product.
__global__ __launch_bounds__(256) void k_km_assign(const float *__restrict__ syn0, uint32_t dp, uint32_t d,
                                                   const uint32_t *__restrict__ order, uint64_t vs,
                                                   const float *__restrict__ cent, uint32_t K,
                                                   int32_t *__restrict__ cl) {
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    for (uint64_t a = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; a < vs; a += nwaves) {
        const float *row = syn0 + (uint64_t)order[a] * dp;
        float x[8];
        const uint32_t ne = (d + 63) / 64;
        for (uint32_t t = 0; t < 8; ++t) x[t] = (t < ne && lane + 64 * t < d) ? row[lane + 64 * t] : 0.f;
        float best = -10.f;
        int32_t bid = 0;
        for (uint32_t k = 0; k < K; ++k) {
            const float *cr = cent + (uint64_t)k * d;
            float p = 0.f;
            for (uint32_t t = 0; t < ne; ++t)
                if (lane + 64 * t < d) p += x[t] * cr[lane + 64 * t];
            for (int o = 32; o > 0; o >>= 1) p += __shfl_xor(p, o);
            if (p > best) {
                best = p;
                bid = (int32_t)k;
            }
        }
        if (lane == 0) cl[a] = bid;
    }
}

}  // namespace

hipError_t kmeans_classes(felog_ctx *c, uint32_t K, uint32_t iters, int32_t *d_cl) {
    const uint64_t vs = c->vocab_size;
    const uint32_t d = c->d;
    double *sum = nullptr;
    unsigned long long *cnt = nullptr;
    float *cent = nullptr;
    HIP_TRY(hipMalloc(&sum, (uint64_t)K * d * 8));
    HIP_TRY(hipMalloc(&cnt, (uint64_t)K * 8));
    HIP_TRY(hipMalloc(&cent, (uint64_t)K * d * 4));
    const unsigned g = grid_for(vs * 64, 256, 1u << 16);
    k_km_init<<<grid_for(vs, 256, 65535), 256, 0, c->stream>>>(d_cl, vs, K);
    for (uint32_t it = 0; it < iters; ++it) {
        k_km_clear<<<grid_for((uint64_t)K * d, 256, 65535), 256, 0, c->stream>>>(sum, cnt, (uint64_t)K * d, K);
        k_km_accum<<<g, 256, 0, c->stream>>>((const float *)c->d_syn0, c->dp, d, c->d_vocab_order, vs, d_cl, sum, cnt);
        k_km_norm<<<K < 65535 ? K : 65535, 64, 0, c->stream>>>(sum, cnt, K, d, cent);
        k_km_assign<<<g, 256, 0, c->stream>>>((const float *)c->d_syn0, c->dp, d, c->d_vocab_order, vs, cent, K, d_cl);
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(c->stream));
    dfree(sum);
    dfree(cnt);
    dfree(cent);
    return hipSuccess;
}

}  // namespace felog
