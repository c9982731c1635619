// graph_kernels.hip - device construction of the walk graph: directed edge
// list (undirecting), CSR with neighbours in input order, per-vertex alias
// tables, the degree-ordered vocabulary and table initialisation.
//
// Reference behaviour restated here:
//   undirecting       include/graph.hpp:741-772 (edge i -> slots 2i, 2i+1)
//   out-degree        include/graph.hpp:775-805
//   CSR, input order  include/graph.hpp:531-575
//   vocabulary order  src/word2vec.cu:697-755, 834-862 (degree descending)
//   syn0 init (LCG)   src/word2vec.cu:864-902
// Alias tables are new (the reference declares AliasBucket,
// include/walk.hpp:87-103, but never builds one).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <algorithm>
#include <cmath>
#include <vector>
#include "felog_dev.h"
#include "felog_internal.h"

namespace felog {

namespace {

__global__ void k_undirect(const uint32_t *__restrict__ src, const uint32_t *__restrict__ dst,
                           const float *__restrict__ w, uint64_t m, int undirected, uint32_t V,
                           uint32_t *__restrict__ osrc, uint32_t *__restrict__ odst,
                           float *__restrict__ ow, unsigned long long *bad) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride) {
        uint32_t s = src[i], d = dst[i];
        const float x = w[i];
        if (s >= V || d >= V) {
            atomicAdd(bad, 1ull);
            s = s % V;  // keep every later index in range; the build fails on bad != 0
            d = d % V;
        }
        if (undirected) {
            osrc[2 * i] = s; odst[2 * i] = d; ow[2 * i] = x;
            osrc[2 * i + 1] = d; odst[2 * i + 1] = s; ow[2 * i + 1] = x;
        } else {
            osrc[i] = s; odst[i] = d; ow[i] = x;
        }
    }
}

__global__ void k_degree(const uint32_t *__restrict__ src, uint64_t E, uint32_t *__restrict__ deg) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < E; i += stride)
        atomicAdd(&deg[src[i]], 1u);
}

__global__ void k_pack(const uint32_t *__restrict__ dst, const float *__restrict__ w, uint64_t E,
                       uint64_t *__restrict__ out) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < E; i += stride)
        out[i] = (uint64_t)dst[i] | ((uint64_t)__float_as_uint(w[i]) << 32);
}

__global__ void k_unpack(const uint64_t *__restrict__ packed, uint64_t E, uint32_t *__restrict__ col,
                         float *__restrict__ w) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < E; i += stride) {
        const uint64_t x = packed[i];
        col[i] = (uint32_t)x;
        w[i] = __uint_as_float((uint32_t)(x >> 32));
    }
}

__global__ void k_vrec(const uint64_t *__restrict__ row_ptr, uint32_t V, VertexRec *__restrict__ vrec,
                       uint32_t *__restrict__ deg) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += stride) {
        const uint64_t b = row_ptr[v], e = row_ptr[v + 1];
        vrec[v] = VertexRec{(uint32_t)b, (uint32_t)(e - b)};
        deg[v] = (uint32_t)(e - b);
    }
}

// Walk vertex records: vrec with kSelfFirst on degree-2 vertices whose
// first neighbour is themselves.
__global__ void k_vwalk(const VertexRec *__restrict__ vrec, const uint32_t *__restrict__ col, uint32_t V,
                        VertexRec *__restrict__ vwalk) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += stride) {
        VertexRec r = vrec[v];
        if (r.deg == 2 && col[r.begin] == v) r.deg |= kSelfFirst;
        vwalk[v] = r;
    }
}

// Both neighbours' walk records into every edge record.
__global__ void k_edge_vr(const VertexRec *__restrict__ vwalk, uint64_t E, EdgeRec *__restrict__ erec) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E; e += stride) {
        EdgeRec r = erec[e];
        r.col_vr = vwalk[r.col];
        r.alias_vr = vwalk[r.alias_col];
        erec[e] = r;
    }
}

// Threshold of an alias column: P(keep column) = thresh / 2^32.
__device__ __forceinline__ uint32_t alias_quantise(double p) {
    if (p >= 1.0) return 0xFFFFFFFFu;
    if (!(p > 0.0)) return 0u;
    return (uint32_t)(p * 4294967296.0);
}

// Vose's alias construction, one thread per vertex.  Vertices are taken in
// degree-descending order so the lanes of a wavefront have similar work.
// The two work stacks share the segment stk[b, b+n): small entries grow
// upward from b, large entries downward from b+n-1.  Pops take the most
// recent push.  FP64 throughout: weights are summed in column order, then
// p_j = w_j * n / sum, and each pairing updates p_l = (p_l + p_s) - 1.
__global__ void k_alias(const uint32_t *__restrict__ order, uint32_t V,
                        const uint64_t *__restrict__ row_ptr, const float *__restrict__ w,
                        const uint32_t *__restrict__ col, double *__restrict__ p,
                        uint32_t *__restrict__ stk, EdgeRec *__restrict__ erec) {
#pragma clang fp contract(off)
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < V; i += stride) {
        const uint32_t v = order[i];
        const uint64_t b = row_ptr[v];
        const uint32_t n = (uint32_t)(row_ptr[v + 1] - b);
        if (n == 0) continue;
        double sum = 0.0;
        for (uint32_t j = 0; j < n; ++j) {
            const float x = w[b + j];
            sum = sum + (x > 0.f ? (double)x : 0.0);
        }
        uint32_t ns = 0, nl = 0;
        const double dn = (double)n;
        for (uint32_t j = 0; j < n; ++j) {
            const float x = w[b + j];
            const double pj = sum > 0.0 ? ((x > 0.f ? (double)x : 0.0) * dn) / sum : 1.0;
            p[b + j] = pj;
            if (pj < 1.0) stk[b + ns++] = j;
            else stk[b + n - 1 - nl++] = j;
        }
        while (ns > 0 && nl > 0) {
            const uint32_t s = stk[b + --ns];
            const uint32_t l = stk[b + n - nl];
            --nl;
            const double ps = p[b + s];
            erec[b + s] = EdgeRec{alias_quantise(ps), col[b + s], col[b + l], l, {}, {}};
            const double pl = (p[b + l] + ps) - 1.0;
            p[b + l] = pl;
            if (pl < 1.0) stk[b + ns++] = l;
            else stk[b + n - 1 - nl++] = l;
        }
        while (nl > 0) {
            const uint32_t l = stk[b + n - nl];
            --nl;
            erec[b + l] = EdgeRec{0xFFFFFFFFu, col[b + l], col[b + l], l, {}, {}};
        }
        while (ns > 0) {
            const uint32_t s = stk[b + --ns];
            erec[b + s] = EdgeRec{0xFFFFFFFFu, col[b + s], col[b + s], s, {}, {}};
        }
    }
}

// syn0 initialisation of src/word2vec.cu:893-896: element e of the
// vocabulary-ordered table takes x_{e+1} of the LCG x <- x*25214903917 + 11
// started at 1, as ((x & 0xFFFF)/65536 - 0.5)/d.  jump[k] holds the affine
// map applied 2^k times, so x_n is composed from the bits of n.
struct Affine { unsigned long long a, c; };

__device__ __forceinline__ unsigned long long lcg_nth(const Affine *__restrict__ jump, uint64_t n) {
    unsigned long long x = 1ull;
    for (int k = 0; n != 0; ++k, n >>= 1)
        if (n & 1) x = jump[k].a * x + jump[k].c;
    return x;
}

template <typename T>
__global__ void k_init_syn0(const uint32_t *__restrict__ rank, uint32_t V, uint32_t d, uint32_t dp,
                            const Affine *__restrict__ jump, T *__restrict__ syn0) {
    const uint64_t total = (uint64_t)V * dp;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        const uint32_t v = (uint32_t)(i / dp), b = (uint32_t)(i % dp);
        float val = 0.f;
        if (b < d) {
            const uint64_t e = (uint64_t)rank[v] * d + b;
            const unsigned long long x = lcg_nth(jump, e + 1);
            val = (float)(((double)(x & 0xFFFFull) / 65536.0 - 0.5) / (double)d);
        }
        if constexpr (sizeof(T) == 4) syn0[i] = val;
        else syn0[i] = f32_to_bf16_rne(val);
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

hipError_t launch_undirect(felog_ctx *c, const uint32_t *src, const uint32_t *dst, const float *w,
                           uint64_t m, bool undirected) {
    const uint64_t E = undirected ? 2 * m : m;
    HIP_TRY(hipMalloc(&c->d_esrc, E * 4));
    HIP_TRY(hipMalloc(&c->d_edst, E * 4));
    HIP_TRY(hipMalloc(&c->d_ew, E * 4));
    uint32_t *s_src = nullptr, *s_dst = nullptr;
    float *s_w = nullptr;
    HIP_TRY(hipMalloc(&s_src, m * 4));
    HIP_TRY(hipMalloc(&s_dst, m * 4));
    HIP_TRY(hipMalloc(&s_w, m * 4));
    HIP_TRY(hipMemcpyAsync(s_src, src, m * 4, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(s_dst, dst, m * 4, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(s_w, w, m * 4, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemsetAsync(c->d_counters, 0, kNumCounters * 8, c->stream));
    k_undirect<<<grid_for(m, 256, 65535), 256, 0, c->stream>>>(
        s_src, s_dst, s_w, m, undirected ? 1 : 0, c->V, c->d_esrc, c->d_edst, c->d_ew,
        (unsigned long long *)(c->d_counters + C_BADEDGE));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(c->stream));
    dfree(s_src);
    dfree(s_dst);
    dfree(s_w);
    c->E = E;
    return hipSuccess;
}

hipError_t build_csr(felog_ctx *c) {
    const uint64_t E = c->E;
    const uint32_t V = c->V;
    HIP_TRY(hipMalloc(&c->d_row_ptr, (uint64_t)(V + 1) * 8));
    HIP_TRY(hipMalloc(&c->d_deg, (uint64_t)V * 4));
    HIP_TRY(hipMalloc(&c->d_vrec, (uint64_t)V * sizeof(VertexRec)));
    HIP_TRY(hipMemsetAsync(c->d_deg, 0, (uint64_t)V * 4, c->stream));
    HIP_TRY(hipMalloc(&c->d_col, E * 4 + 4));
    HIP_TRY(hipMalloc(&c->d_w, E * 4 + 4));

    // Stable LSD radix sort of (src, dst|w) keeps each vertex's neighbours in
    // input order.
    uint64_t *vals_a = nullptr, *vals_b = nullptr;
    uint32_t *keys_b = nullptr;
    HIP_TRY(hipMalloc(&vals_a, E * 8 + 8));
    HIP_TRY(hipMalloc(&vals_b, E * 8 + 8));
    HIP_TRY(hipMalloc(&keys_b, E * 4 + 4));
    k_pack<<<grid_for(E, 256, 65535), 256, 0, c->stream>>>(c->d_edst, c->d_ew, E, vals_a);
    HIP_TRY(hipGetLastError());
    int bits = 1;
    while (bits < 32 && (1ull << bits) < (uint64_t)V) ++bits;
    hipcub::DoubleBuffer<uint32_t> keys(c->d_esrc, keys_b);
    hipcub::DoubleBuffer<uint64_t> vals(vals_a, vals_b);
    size_t tmp = 0;
    HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, keys, vals, E, 0, bits, c->stream));
    void *d_tmp = nullptr;
    HIP_TRY(hipMalloc(&d_tmp, tmp + 16));
    HIP_TRY(hipcub::DeviceRadixSort::SortPairs(d_tmp, tmp, keys, vals, E, 0, bits, c->stream));
    k_degree<<<grid_for(E, 256, 65535), 256, 0, c->stream>>>(keys.Current(), E, c->d_deg);
    HIP_TRY(hipGetLastError());
    k_unpack<<<grid_for(E, 256, 65535), 256, 0, c->stream>>>(vals.Current(), E, c->d_col, c->d_w);
    HIP_TRY(hipGetLastError());
    // row_ptr = exclusive prefix sum of the degrees (u64).
    HIP_TRY(hipMemsetAsync(c->d_row_ptr, 0, 8, c->stream));
    size_t tmp2 = 0;
    HIP_TRY(hipcub::DeviceScan::InclusiveSum(nullptr, tmp2, c->d_deg, c->d_row_ptr + 1, (uint64_t)V,
                                             c->stream));
    void *d_tmp2 = nullptr;
    HIP_TRY(hipMalloc(&d_tmp2, tmp2 + 16));
    HIP_TRY(hipcub::DeviceScan::InclusiveSum(d_tmp2, tmp2, c->d_deg, c->d_row_ptr + 1, (uint64_t)V,
                                             c->stream));
    k_vrec<<<grid_for(V, 256, 65535), 256, 0, c->stream>>>(c->d_row_ptr, V, c->d_vrec, c->d_deg);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(c->stream));
    dfree(d_tmp);
    dfree(d_tmp2);
    dfree(vals_a);
    dfree(vals_b);
    dfree(keys_b);
    dfree(c->d_esrc);
    dfree(c->d_edst);
    dfree(c->d_ew);
    c->d_esrc = c->d_edst = nullptr;
    c->d_ew = nullptr;
    return hipSuccess;
}

// Vocabulary: vertices by degree descending, ties by id ascending (stable
// sort of ids 0..V-1 on the complemented degree).  The reference's qsort
// leaves ties unspecified (src/word2vec.cu:737).
static hipError_t sort_vocab(felog_ctx *c) {
    const uint32_t V = c->V;
    uint32_t *keys_a = nullptr, *keys_b = nullptr, *ids_a = nullptr, *ids_b = nullptr;
    HIP_TRY(hipMalloc(&keys_a, (uint64_t)V * 4));
    HIP_TRY(hipMalloc(&keys_b, (uint64_t)V * 4));
    HIP_TRY(hipMalloc(&ids_a, (uint64_t)V * 4));
    HIP_TRY(hipMalloc(&ids_b, (uint64_t)V * 4));
    std::vector<uint32_t> iota(V);
    for (uint32_t i = 0; i < V; ++i) iota[i] = i;
    HIP_TRY(hipMemcpyAsync(ids_a, iota.data(), (uint64_t)V * 4, hipMemcpyHostToDevice, c->stream));
    // keys = ~deg
    std::vector<uint32_t> deg(V);
    HIP_TRY(hipMemcpyAsync(deg.data(), c->d_deg, (uint64_t)V * 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (uint32_t i = 0; i < V; ++i) deg[i] = ~deg[i];
    HIP_TRY(hipMemcpyAsync(keys_a, deg.data(), (uint64_t)V * 4, hipMemcpyHostToDevice, c->stream));
    hipcub::DoubleBuffer<uint32_t> keys(keys_a, keys_b), ids(ids_a, ids_b);
    size_t tmp = 0;
    HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, keys, ids, (uint64_t)V, 0, 32, c->stream));
    void *d_tmp = nullptr;
    HIP_TRY(hipMalloc(&d_tmp, tmp + 16));
    HIP_TRY(hipcub::DeviceRadixSort::SortPairs(d_tmp, tmp, keys, ids, (uint64_t)V, 0, 32, c->stream));
    HIP_TRY(hipMemcpyAsync(c->d_vocab_order, ids.Current(), (uint64_t)V * 4, hipMemcpyDeviceToDevice,
                           c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    dfree(d_tmp);
    dfree(keys_a);
    dfree(keys_b);
    dfree(ids_a);
    dfree(ids_b);
    return hipSuccess;
}

hipError_t build_alias(felog_ctx *c) {
    const uint64_t E = c->E;
    HIP_TRY(hipMalloc(&c->d_erec, E * sizeof(EdgeRec) + 16));
    double *p = nullptr;
    uint32_t *stk = nullptr;
    HIP_TRY(hipMalloc(&p, E * 8 + 8));
    HIP_TRY(hipMalloc(&stk, E * 4 + 4));
    k_alias<<<grid_for(c->V, 64, 1u << 20), 64, 0, c->stream>>>(c->d_vocab_order, c->V, c->d_row_ptr,
                                                                c->d_w, c->d_col, p, stk, c->d_erec);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMalloc(&c->d_vwalk, (uint64_t)c->V * sizeof(VertexRec) + 8));
    k_vwalk<<<grid_for(c->V, 256, 65535), 256, 0, c->stream>>>(c->d_vrec, c->d_col, c->V, c->d_vwalk);
    HIP_TRY(hipGetLastError());
    if (E) {
        k_edge_vr<<<grid_for(E, 256, 1u << 16), 256, 0, c->stream>>>(c->d_vwalk, E, c->d_erec);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    dfree(p);
    dfree(stk);
    return hipSuccess;
}

// Host Vose over V buckets (same pairing rule as k_alias), for the unigram
// negative sampler: p_v proportional to deg^0.75 over vocabulary vertices
// (src/word2vec.cu:674-694 draws from the same law through a 1e8 table).
static void host_alias(const std::vector<double> &wts, std::vector<uint32_t> &thr,
                       std::vector<uint32_t> &ali) {
    const uint32_t n = (uint32_t)wts.size();
    thr.assign(n, 0xFFFFFFFFu);
    ali.resize(n);
    double sum = 0.0;
    for (uint32_t j = 0; j < n; ++j) sum = sum + (wts[j] > 0.0 ? wts[j] : 0.0);
    std::vector<double> p(n);
    std::vector<uint32_t> stk(n);
    uint32_t ns = 0, nl = 0;
    for (uint32_t j = 0; j < n; ++j) {
        const double pj = sum > 0.0 ? ((wts[j] > 0.0 ? wts[j] : 0.0) * (double)n) / sum : 1.0;
        p[j] = pj;
        if (pj < 1.0) stk[ns++] = j;
        else stk[n - 1 - nl++] = j;
    }
    auto q = [](double x) -> uint32_t {
        if (x >= 1.0) return 0xFFFFFFFFu;
        if (!(x > 0.0)) return 0u;
        return (uint32_t)(x * 4294967296.0);
    };
    while (ns > 0 && nl > 0) {
        const uint32_t s = stk[--ns];
        const uint32_t l = stk[n - nl];
        --nl;
        thr[s] = q(p[s]);
        ali[s] = l;
        const double pl = (p[l] + p[s]) - 1.0;
        p[l] = pl;
        if (pl < 1.0) stk[ns++] = l;
        else stk[n - 1 - nl++] = l;
    }
    while (nl > 0) {
        const uint32_t l = stk[n - nl];
        --nl;
        thr[l] = 0xFFFFFFFFu;
        ali[l] = l;
    }
    while (ns > 0) {
        const uint32_t s = stk[--ns];
        thr[s] = 0xFFFFFFFFu;
        ali[s] = s;
    }
}

hipError_t build_vocab_and_tables(felog_ctx *c, std::vector<uint32_t> &deg) {
    const uint32_t V = c->V;
    HIP_TRY(hipMalloc(&c->d_vocab_order, (uint64_t)V * 4));
    HIP_TRY(hipMalloc(&c->d_vocab_rank, (uint64_t)V * 4));
    HIP_TRY(hipMalloc(&c->d_in_vocab, (uint64_t)V));
    HIP_TRY(hipMalloc(&c->d_uni, (uint64_t)V * 8));
    HIP_TRY(hipMalloc(&c->d_keep_thr, (uint64_t)V * 4));
    HIP_TRY(sort_vocab(c));
    std::vector<uint32_t> order(V), rank(V);
    deg.resize(V);
    HIP_TRY(hipMemcpy(order.data(), c->d_vocab_order, (uint64_t)V * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(deg.data(), c->d_deg, (uint64_t)V * 4, hipMemcpyDeviceToHost));
    for (uint32_t i = 0; i < V; ++i) rank[order[i]] = i;
    // src/word2vec.cu:740-747: vocab keeps degree >= min-count plus index 0;
    // train_words sums the kept degrees.
    std::vector<uint8_t> inv(V);
    uint64_t vs = 0, tw = 0;
    for (uint32_t v = 0; v < V; ++v) {
        inv[v] = (deg[v] >= c->opt.min_count || rank[v] == 0) ? 1 : 0;
        if (inv[v]) { ++vs; tw += deg[v]; }
    }
    c->vocab_size = vs;
    c->train_words = tw;
    // unigram^0.75 over the vocabulary
    std::vector<double> uw(V);
    for (uint32_t v = 0; v < V; ++v) uw[v] = inv[v] ? std::pow((double)deg[v], 0.75) : 0.0;
    std::vector<uint32_t> uthr, uali;
    host_alias(uw, uthr, uali);
    // subsampling keep thresholds (src/word2vec.cu:180-212): keep with
    // probability min(1, (sqrt(c/sT)+1)*sT/c), c = degree, sT = sample*train_words.
    std::vector<uint32_t> keep(V, 0xFFFFFFFFu);
    const double sT = (double)c->opt.sample * (double)tw;
    if (c->opt.sample > 0.f && tw > 0) {
        for (uint32_t v = 0; v < V; ++v) {
            const double cnt = (double)deg[v];
            if (cnt <= 0.0) continue;
            double prob = (std::sqrt(cnt / sT) + 1.0) * (sT / cnt);
            if (prob < 1.0) keep[v] = prob <= 0.0 ? 0u : (uint32_t)(prob * 4294967296.0);
        }
    }
    // one gather per token in the SGNS filter: vertices outside the
    // vocabulary are never kept (threshold 0); one per negative: the
    // sampler's threshold and alias side by side
    for (uint32_t v = 0; v < V; ++v)
        if (!inv[v]) keep[v] = 0;
    std::vector<uint2> uni(V);
    for (uint32_t v = 0; v < V; ++v) uni[v] = make_uint2(uthr[v], uali[v]);
    HIP_TRY(hipMemcpy(c->d_vocab_rank, rank.data(), (uint64_t)V * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->d_in_vocab, inv.data(), (uint64_t)V, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->d_uni, uni.data(), (uint64_t)V * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->d_keep_thr, keep.data(), (uint64_t)V * 4, hipMemcpyHostToDevice));
    // n log2 n table for the walk statistics
    const uint32_t L = c->opt.max_walk_len;
    std::vector<double> T(L + 2), dT(L + 1);
    for (uint32_t n = 0; n < L + 2; ++n) T[n] = n <= 1 ? 0.0 : (double)n * std::log2((double)n);
    for (uint32_t f = 0; f < L + 1; ++f) dT[f] = T[f + 1] - T[f];
    HIP_TRY(hipMalloc(&c->d_T, (L + 2) * 8));
    HIP_TRY(hipMalloc(&c->d_dT, (L + 1) * 8));
    HIP_TRY(hipMemcpy(c->d_T, T.data(), (L + 2) * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->d_dT, dT.data(), (L + 1) * 8, hipMemcpyHostToDevice));
    return hipSuccess;
}

hipError_t init_tables(felog_ctx *c) {
    const uint64_t n = (uint64_t)c->V * c->dp;
    const size_t es = c->opt.precision == FELOG_BF16 ? 2 : 4;
    // one allocation: syn1neg directly follows syn0, so a row of either
    // table is addressed from syn0 by one row index (syn1neg row v = V + v)
    if (!c->d_syn0) {
        HIP_TRY(hipMalloc(&c->d_syn0, 2 * n * es));
        c->d_syn1 = (unsigned char *)c->d_syn0 + n * es;
    }
    Affine jump[64];
    unsigned long long a = 25214903917ull, cc = 11ull;
    for (int k = 0; k < 64; ++k) {
        jump[k] = Affine{a, cc};
        cc = a * cc + cc;  // f^(2m) = f^m o f^m
        a = a * a;
    }
    Affine *d_jump = nullptr;
    HIP_TRY(hipMalloc(&d_jump, sizeof(jump)));
    HIP_TRY(hipMemcpyAsync(d_jump, jump, sizeof(jump), hipMemcpyHostToDevice, c->stream));
    if (es == 4)
        k_init_syn0<float><<<grid_for(n, 256, 1u << 20), 256, 0, c->stream>>>(
            c->d_vocab_rank, c->V, c->d, c->dp, d_jump, (float *)c->d_syn0);
    else
        k_init_syn0<uint16_t><<<grid_for(n, 256, 1u << 20), 256, 0, c->stream>>>(
            c->d_vocab_rank, c->V, c->d, c->dp, d_jump, (uint16_t *)c->d_syn0);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemsetAsync(c->d_syn1, 0, n * es, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    dfree(d_jump);
    return hipSuccess;
}

}  // namespace felog
