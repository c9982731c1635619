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
// include/walk.hpp:87-103, but never builds one): an exact integer
// construction, parallel over columns (below).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <rocprim/iterator/counting_iterator.hpp>
#include <rocprim/iterator/transform_iterator.hpp>
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

// One thread per 8 consecutive elements of a row (dp is a multiple of 32):
// one jump to the first element's LCG state, then single steps.
template <typename T>
__global__ void k_init_syn0(const uint32_t *__restrict__ rank, uint32_t V, uint32_t d, uint32_t dp,
                            const Affine *__restrict__ jump, T *__restrict__ syn0) {
    const uint32_t per_row = dp / 8;
    const uint64_t total = (uint64_t)V * per_row;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        const uint32_t v = (uint32_t)(i / per_row), b0 = (uint32_t)(i % per_row) * 8;
        float val[8];
        unsigned long long x = 0;
        if (b0 < d) x = lcg_nth(jump, (uint64_t)rank[v] * d + b0 + 1);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            val[k] = 0.f;
            if (b0 + k < d) {
                val[k] = (float)(((double)(x & 0xFFFFull) / 65536.0 - 0.5) / (double)d);
                x = x * 25214903917ull + 11ull;
            }
        }
        T *row = syn0 + (uint64_t)v * dp + b0;
        if constexpr (sizeof(T) == 4) {
            reinterpret_cast<float4 *>(row)[0] = make_float4(val[0], val[1], val[2], val[3]);
            reinterpret_cast<float4 *>(row)[1] = make_float4(val[4], val[5], val[6], val[7]);
        } else {
            uint32_t p[4];
#pragma unroll
            for (int k = 0; k < 4; ++k)
                p[k] = f32x2_to_bf16x2_rne(val[2 * k], val[2 * k + 1]);
            reinterpret_cast<uint4 *>(row)[0] = make_uint4(p[0], p[1], p[2], p[3]);
        }
    }
}

inline unsigned grid_for(uint64_t n, unsigned block, unsigned cap) {
    uint64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (unsigned)g;
}

// ---------------------------------------------------------------------------
// Exact alias tables over many weight segments at once (a vertex's CSR row,
// or the whole vocabulary for the negative sampler).  For a segment of n
// columns with weights w_j (negatives count as 0):
//   S    canonical FP64 sum: 4096-column chunks, each the xor-butterfly of 64
//        lane partials (lane l sums columns l, l+64, ... in order); the
//        chunk sums are combined by the same rule;
//   q_j  trunc(((w_j n) / S) 2^32), an integer share of n 2^32; the
//        remainder goes to the largest q (lowest column on ties); S = 0
//        gives every column exactly 2^32;
//   pairing  the integer sweep: light columns (q < 2^32) in order take the
//        heavy column (q >= 2^32) current at their position as alias, and
//        a heavy column whose residual falls below 2^32 becomes a bucket
//        aliased to the next heavy column.  With the light deficits
//        d = 2^32 - q and heavy surpluses s = q - 2^32 as prefix sums D, S
//        (per segment, in column order) the sweep's state at any column is
//        a function of D and S, so every column is paired independently:
//          light with exclusive deficit prefix D: alias = the first heavy
//            column whose inclusive surplus prefix is >= D (D = 0: the
//            first heavy column), threshold q;
//          heavy with inclusive surplus prefix S: if some light's
//            inclusive deficit prefix exceeds S (and a later heavy exists),
//            threshold 2^32 + S - (the first such prefix), alias = the next
//            heavy column; else a whole bucket.
// All integer steps are exact, so any evaluation order gives the tables of
// the sequential sweep bit for bit.
// ---------------------------------------------------------------------------
constexpr uint64_t kUnit = 1ull << 32;
constexpr uint64_t kHeavyBit = 1ull << 63;
constexpr uint32_t kChunk = kHubChunk;

__device__ __forceinline__ double wclamp(float x) { return x > 0.f ? (double)x : 0.0; }
__device__ __forceinline__ double wclamp(double x) { return x > 0.0 ? x : 0.0; }

// The xor butterfly over the 64 lanes (every lane ends with the same sum).
__device__ __forceinline__ double butterfly(double p) {
#pragma clang fp contract(off)
    for (int o = 32; o > 0; o >>= 1) p = p + __shfl_xor(p, o);
    return p;
}

// Lane partial of the canonical sum over x[0, n) (n <= 4096 for a chunk).
template <typename W>
__device__ __forceinline__ double lane_partial(const W *__restrict__ x, uint64_t n, uint32_t lane) {
#pragma clang fp contract(off)
    double p = 0.0;
    for (uint64_t j = lane; j < n; j += 64) p = p + wclamp(x[j]);
    return p;
}

__device__ __forceinline__ uint64_t share(double x, double dn, double S) {
#pragma clang fp contract(off)
    return (uint64_t)(((x * dn) / S) * 4294967296.0);
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t x) {
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    return x;
}

// (q, j) of the larger q, the lower j on ties.
__device__ __forceinline__ void max_lowest(uint64_t &q, uint64_t &j, uint64_t q2, uint64_t j2) {
    if (q2 > q || (q2 == q && j2 < j)) {
        q = q2;
        j = j2;
    }
}

__device__ __forceinline__ void wave_argmax(uint64_t &q, uint64_t &j) {
    for (int o = 32; o > 0; o >>= 1) max_lowest(q, j, __shfl_xor(q, o), __shfl_xor(j, o));
}

// Encoded column: light -> its deficit, heavy -> its surplus | kHeavyBit.
__device__ __forceinline__ uint64_t encode(uint64_t q) { return q < kUnit ? kUnit - q : (q - kUnit) | kHeavyBit; }

// The butterfly within 16-lane groups.  For a segment of at most 16
// columns held by lanes 0..15, the 64-lane butterfly's first two steps add
// +0 (the other lanes hold none of its columns), so this is its sum bit for
// bit.
__device__ __forceinline__ double butterfly16(double p) {
#pragma clang fp contract(off)
    for (int o = 8; o > 0; o >>= 1) p = p + __shfl_xor(p, o);
    return p;
}

// Shares and the encoded columns of the segment [b, b + n) from its sum S:
// lanes step through the columns (lane l: l, l + step, ...), red() reduces
// over the lanes that share the segment.
template <typename W, typename Sum, typename Arg>
__device__ __forceinline__ void encode_segment(const W *__restrict__ w, uint64_t *__restrict__ val, uint64_t b,
                                               uint64_t n, double S, uint32_t l, uint32_t step, Sum sum_u64,
                                               Arg argmax) {
    const double dn = (double)n;
    uint64_t sq = 0, qm = 0, jm = ~0ull;
    if (S > 0.0) {
        for (uint64_t j = l; j < n; j += step) {
            const uint64_t q = share(wclamp(w[b + j]), dn, S);
            sq += q;
            max_lowest(qm, jm, q, j);
        }
        sq = sum_u64(sq);
        argmax(qm, jm);
    }
    const uint64_t rem = n * kUnit - sq;
    for (uint64_t j = l; j < n; j += step) {
        uint64_t q = kUnit;
        if (S > 0.0) q = share(wclamp(w[b + j]), dn, S) + (j == jm ? rem : 0ull);
        val[b + j] = encode(q);
    }
}

// Segments of at most one chunk.  A wavefront takes four consecutive
// segments: each of at most 16 columns is done by one 16-lane group, the
// others one after the other by the whole wavefront (sum, shares and the
// encoded columns, one pass per step).
template <typename W>
__global__ __launch_bounds__(256) void k_alias_small(const uint64_t *__restrict__ seg, uint32_t nseg,
                                                     const W *__restrict__ w, uint64_t *__restrict__ val) {
    const uint32_t lane = threadIdx.x & 63, grp = lane >> 4, l16 = lane & 15;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
    auto sum16 = [](uint64_t x) {
        for (int o = 8; o > 0; o >>= 1) x += __shfl_xor(x, o);
        return x;
    };
    auto argmax16 = [](uint64_t &q, uint64_t &j) {
        for (int o = 8; o > 0; o >>= 1) max_lowest(q, j, __shfl_xor(q, o), __shfl_xor(j, o));
    };
    for (uint32_t s0 = 4 * wave; s0 < nseg; s0 += 4 * nwaves) {
        const uint32_t s = s0 + grp;
        uint64_t b = 0, n = 0;
        if (s < nseg) {
            b = seg[s];
            n = seg[s + 1] - b;
        }
        if (n <= 16) {
            if (n > 0) {
                const double S = butterfly16(l16 < n ? wclamp(w[b + l16]) : 0.0);
                encode_segment(w, val, b, n, S, l16, 16, sum16, argmax16);
            }
        }
        // the larger segments of the four, by the whole wavefront
        for (uint64_t m = __ballot(l16 == 0 && n > 16 && n <= kChunk); m; m &= m - 1) {
            const uint32_t g = (uint32_t)__builtin_ctzll(m) >> 4;
            const uint64_t bg = seg[s0 + g], ng = seg[s0 + g + 1] - bg;
            const double S = butterfly(lane_partial(w + bg, ng, lane));
            encode_segment(w, val, bg, ng, S, lane, 64, [](uint64_t x) { return wave_sum_u64(x); },
                           [](uint64_t &q, uint64_t &j) { wave_argmax(q, j); });
        }
    }
}

// Larger segments ("hubs", hub k = segment hub[k]) are cut into chunks;
// chunk t of the launch belongs to the hub k with cstart[k] <= t <
// cstart[k+1] (hub_of_chunk).
template <typename W>
__global__ __launch_bounds__(256) void k_alias_chunk_sum(const uint64_t *__restrict__ seg,
                                                         const uint32_t *__restrict__ hub,
                                                         const uint64_t *__restrict__ cstart, uint32_t nh,
                                                         const W *__restrict__ w, double *__restrict__ csum) {
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t nchunks = cstart[nh];
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    for (uint64_t t = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; t < nchunks; t += nwaves) {
        const uint32_t k = hub_of_chunk(cstart, nh, t);
        const uint64_t b = seg[hub[k]], n = seg[hub[k] + 1] - b;
        const uint64_t c0 = (t - cstart[k]) * kChunk;
        const uint64_t cn = n - c0 < kChunk ? n - c0 : kChunk;
        const double p = butterfly(lane_partial(w + b + c0, cn, lane));
        if (lane == 0) csum[t] = p;
    }
}

// One wavefront per hub: the canonical sum of its chunk sums.
__global__ __launch_bounds__(256) void k_alias_hub_sum(const uint64_t *__restrict__ cstart, uint32_t nh,
                                                       const double *__restrict__ csum, double *__restrict__ hsum) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
    for (uint32_t k = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; k < nh; k += nwaves) {
        const double p = butterfly(lane_partial(csum + cstart[k], cstart[k + 1] - cstart[k], lane));
        if (lane == 0) hsum[k] = p;
    }
}

// Per chunk: the sum of its shares and its largest share (lowest column).
template <typename W>
__global__ __launch_bounds__(256) void k_alias_chunk_q(const uint64_t *__restrict__ seg,
                                                       const uint32_t *__restrict__ hub,
                                                       const uint64_t *__restrict__ cstart, uint32_t nh,
                                                       const W *__restrict__ w, const double *__restrict__ hsum,
                                                       uint64_t *__restrict__ cq) {
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t nchunks = cstart[nh];
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    for (uint64_t t = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; t < nchunks; t += nwaves) {
        const uint32_t k = hub_of_chunk(cstart, nh, t);
        const uint64_t b = seg[hub[k]], n = seg[hub[k] + 1] - b;
        const uint64_t c0 = (t - cstart[k]) * kChunk;
        const uint64_t cn = n - c0 < kChunk ? n - c0 : kChunk;
        const double S = hsum[k], dn = (double)n;
        uint64_t sq = 0, qm = 0, jm = ~0ull;
        if (S > 0.0) {
            for (uint64_t j = c0 + lane; j < c0 + cn; j += 64) {
                const uint64_t q = share(wclamp(w[b + j]), dn, S);
                sq += q;
                max_lowest(qm, jm, q, j);
            }
            sq = wave_sum_u64(sq);
            wave_argmax(qm, jm);
        }
        if (lane == 0) {
            cq[3 * t] = sq;
            cq[3 * t + 1] = qm;
            cq[3 * t + 2] = jm;
        }
    }
}

// One wavefront per hub: remainder and the column that takes it.
__global__ __launch_bounds__(256) void k_alias_hub_q(const uint64_t *__restrict__ seg, const uint32_t *__restrict__ hub,
                                                     const uint64_t *__restrict__ cstart, uint32_t nh,
                                                     const uint64_t *__restrict__ cq, uint64_t *__restrict__ hrem) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
    for (uint32_t k = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; k < nh; k += nwaves) {
        uint64_t sq = 0, qm = 0, jm = ~0ull;
        for (uint64_t t = cstart[k] + lane; t < cstart[k + 1]; t += 64) {
            sq += cq[3 * t];
            max_lowest(qm, jm, cq[3 * t + 1], cq[3 * t + 2]);
        }
        sq = wave_sum_u64(sq);
        wave_argmax(qm, jm);
        if (lane == 0) {
            const uint64_t n = seg[hub[k] + 1] - seg[hub[k]];
            hrem[2 * k] = n * kUnit - sq;
            hrem[2 * k + 1] = jm;
        }
    }
}

template <typename W>
__global__ __launch_bounds__(256) void k_alias_chunk_val(const uint64_t *__restrict__ seg,
                                                         const uint32_t *__restrict__ hub,
                                                         const uint64_t *__restrict__ cstart, uint32_t nh,
                                                         const W *__restrict__ w, const double *__restrict__ hsum,
                                                         const uint64_t *__restrict__ hrem,
                                                         uint64_t *__restrict__ val) {
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t nchunks = cstart[nh];
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    for (uint64_t t = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; t < nchunks; t += nwaves) {
        const uint32_t k = hub_of_chunk(cstart, nh, t);
        const uint64_t b = seg[hub[k]], n = seg[hub[k] + 1] - b;
        const uint64_t c0 = (t - cstart[k]) * kChunk;
        const uint64_t cn = n - c0 < kChunk ? n - c0 : kChunk;
        const double S = hsum[k], dn = (double)n;
        const uint64_t rem = hrem[2 * k], jm = hrem[2 * k + 1];
        for (uint64_t j = c0 + lane; j < c0 + cn; j += 64) {
            uint64_t q = kUnit;
            if (S > 0.0) q = share(wclamp(w[b + j]), dn, S) + (j == jm ? rem : 0ull);
            val[b + j] = encode(q);
        }
    }
}

__global__ void k_alias_hub_chunks(const uint64_t *__restrict__ seg, const uint32_t *__restrict__ hub, uint32_t nh,
                                   uint64_t *__restrict__ cnt) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k <= nh; k += stride)
        cnt[k] = k < nh ? (seg[hub[k] + 1] - seg[hub[k]] + kChunk - 1) / kChunk : 0;
}

struct LightDeficit {
    __host__ __device__ uint64_t operator()(uint64_t v) const { return (v & kHeavyBit) ? 0ull : v; }
};
struct HeavySurplus {
    __host__ __device__ uint64_t operator()(uint64_t v) const { return (v & kHeavyBit) ? (v & ~kHeavyBit) : 0ull; }
};
struct LightCount {
    __host__ __device__ uint32_t operator()(uint64_t v) const { return (v & kHeavyBit) ? 0u : 1u; }
};
struct IsHub {
    const uint64_t *seg;
    __host__ __device__ bool operator()(uint32_t s) const { return seg[s + 1] - seg[s] > kChunk; }
};

// First index i in [lo, hi) with pred(i) true (pred false then true).
template <typename P>
__device__ __forceinline__ uint64_t first_true(uint64_t lo, uint64_t hi, P pred) {
    while (lo < hi) {
        const uint64_t mid = lo + (hi - lo) / 2;
        if (pred(mid)) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

// Pairing of every column (flat over the columns of all segments).
// seg_of[e] is the segment of column e (null: one segment).  out(e, j,
// threshold, alias column) stores the result.
template <typename Out>
__global__ __launch_bounds__(256) void k_alias_pair(const uint64_t *__restrict__ seg, const uint32_t *__restrict__ seg_of,
                                                    uint64_t total, const uint64_t *__restrict__ val,
                                                    const uint64_t *__restrict__ LD, const uint64_t *__restrict__ SH,
                                                    const uint32_t *__restrict__ NL, Out out) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
        const uint32_t s = seg_of ? seg_of[e] : 0u;
        const uint64_t b = seg[s], end = seg[s + 1];
        const uint64_t bL = b ? LD[b - 1] : 0ull, bS = b ? SH[b - 1] : 0ull;
        const uint32_t bN = b ? NL[b - 1] : 0u;
        // heavy columns in [b, i]
        auto heavies = [&](uint64_t i) -> uint64_t { return (i - b + 1) - (uint64_t)(NL[i] - bN); };
        const uint64_t v = val[e];
        uint32_t thr = 0xFFFFFFFFu;
        uint64_t ali = e;
        if (!(v & kHeavyBit)) {
            const uint64_t D = (LD[e] - bL) - v;  // deficits of the lights before e
            ali = D == 0 ? first_true(b, end, [&](uint64_t i) { return heavies(i) >= 1; })
                         : first_true(b, end, [&](uint64_t i) { return SH[i] - bS >= D; });
            thr = (uint32_t)(kUnit - v);
        } else {
            const uint64_t S = SH[e] - bS;
            const uint64_t h = heavies(e);  // 1-based rank of e among the heavy columns
            const uint64_t nH = heavies(end - 1);
            if (h < nH && LD[end - 1] - bL > S) {
                const uint64_t i = first_true(b, end, [&](uint64_t x) { return LD[x] - bL > S; });
                thr = (uint32_t)(kUnit + S - (LD[i] - bL));
                ali = first_true(e + 1, end, [&](uint64_t x) { return heavies(x) >= h + 1; });
            }
        }
        out(e, e - b, thr, ali - b);
    }
}

// Outputs: the walk's edge records, or the negative sampler's {thr, alias}.
struct EdgeOut {
    const uint64_t *seg;
    const uint32_t *col;
    const VertexRec *vwalk;
    EdgeRec *erec;
    __device__ void operator()(uint64_t e, uint64_t j, uint32_t thr, uint64_t a) const {
        const uint64_t b = e - j;
        const uint32_t c0 = col[e], c1 = col[b + a];
        erec[e] = EdgeRec{thr, c0, c1, (uint32_t)a, vwalk[c0], vwalk[c1]};
    }
};
struct PairOut {
    uint2 *out;
    __device__ void operator()(uint64_t e, uint64_t, uint32_t thr, uint64_t a) const {
        out[e] = make_uint2(thr, (uint32_t)a);
    }
};

// Sort keys of the vocabulary (complemented degree, ids), the totals
// cleared and the one-segment offsets {0, V} of the negative sampler.
__global__ void k_vocab_keys(const uint32_t *__restrict__ deg, uint32_t V, uint32_t *__restrict__ keys,
                             uint32_t *__restrict__ ids, uint64_t *__restrict__ totals, uint64_t *__restrict__ seg01) {
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0) {
        totals[0] = totals[1] = 0;
        seg01[0] = 0;
        seg01[1] = V;
    }
    for (uint32_t v = t; v < V; v += stride) {
        keys[v] = ~deg[v];
        ids[v] = v;
    }
}

// rank (a scatter through the sorted order), and per vertex in id order
// the vocabulary flag (degree >= min-count, or rank 0), the negative
// sampler's weight deg^0.75 = sqrt(deg) * sqrt(sqrt(deg)) (correctly
// rounded square roots: the host computes the same bits) and the totals
// (vocabulary size, train words = sum of the vocabulary's degrees).
__global__ __launch_bounds__(256) void k_vocab_rank(const uint32_t *__restrict__ order, const uint32_t *__restrict__ deg, uint32_t V,
                             uint32_t min_count, uint32_t *__restrict__ rank, uint8_t *__restrict__ inv,
                             double *__restrict__ uw, unsigned long long *totals) {
#pragma clang fp contract(off)
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t top = order[0];
    unsigned long long vs = 0, tw = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < V; i += stride) {
        rank[order[i]] = i;
        const uint32_t d = deg[i];
        const bool in = d >= min_count || i == top;
        inv[i] = in ? 1 : 0;
        const double x = (double)d;
        uw[i] = in ? __builtin_sqrt(x) * __builtin_sqrt(__builtin_sqrt(x)) : 0.0;
        vs += in ? 1ull : 0ull;
        tw += in ? d : 0ull;
    }
    // one pair of atomics per workgroup: adds to one address serialise
    for (int o = 32; o > 0; o >>= 1) {
        vs += __shfl_xor(vs, o);
        tw += __shfl_xor(tw, o);
    }
    __shared__ unsigned long long part[2][4];
    const uint32_t wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        part[0][wv] = vs;
        part[1][wv] = tw;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        vs = part[0][0] + part[0][1] + part[0][2] + part[0][3];
        tw = part[1][0] + part[1][1] + part[1][2] + part[1][3];
        if (vs | tw) {
            atomicAdd(&totals[0], vs);
            atomicAdd(&totals[1], tw);
        }
    }
}

// Subsampling keep thresholds: keep with probability
// min(1, (sqrt(c/sT) + 1) sT/c), c = degree, sT = sample * train words, on
// the u32 scale; 0 outside the vocabulary (one gather per token decides).
__global__ void k_keep(const uint32_t *__restrict__ deg, const uint8_t *__restrict__ inv, uint32_t V, float sample,
                       const uint64_t *__restrict__ train_words, uint32_t *__restrict__ keep) {
#pragma clang fp contract(off)
    const uint64_t tw = *train_words;
    const double sT = (double)sample * (double)tw;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += stride) {
        uint32_t k = 0xFFFFFFFFu;
        const double cnt = (double)deg[v];
        if (sample > 0.f && tw > 0 && cnt > 0.0) {
            const double prob = (__builtin_sqrt(cnt / sT) + 1.0) * (sT / cnt);
            if (prob < 1.0) k = prob <= 0.0 ? 0u : (uint32_t)(prob * 4294967296.0);
        }
        keep[v] = inv[v] ? k : 0u;
    }
}

// row_ptr from the sorted sources: entry e starts every vertex v with
// key[e-1] < v <= key[e] (key[-1] = -1, key[E] = V), so vertices without
// edges get empty rows.
__global__ void k_row_ptr(const uint32_t *__restrict__ key, uint64_t E, uint32_t V, uint64_t *__restrict__ row_ptr) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e <= E; e += stride) {
        const uint64_t lo = e ? (uint64_t)key[e - 1] + 1 : 0;
        const uint64_t hi = e < E ? (uint64_t)key[e] : (uint64_t)V;
        for (uint64_t v = lo; v <= hi; ++v) row_ptr[v] = e;
    }
}

}  // namespace

#define HIP_TRY(x)                              \
    do {                                        \
        hipError_t e_ = (x);                    \
        if (e_ != hipSuccess) return e_;        \
    } while (0)

hipError_t Scratch::reserve(size_t bytes) {
    const hipError_t e = hipMalloc(&base, bytes);
    if (e != hipSuccess) {
        base = nullptr;
        return e;
    }
    cap = bytes;
    top = 0;
    return hipSuccess;
}

hipError_t Scratch::take(void **p, size_t bytes) {
    const size_t o = (top + 255) & ~(size_t)255;
    if (base && o + bytes <= cap) {
        *p = base + o;
        top = o + bytes;
        return hipSuccess;
    }
    const hipError_t e = hipMalloc(p, bytes);
    if (e == hipSuccess) extra.push_back(*p);
    return e;
}

void Scratch::done(felog_ctx *c) {
    for (void *p : extra) dfree(p);
    extra.clear();
    if (c) dkeep(c, base, cap);
    else dfree(base);
    base = nullptr;
    cap = top = 0;
}

hipError_t dalloc(felog_ctx *c, void **p, size_t bytes) {
    size_t best = c->spare.size();
    for (size_t i = 0; i < c->spare.size(); ++i) {
        const size_t b = c->spare[i].second;
        if (b >= bytes && b / 2 <= bytes && (best == c->spare.size() || b < c->spare[best].second)) best = i;
    }
    if (best < c->spare.size()) {
        *p = c->spare[best].first;
        c->spare.erase(c->spare.begin() + (ptrdiff_t)best);
        return hipSuccess;
    }
    return hipMalloc(p, bytes);
}

// The largest of the three stages: CSR (second key buffer, two value
// buffers, sort temporary), vocabulary (the sorted sources kept, the
// vocabulary sort and the sampler's alias build over V) and the edge alias
// build (the sorted sources, 28 bytes per entry, scan temporaries and chunk
// plans, which the slack covers).
size_t build_scratch_bytes(felog_ctx *c) {
    const uint64_t E = c->E, V = c->V;
    auto al = [](uint64_t b) -> size_t { return (size_t)((b + 255) & ~(uint64_t)255); };
    int bits = 1;
    while (bits < 32 && (1ull << bits) < V) ++bits;
    size_t sort_e = 0, sort_v = 0;
    {
        hipcub::DoubleBuffer<uint32_t> k(nullptr, nullptr);
        hipcub::DoubleBuffer<uint64_t> v(nullptr, nullptr);
        if (hipcub::DeviceRadixSort::SortPairs(nullptr, sort_e, k, v, E, 0, bits, c->stream) != hipSuccess)
            sort_e = 0;
    }
    {
        hipcub::DoubleBuffer<uint32_t> k(nullptr, nullptr), v(nullptr, nullptr);
        if (hipcub::DeviceRadixSort::SortPairs(nullptr, sort_v, k, v, V, 0, 32, c->stream) != hipSuccess)
            sort_v = 0;
    }
    const size_t keys = al(4 * E + 4);
    const size_t slack_e = al(E / 8) + ((size_t)64 << 20), slack_v = al(V / 8) + ((size_t)16 << 20);
    const size_t csr = keys + 2 * al(8 * E + 8) + al(sort_e + 16);
    const size_t vocab = keys + 4 * al(4 * V) + 512 + al(8 * V) + al(sort_v + 16) + 3 * al(8 * V) + 2 * al(4 * V) +
                         slack_v;
    const size_t alias = keys + 3 * al(8 * E) + al(4 * E) + al(4 * V) + slack_e;
    size_t m = csr > vocab ? csr : vocab;
    return m > alias ? m : alias;
}

hipError_t launch_undirect(felog_ctx *c, const uint32_t *src, const uint32_t *dst, const float *w,
                           uint64_t m, bool undirected) {
    const uint64_t E = undirected ? 2 * m : m;
    HIP_TRY(hipMalloc(&c->d_esrc, E * 4 + 4));
    HIP_TRY(hipMalloc(&c->d_edst, E * 4 + 4));
    HIP_TRY(hipMalloc(&c->d_ew, E * 4 + 4));
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

// Exact alias tables of nseg weight segments: seg holds nseg + 1 column
// offsets on the device (total = the last one), w the weights, seg_of the
// segment of every column (null when nseg == 1); out stores each column's
// threshold and alias.  Temporary device memory: 28 bytes per column.  keep
// (when set) receives the chunk plan of the segments longer than one chunk.
template <typename W, typename Out>
static hipError_t alias_segments(felog_ctx *c, Scratch &sc, const uint64_t *seg, uint32_t nseg, uint64_t total,
                                 const W *w, const uint32_t *seg_of, Out out, HubPlan *keep) {
    if (total == 0 || nseg == 0) return hipSuccess;
    hipStream_t st = c->stream;
    const size_t m0 = sc.mark();
    uint64_t *val = nullptr, *LD = nullptr, *SH = nullptr;
    uint32_t *NL = nullptr, *hub = nullptr, *d_nh = nullptr;
    HIP_TRY(sc.take(&val, total * 8));
    HIP_TRY(sc.take(&LD, total * 8));
    HIP_TRY(sc.take(&SH, total * 8));
    HIP_TRY(sc.take(&NL, total * 4));
    // a kept plan outlives the build: its own allocations
    if (keep) HIP_TRY(hipMalloc(&hub, (uint64_t)nseg * 4));
    else HIP_TRY(sc.take(&hub, (uint64_t)nseg * 4));
    HIP_TRY(sc.take(&d_nh, 8));
    k_alias_small<W><<<grid_for((uint64_t)nseg * 64, 256, 1u << 16), 256, 0, st>>>(seg, nseg, w, val);
    HIP_TRY(hipGetLastError());
    // segments longer than one chunk
    size_t tb = 0;
    const rocprim::counting_iterator<uint32_t> ids(0);
    HIP_TRY(hipcub::DeviceSelect::If(nullptr, tb, ids, hub, d_nh, nseg, IsHub{seg}, st));
    size_t tb2 = 0, tb3 = 0, tb4 = 0;
    HIP_TRY(hipcub::DeviceScan::InclusiveSum(nullptr, tb2, rocprim::make_transform_iterator(val, LightDeficit{}), LD,
                                             total, st));
    HIP_TRY(hipcub::DeviceScan::InclusiveSum(nullptr, tb3, rocprim::make_transform_iterator(val, LightCount{}), NL,
                                             total, st));
    HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, tb4, (uint64_t *)nullptr, (uint64_t *)nullptr,
                                             (uint64_t)nseg + 1, st));
    size_t tmax = tb;
    for (size_t x : {tb2, tb3, tb4}) tmax = x > tmax ? x : tmax;
    void *tmp = nullptr;
    HIP_TRY(sc.take(&tmp, tmax + 16));
    tb = tmax;
    HIP_TRY(hipcub::DeviceSelect::If(tmp, tb, ids, hub, d_nh, nseg, IsHub{seg}, st));
    uint32_t nh = 0;
    HIP_TRY(hipMemcpyAsync(&nh, d_nh, 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    uint64_t *cnt = nullptr, *cstart = nullptr, *cq = nullptr, *hrem = nullptr;
    double *csum = nullptr, *hsum = nullptr;
    uint64_t nchunks = 0;
    if (nh) {
        HIP_TRY(sc.take(&cnt, ((uint64_t)nh + 1) * 8));
        if (keep) HIP_TRY(hipMalloc(&cstart, ((uint64_t)nh + 1) * 8));
        else HIP_TRY(sc.take(&cstart, ((uint64_t)nh + 1) * 8));
        k_alias_hub_chunks<<<grid_for((uint64_t)nh + 1, 256, 4096), 256, 0, st>>>(seg, hub, nh, cnt);
        HIP_TRY(hipGetLastError());
        tb = tmax;
        HIP_TRY(hipcub::DeviceScan::ExclusiveSum(tmp, tb, cnt, cstart, (uint64_t)nh + 1, st));
        HIP_TRY(hipMemcpyAsync(&nchunks, cstart + nh, 8, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        HIP_TRY(sc.take(&csum, nchunks * 8));
        HIP_TRY(sc.take(&cq, nchunks * 24));
        HIP_TRY(sc.take(&hsum, (uint64_t)nh * 8));
        HIP_TRY(sc.take(&hrem, (uint64_t)nh * 16));
        const unsigned gc = grid_for(nchunks * 64, 256, 1u << 16), gh = grid_for((uint64_t)nh * 64, 256, 1u << 16);
        k_alias_chunk_sum<W><<<gc, 256, 0, st>>>(seg, hub, cstart, nh, w, csum);
        k_alias_hub_sum<<<gh, 256, 0, st>>>(cstart, nh, csum, hsum);
        k_alias_chunk_q<W><<<gc, 256, 0, st>>>(seg, hub, cstart, nh, w, hsum, cq);
        k_alias_hub_q<<<gh, 256, 0, st>>>(seg, hub, cstart, nh, cq, hrem);
        k_alias_chunk_val<W><<<gc, 256, 0, st>>>(seg, hub, cstart, nh, w, hsum, hrem, val);
        HIP_TRY(hipGetLastError());
    }
    // per-segment prefixes come from global ones minus the value before the
    // segment's first column: every value is an exact integer
    tb = tmax;
    HIP_TRY(hipcub::DeviceScan::InclusiveSum(tmp, tb, rocprim::make_transform_iterator(val, LightDeficit{}), LD,
                                             total, st));
    tb = tmax;
    HIP_TRY(hipcub::DeviceScan::InclusiveSum(tmp, tb, rocprim::make_transform_iterator(val, HeavySurplus{}), SH,
                                             total, st));
    tb = tmax;
    HIP_TRY(hipcub::DeviceScan::InclusiveSum(tmp, tb, rocprim::make_transform_iterator(val, LightCount{}), NL,
                                             total, st));
    k_alias_pair<Out><<<grid_for(total, 256, 1u << 20), 256, 0, st>>>(seg, seg_of, total, val, LD, SH, NL, out);
    HIP_TRY(hipGetLastError());
    if (keep) {
        if (nh) {
            *keep = HubPlan{hub, cstart, nh, nchunks};
        } else {
            dfree(hub);
            dfree(cstart);
        }
    }
    sc.release(m0);
    return hipSuccess;
}

hipError_t build_csr(felog_ctx *c, Scratch &sc, uint32_t **edge_src) {
    const uint64_t E = c->E;
    const uint32_t V = c->V;
    HIP_TRY(hipMalloc(&c->d_row_ptr, (uint64_t)(V + 1) * 8));
    HIP_TRY(hipMalloc(&c->d_deg, (uint64_t)V * 4));
    HIP_TRY(hipMalloc(&c->d_vrec, (uint64_t)V * sizeof(VertexRec)));

    // Stable LSD radix sort of (src, dst|w) keeps each vertex's neighbours in
    // input order.  The second key buffer sits at the bottom of the scratch:
    // it may end up holding the sorted sources, which the alias build reads.
    uint64_t *vals_a = nullptr, *vals_b = nullptr;
    uint32_t *keys_b = nullptr;
    HIP_TRY(sc.take(&keys_b, E * 4 + 4));
    const size_t m0 = sc.mark();
    HIP_TRY(sc.take(&vals_a, E * 8 + 8));
    HIP_TRY(sc.take(&vals_b, E * 8 + 8));
    k_pack<<<grid_for(E, 256, 65535), 256, 0, c->stream>>>(c->d_edst, c->d_ew, E, vals_a);
    HIP_TRY(hipGetLastError());
    // the input's destination and weight blocks (E * 4 + 4 bytes each) are
    // free once packed: they become the CSR's columns and weights
    c->d_col = c->d_edst;
    c->d_w = c->d_ew;
    c->d_edst = nullptr;
    c->d_ew = nullptr;
    int bits = 1;
    while (bits < 32 && (1ull << bits) < (uint64_t)V) ++bits;
    hipcub::DoubleBuffer<uint32_t> keys(c->d_esrc, keys_b);
    hipcub::DoubleBuffer<uint64_t> vals(vals_a, vals_b);
    size_t tmp = 0;
    HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, keys, vals, E, 0, bits, c->stream));
    void *d_tmp = nullptr;
    HIP_TRY(sc.take(&d_tmp, tmp + 16));
    HIP_TRY(hipcub::DeviceRadixSort::SortPairs(d_tmp, tmp, keys, vals, E, 0, bits, c->stream));
    // row_ptr from the sorted sources; degrees and vertex records from it
    k_row_ptr<<<grid_for(E + 1, 256, 65535), 256, 0, c->stream>>>(keys.Current(), E, V, c->d_row_ptr);
    HIP_TRY(hipGetLastError());
    k_unpack<<<grid_for(E, 256, 65535), 256, 0, c->stream>>>(vals.Current(), E, c->d_col, c->d_w);
    HIP_TRY(hipGetLastError());
    k_vrec<<<grid_for(V, 256, 65535), 256, 0, c->stream>>>(c->d_row_ptr, V, c->d_vrec, c->d_deg);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(c->stream));
    // the sorted sources are the vertex of every CSR entry: kept for the
    // alias build (c->d_esrc stays set when it holds them)
    *edge_src = keys.Current();
    if (keys.Current() != c->d_esrc) {
        dkeep(c, c->d_esrc, E * 4 + 4);
        c->d_esrc = nullptr;
    }
    sc.release(m0);
    return hipSuccess;
}

// Edge records (alias columns and both neighbours' walk records, written
// once per entry by the pairing kernel).
hipError_t build_alias(felog_ctx *c, Scratch &sc, const uint32_t *edge_src) {
    const uint64_t E = c->E;
    HIP_TRY(hipMalloc(&c->d_vwalk, (uint64_t)c->V * sizeof(VertexRec) + 8));
    k_vwalk<<<grid_for(c->V, 256, 65535), 256, 0, c->stream>>>(c->d_vrec, c->d_col, c->V, c->d_vwalk);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMalloc(&c->d_erec, E * sizeof(EdgeRec) + 16));
    HIP_TRY(alias_segments(c, sc, c->d_row_ptr, c->V, E, (const float *)c->d_w, edge_src,
                           EdgeOut{c->d_row_ptr, c->d_col, c->d_vwalk, c->d_erec}, &c->hubs));
    return hipSuccess;
}

// Vocabulary (vertices by degree descending, ties by id ascending: a stable
// sort of ids 0..V-1 on the complemented degree; the reference's qsort
// leaves ties unspecified, src/word2vec.cu:737), the vocabulary flags and
// totals (src/word2vec.cu:740-747), the negative sampler over deg^0.75
// (src/word2vec.cu:674-694 draws from the same law through a 1e8 table) and
// the subsampling thresholds (src/word2vec.cu:180-212), all on the device.
hipError_t build_vocab_and_tables(felog_ctx *c, Scratch &sc) {
    const uint32_t V = c->V;
    hipStream_t st = c->stream;
    HIP_TRY(hipMalloc(&c->d_vocab_order, (uint64_t)V * 4));
    HIP_TRY(hipMalloc(&c->d_vocab_rank, (uint64_t)V * 4));
    HIP_TRY(hipMalloc(&c->d_in_vocab, (uint64_t)V));
    HIP_TRY(hipMalloc(&c->d_uni, (uint64_t)V * 8));
    HIP_TRY(hipMalloc(&c->d_keep_thr, (uint64_t)V * 4));
    const size_t m0 = sc.mark();
    uint32_t *keys_a = nullptr, *keys_b = nullptr, *ids_a = nullptr, *ids_b = nullptr;
    uint64_t *totals = nullptr, *seg01 = nullptr;
    double *uw = nullptr;
    HIP_TRY(sc.take(&keys_a, (uint64_t)V * 4));
    HIP_TRY(sc.take(&keys_b, (uint64_t)V * 4));
    HIP_TRY(sc.take(&ids_a, (uint64_t)V * 4));
    HIP_TRY(sc.take(&ids_b, (uint64_t)V * 4));
    HIP_TRY(sc.take(&totals, 16));
    HIP_TRY(sc.take(&seg01, 16));
    HIP_TRY(sc.take(&uw, (uint64_t)V * 8));
    k_vocab_keys<<<grid_for(V, 256, 65535), 256, 0, st>>>(c->d_deg, V, keys_a, ids_a, totals, seg01);
    HIP_TRY(hipGetLastError());
    hipcub::DoubleBuffer<uint32_t> keys(keys_a, keys_b), ids(ids_a, ids_b);
    size_t tmp = 0;
    HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, keys, ids, (uint64_t)V, 0, 32, st));
    void *d_tmp = nullptr;
    HIP_TRY(sc.take(&d_tmp, tmp + 16));
    HIP_TRY(hipcub::DeviceRadixSort::SortPairs(d_tmp, tmp, keys, ids, (uint64_t)V, 0, 32, st));
    HIP_TRY(hipMemcpyAsync(c->d_vocab_order, ids.Current(), (uint64_t)V * 4, hipMemcpyDeviceToDevice, st));
    k_vocab_rank<<<grid_for(V, 256, 1024), 256, 0, st>>>(c->d_vocab_order, c->d_deg, V, c->opt.min_count,
                                                         c->d_vocab_rank, c->d_in_vocab, uw,
                                                         (unsigned long long *)totals);
    HIP_TRY(hipGetLastError());
    k_keep<<<grid_for(V, 256, 65535), 256, 0, st>>>(c->d_deg, c->d_in_vocab, V, c->opt.sample, totals + 1,
                                                   c->d_keep_thr);
    HIP_TRY(hipGetLastError());
    uint64_t h_tot[2] = {0, 0};
    HIP_TRY(hipMemcpyAsync(h_tot, totals, 16, hipMemcpyDeviceToHost, st));
    HIP_TRY(alias_segments(c, sc, seg01, 1, V, (const double *)uw, nullptr, PairOut{c->d_uni}, nullptr));
    HIP_TRY(hipStreamSynchronize(st));  // h_tot
    c->vocab_size = h_tot[0];
    c->train_words = h_tot[1];
    sc.release(m0);
    // n log2 n table for the walk statistics
    const uint32_t L = c->opt.max_walk_len;
    // (held by the context until the build's last wait: asynchronous copies)
    std::vector<double> &T = c->h_T;
    T.assign(2 * L + 3, 0.0);
    double *dT = T.data() + L + 2;
    for (uint32_t n = 0; n < L + 2; ++n) T[n] = n <= 1 ? 0.0 : (double)n * std::log2((double)n);
    for (uint32_t f = 0; f < L + 1; ++f) dT[f] = T[f + 1] - T[f];
    HIP_TRY(hipMalloc(&c->d_T, (L + 2) * 8));
    HIP_TRY(hipMalloc(&c->d_dT, (L + 1) * 8));
    HIP_TRY(hipMemcpyAsync(c->d_T, T.data(), (L + 2) * 8, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(c->d_dT, dT, (L + 1) * 8, hipMemcpyHostToDevice, st));
    return hipSuccess;
}

hipError_t init_tables(felog_ctx *c, Scratch *sc) {
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
    const size_t m0 = sc ? sc->mark() : 0;
    if (sc) HIP_TRY(sc->take(&d_jump, sizeof(jump)));
    else HIP_TRY(hipMalloc(&d_jump, sizeof(jump)));
    HIP_TRY(hipMemcpyAsync(d_jump, jump, sizeof(jump), hipMemcpyHostToDevice, c->stream));
    if (es == 4)
        k_init_syn0<float><<<grid_for(n / 8, 256, 1u << 20), 256, 0, c->stream>>>(
            c->d_vocab_rank, c->V, c->d, c->dp, d_jump, (float *)c->d_syn0);
    else
        k_init_syn0<uint16_t><<<grid_for(n / 8, 256, 1u << 20), 256, 0, c->stream>>>(
            c->d_vocab_rank, c->V, c->d, c->dp, d_jump, (uint16_t *)c->d_syn0);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemsetAsync(c->d_syn1, 0, n * es, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (sc) sc->release(m0);
    else dfree(d_jump);
    return hipSuccess;
}

}  // namespace felog
