// graph_tools.hip - the steps either side of the walk -> SGNS path that the
// reference runs as separate tools, on the device:
//   - common-neighbour counts (src/tools/gcn.cpp:304-357): the reference
//     sorts every adjacency list and intersects the two lists of each edge;
//     here the CSR is sorted once by (vertex, neighbour) with a device radix
//     sort, and one wavefront per pair intersects by binary search of the
//     shorter list's entries in the longer list (64 lookups at a time).  A
//     pair's count is the number of distinct common neighbours other than
//     its two endpoints, the reference's count on graphs without self loops
//     (its correction also drops a repeated or endpoint entry of the
//     intersection, :330-344).  The counts feed the reference's transition
//     (felog_use_common_neighbours) as its edge data does in the reference.
//   - held-out link-prediction AUC: the cosine of the syn0 rows of every
//     test pair, then the Mann-Whitney statistic with ties at half weight
//     (sort by score, average ranks per run of equal scores).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include "felog_dev.h"
#include "felog_internal.h"

namespace felog {
namespace {

inline unsigned grid_for(uint64_t n, unsigned block, unsigned cap) {
    uint64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (unsigned)g;
}

// (vertex << 32) | neighbour for every CSR entry; one wavefront per vertex.
__global__ __launch_bounds__(256) void k_cn_keys(const uint64_t *__restrict__ row_ptr, const uint32_t *__restrict__ col,
                                                 uint32_t V, uint64_t *__restrict__ keys) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
    for (uint32_t v = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; v < V; v += nwaves)
        for (uint64_t e = row_ptr[v] + lane; e < row_ptr[v + 1]; e += 64) keys[e] = ((uint64_t)v << 32) | col[e];
}

__global__ void k_cn_low(const uint64_t *__restrict__ keys, uint64_t E, uint32_t *__restrict__ scol) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E; e += stride) scol[e] = (uint32_t)keys[e];
}

// Pair i = (u[i], v[i]), or with u null the CSR entry i (its row and
// column): distinct common neighbours of the two endpoints, endpoints
// excluded.  The shorter sorted list is walked 64 entries at a time; each
// lane looks its entry up in the longer one.
__global__ __launch_bounds__(256) void k_cn_pairs(const uint64_t *__restrict__ row_ptr,
                                                  const uint32_t *__restrict__ scol, const uint32_t *__restrict__ u,
                                                  const uint32_t *__restrict__ v, const uint32_t *__restrict__ erow,
                                                  const uint32_t *__restrict__ ecol, uint64_t n,
                                                  float *__restrict__ out_f, int32_t *__restrict__ out_i) {
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    for (uint64_t i = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; i < n; i += nwaves) {
        const uint32_t a = u ? u[i] : erow[i], b = u ? v[i] : ecol[i];
        uint64_t sa = row_ptr[a], na = row_ptr[a + 1] - sa, sb = row_ptr[b], nb = row_ptr[b + 1] - sb;
        if (na > nb) {
            const uint64_t t0 = sa, t1 = na;
            sa = sb;
            na = nb;
            sb = t0;
            nb = t1;
        }
        uint32_t cnt = 0;
        for (uint64_t j0 = 0; j0 < na; j0 += 64) {
            const uint64_t j = j0 + lane;
            bool hit = false;
            if (j < na) {
                const uint32_t x = scol[sa + j];
                // first of a run of equal entries, not an endpoint
                if (x != a && x != b && (j == 0 || scol[sa + j - 1] != x)) {
                    uint64_t lo = 0, hi = nb;
                    while (lo < hi) {
                        const uint64_t mid = lo + (hi - lo) / 2;
                        if (scol[sb + mid] < x) lo = mid + 1;
                        else hi = mid;
                    }
                    hit = lo < nb && scol[sb + lo] == x;
                }
            }
            cnt += (uint32_t)__popcll(__ballot(hit));
        }
        if (lane == 0) {
            if (out_f) out_f[i] = (float)cnt;
            if (out_i) out_i[i] = (int32_t)cnt;
        }
    }
}

// Row of every CSR entry (one wavefront per vertex).
__global__ __launch_bounds__(256) void k_cn_rows(const uint64_t *__restrict__ row_ptr, uint32_t V,
                                                 uint32_t *__restrict__ erow) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
    for (uint32_t v = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; v < V; v += nwaves)
        for (uint64_t e = row_ptr[v] + lane; e < row_ptr[v + 1]; e += 64) erow[e] = v;
}

// Cosine of the syn0 rows of each pair: 16 lanes per pair, four pairs per
// wavefront.
template <typename T>
__device__ __forceinline__ float ldf(const T *p);
template <>
__device__ __forceinline__ float ldf<float>(const float *p) { return *p; }
template <>
__device__ __forceinline__ float ldf<uint16_t>(const uint16_t *p) { return bf16_to_f32(*p); }

template <typename T>
__global__ __launch_bounds__(256) void k_pair_cos(const T *__restrict__ syn0, uint32_t dp, uint32_t d,
                                                  const uint32_t *__restrict__ pairs, uint64_t n,
                                                  float *__restrict__ score) {
    const uint32_t l = threadIdx.x & 15;
    const uint64_t grp = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 4;
    const uint64_t ngrp = ((uint64_t)gridDim.x * blockDim.x) >> 4;
    // a pair's 16 lanes stay together, so the 16-lane reductions never read
    // a lane that left the loop
    for (uint64_t i = grp; i < n; i += ngrp) {
        const T *x = syn0 + (uint64_t)pairs[2 * i] * dp;
        const T *y = syn0 + (uint64_t)pairs[2 * i + 1] * dp;
        float xy = 0.f, xx = 0.f, yy = 0.f;
        for (uint32_t e = l; e < d; e += 16) {
            const float a = ldf(x + e), b = ldf(y + e);
            xy += a * b;
            xx += a * a;
            yy += b * b;
        }
        for (int o = 8; o > 0; o >>= 1) {
            xy += __shfl_xor(xy, o);
            xx += __shfl_xor(xx, o);
            yy += __shfl_xor(yy, o);
        }
        if (l == 0) {
            const float den = sqrtf(xx) * sqrtf(yy);
            score[i] = den > 0.f ? xy / den : 0.f;
        }
    }
}

// Runs of equal sorted scores: the run's first position and its size and
// positives, then the positives' rank sum with each run at its average rank.
__global__ void k_auc_runs(const float *__restrict__ s, const uint8_t *__restrict__ lab, uint64_t n,
                           unsigned long long *__restrict__ run_start, unsigned long long *__restrict__ run_n,
                           unsigned long long *__restrict__ run_pos, const uint64_t *__restrict__ run_id) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint64_t g = run_id[i] - 1;
        if (i == 0 || s[i] != s[i - 1]) run_start[g] = i;
        atomicAdd(&run_n[g], 1ull);
        if (lab[i]) atomicAdd(&run_pos[g], 1ull);
    }
}

__global__ void k_auc_heads(const float *__restrict__ s, uint64_t n, uint64_t *__restrict__ head) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        head[i] = (i == 0 || s[i] != s[i - 1]) ? 1ull : 0ull;
}

__global__ void k_auc_sum(const unsigned long long *__restrict__ run_start, const unsigned long long *__restrict__ run_n,
                          const unsigned long long *__restrict__ run_pos, const uint64_t *__restrict__ nruns,
                          double *__restrict__ out) {
    __shared__ double part[256];
    double acc = 0.0, npos = 0.0;
    const uint64_t R = *nruns;
    for (uint64_t g = threadIdx.x; g < R; g += blockDim.x) {
        // ranks start + 1 .. start + size: average start + (size + 1) / 2
        acc += (double)run_pos[g] * ((double)run_start[g] + 0.5 * ((double)run_n[g] + 1.0));
        npos += (double)run_pos[g];
    }
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) part[threadIdx.x] += part[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = part[0];
    __syncthreads();
    part[threadIdx.x] = npos;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) part[threadIdx.x] += part[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[1] = part[0];
}

__global__ void k_last_u64(const uint64_t *__restrict__ x, uint64_t n, uint64_t *__restrict__ out) { *out = x[n - 1]; }

}  // namespace

#define HIP_TRY(x)                              \
    do {                                        \
        hipError_t e_ = (x);                    \
        if (e_ != hipSuccess) return e_;        \
    } while (0)

// The CSR with every row's neighbours in ascending order (built once).
static hipError_t sorted_csr(felog_ctx *c) {
    if (c->d_scol || c->E == 0) return hipSuccess;
    const uint64_t E = c->E;
    uint64_t *ka = nullptr, *kb = nullptr;
    void *tmp = nullptr;
    size_t tb = 0;
    HIP_TRY(hipMalloc(&ka, E * 8));
    HIP_TRY(hipMalloc(&kb, E * 8));
    k_cn_keys<<<grid_for((uint64_t)c->V * 64, 256, 1u << 16), 256, 0, c->stream>>>(c->d_row_ptr, c->d_col, c->V, ka);
    HIP_TRY(hipGetLastError());
    int vbits = 1;
    while (vbits < 32 && (1ull << vbits) < (uint64_t)c->V) ++vbits;
    hipcub::DoubleBuffer<uint64_t> keys(ka, kb);
    HIP_TRY(hipcub::DeviceRadixSort::SortKeys(nullptr, tb, keys, E, 0, 32 + vbits, c->stream));
    HIP_TRY(hipMalloc(&tmp, tb + 16));
    HIP_TRY(hipcub::DeviceRadixSort::SortKeys(tmp, tb, keys, E, 0, 32 + vbits, c->stream));
    HIP_TRY(hipMalloc(&c->d_scol, E * 4));
    k_cn_low<<<grid_for(E, 256, 65535), 256, 0, c->stream>>>(keys.Current(), E, c->d_scol);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(c->stream));
    dfree(ka);
    dfree(kb);
    dfree(tmp);
    return hipSuccess;
}

hipError_t common_neighbours_pairs(felog_ctx *c, const uint32_t *d_u, const uint32_t *d_v, uint64_t n,
                                   int32_t *d_out) {
    HIP_TRY(sorted_csr(c));
    if (n == 0 || c->E == 0) {
        if (n) HIP_TRY(hipMemsetAsync(d_out, 0, n * 4, c->stream));
        return hipStreamSynchronize(c->stream);
    }
    k_cn_pairs<<<grid_for(n * 64, 256, 1u << 16), 256, 0, c->stream>>>(c->d_row_ptr, c->d_scol, d_u, d_v, nullptr,
                                                                      nullptr, n, nullptr, d_out);
    HIP_TRY(hipGetLastError());
    return hipStreamSynchronize(c->stream);
}

hipError_t common_neighbours_csr(felog_ctx *c) {
    HIP_TRY(sorted_csr(c));
    const uint64_t E = c->E;
    if (!c->d_cn) HIP_TRY(hipMalloc(&c->d_cn, (E ? E : 1) * 4));
    if (E == 0) return hipSuccess;
    uint32_t *erow = nullptr;
    HIP_TRY(hipMalloc(&erow, E * 4));
    k_cn_rows<<<grid_for((uint64_t)c->V * 64, 256, 1u << 16), 256, 0, c->stream>>>(c->d_row_ptr, c->V, erow);
    k_cn_pairs<<<grid_for(E * 64, 256, 1u << 16), 256, 0, c->stream>>>(c->d_row_ptr, c->d_scol, nullptr, nullptr,
                                                                      erow, c->d_col, E, c->d_cn, nullptr);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(c->stream));
    dfree(erow);
    return hipSuccess;
}

hipError_t link_auc(felog_ctx *c, const uint32_t *d_pairs, const uint8_t *d_lab, uint64_t n, float *d_score,
                    double *auc) {
    hipStream_t st = c->stream;
    const unsigned g16 = grid_for(n * 16, 256, 1u << 16);
    if (c->opt.precision == FELOG_BF16)
        k_pair_cos<uint16_t><<<g16, 256, 0, st>>>((const uint16_t *)c->d_syn0, c->dp, c->d, d_pairs, n, d_score);
    else
        k_pair_cos<float><<<g16, 256, 0, st>>>((const float *)c->d_syn0, c->dp, c->d, d_pairs, n, d_score);
    HIP_TRY(hipGetLastError());
    float *ks = nullptr;
    uint8_t *kl = nullptr;
    uint64_t *head = nullptr, *rid = nullptr, *nruns = nullptr;
    unsigned long long *rs = nullptr, *rn = nullptr, *rp = nullptr;
    double *d_out = nullptr;
    void *tmp = nullptr;
    size_t t1 = 0, t2 = 0;
    HIP_TRY(hipMalloc(&ks, n * 4));
    HIP_TRY(hipMalloc(&kl, n));
    HIP_TRY(hipMalloc(&head, n * 8));
    HIP_TRY(hipMalloc(&rid, n * 8));
    HIP_TRY(hipMalloc(&nruns, 8));
    HIP_TRY(hipMalloc(&rs, n * 8));
    HIP_TRY(hipMalloc(&rn, n * 8));
    HIP_TRY(hipMalloc(&rp, n * 8));
    HIP_TRY(hipMalloc(&d_out, 16));
    HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, t1, d_score, ks, d_lab, kl, n, 0, 32, st));
    HIP_TRY(hipcub::DeviceScan::InclusiveSum(nullptr, t2, head, rid, n, st));
    HIP_TRY(hipMalloc(&tmp, (t1 > t2 ? t1 : t2) + 16));
    HIP_TRY(hipcub::DeviceRadixSort::SortPairs(tmp, t1, d_score, ks, d_lab, kl, n, 0, 32, st));
    k_auc_heads<<<grid_for(n, 256, 65535), 256, 0, st>>>(ks, n, head);
    HIP_TRY(hipcub::DeviceScan::InclusiveSum(tmp, t2, head, rid, n, st));
    HIP_TRY(hipMemsetAsync(rn, 0, n * 8, st));
    HIP_TRY(hipMemsetAsync(rp, 0, n * 8, st));
    k_last_u64<<<1, 1, 0, st>>>(rid, n, nruns);
    k_auc_runs<<<grid_for(n, 256, 65535), 256, 0, st>>>(ks, kl, n, rs, rn, rp, rid);
    k_auc_sum<<<1, 256, 0, st>>>(rs, rn, rp, nruns, d_out);
    HIP_TRY(hipGetLastError());
    double h[2] = {0.0, 0.0};
    HIP_TRY(hipMemcpyAsync(h, d_out, 16, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    const double P = h[1], N = (double)n - h[1];
    *auc = (P == 0.0 || N == 0.0) ? 0.5 : (h[0] - P * (P + 1.0) / 2.0) / (P * N);
    for (void *p : {(void *)ks, (void *)kl, (void *)head, (void *)rid, (void *)nruns, (void *)rs, (void *)rn,
                    (void *)rp, (void *)d_out, tmp})
        dfree(p);
    return hipSuccess;
}

}  // namespace felog
