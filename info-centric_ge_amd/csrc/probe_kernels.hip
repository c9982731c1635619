// probe_kernels.hip - a measured ceiling for the walk's memory access.
//
// A walk step reads one 32-byte edge record at an address the previous step
// decided.  The peak HBM bandwidth says little about that access; what
// bounds it is how many random lines per second the memory system returns.
// felog_probe_random_reads measures that rate on the device itself: every
// thread reads rows of read_bytes at independent pseudo-random row indices
// of a table of table_bytes (no dependency between reads, four in flight per
// thread), so the result is the throughput the walk's dependent chains can
// approach but not pass.  It has no counterpart in the reference.
#include <hip/hip_runtime.h>
#include "felog_internal.h"

namespace felog {
namespace {

// splitmix64 finaliser (Steele, Lea, Flood, "Fast splittable pseudorandom
// number generators", 2014)
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// row index in [0, rows): the high half of a 64 x 32-bit product
__device__ __forceinline__ uint64_t pick(uint64_t i, uint64_t seed, uint32_t rows) {
    return (mix64(i ^ seed) >> 32) * rows >> 32;
}

template <int V16>
__global__ __launch_bounds__(256) void k_probe_reads(const uint4 *__restrict__ table, uint32_t rows, uint64_t n,
                                                     uint64_t seed, uint32_t *__restrict__ out) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    uint32_t acc = 0;
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    // four independent rows per iteration
    for (; i + 3 * stride < n; i += 4 * stride) {
        uint4 v[4][V16];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const uint4 *row = table + pick(i + u * stride, seed, rows) * V16;
#pragma unroll
            for (int q = 0; q < V16; ++q) v[u][q] = row[q];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int q = 0; q < V16; ++q) acc ^= v[u][q].x ^ v[u][q].w;
    }
    for (; i < n; i += stride) {
        const uint4 *row = table + pick(i, seed, rows) * V16;
#pragma unroll
        for (int q = 0; q < V16; ++q) acc ^= row[q].x ^ row[q].w;
    }
    // the table is zero: this store never happens, but the loads must
    if (acc == 0x9E3779B9u) out[0] = acc;
}

}  // namespace
}  // namespace felog

extern "C" felog_status felog_probe_random_reads(int device, uint64_t table_bytes, uint64_t reads,
                                                 uint32_t read_bytes, double *reads_per_s) {
    using namespace felog;
    if (!reads_per_s || reads == 0) return FELOG_ERR_ARG;
    if (read_bytes != 16 && read_bytes != 32 && read_bytes != 64 && read_bytes != 128) return FELOG_ERR_ARG;
    const uint64_t rows = table_bytes / read_bytes;
    if (rows == 0 || rows > 0xFFFFFFFFull) return FELOG_ERR_ARG;
    if (hipSetDevice(device) != hipSuccess) return FELOG_ERR_HIP;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus < 1)
        return FELOG_ERR_HIP;
    void *table = nullptr;
    uint32_t *out = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    felog_status st = FELOG_OK;
    float ms = 0.f;
    const unsigned grid = (unsigned)cus * 16;
    auto launch = [&](uint64_t seed) {
        const uint32_t r = (uint32_t)rows;
        switch (read_bytes) {
            case 16: k_probe_reads<1><<<grid, 256>>>((const uint4 *)table, r, reads, seed, out); break;
            case 32: k_probe_reads<2><<<grid, 256>>>((const uint4 *)table, r, reads, seed, out); break;
            case 64: k_probe_reads<4><<<grid, 256>>>((const uint4 *)table, r, reads, seed, out); break;
            default: k_probe_reads<8><<<grid, 256>>>((const uint4 *)table, r, reads, seed, out); break;
        }
    };
    constexpr int kRuns = 5;
    if (hipMalloc(&table, rows * read_bytes) != hipSuccess || hipMalloc(&out, 4) != hipSuccess) {
        st = FELOG_ERR_OOM;
    } else if (hipMemset(table, 0, rows * read_bytes) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
               hipEventCreate(&e1) != hipSuccess) {
        st = FELOG_ERR_HIP;
    } else {
        launch(1);  // warm-up
        (void)hipEventRecord(e0, nullptr);
        for (int k = 0; k < kRuns; ++k) launch(2 + k);
        (void)hipEventRecord(e1, nullptr);
        if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess || ms <= 0.f)
            st = FELOG_ERR_HIP;
        else
            *reads_per_s = (double)reads * kRuns / ((double)ms * 1e-3);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    dfree(table);
    dfree(out);
    return st;
}
