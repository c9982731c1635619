// felog_dev.h - device-side building blocks shared by the HIP kernels of
// libfelog_hip.so: Philox4x32-10, the counter layout of every random draw,
// alias sampling, and the FP64 entropy/correlation update of a walker.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace felog {

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC'11).  Counter-based: every draw of the hot path is a
// pure function of (seed, counter), so the host can replay any walk.
// ---------------------------------------------------------------------------
struct u32x4 { uint32_t x, y, z, w; };

__device__ __forceinline__ u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1) {
    constexpr uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    constexpr uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        // one 32x32 -> 64-bit multiply per word gives both halves
        const uint64_t p0 = (uint64_t)M0 * c.x, p1 = (uint64_t)M1 * c.z;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c = u32x4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
        k0 += W0;
        k1 += W1;
    }
    return c;
}

// Counter layout {x: index, y: domain | sub-index, z: stream, w: round}.
// Domains keep the walk, SGNS (subsampling, window, negatives), rounding,
// evaluation and budget-network streams disjoint under one key.
enum : uint32_t {
    DOM_WALK = 0x00000000u,       // y = resample attempt, x = step, z = walker
    // x = index, z = sentence; of the four output words, the first is the
    // subsample coin of raw position x, the second the window shrink of kept
    // centre x, the last two negative j of group g for x = g * negative + j
    // (bucket, coin)
    DOM_SGNS = 0x10000000u,
    DOM_EVAL = 0x40000000u,       // y |= draw index, x = vertex, z = ~0
    DOM_ROUND_BF16 = 0x50000000u, // x = 0, z = tile slot: one draw per tile, hashed per thread
    DOM_MLP = 0x60000000u,        // x = sample slot of a budget-MLP step, w = step
};

__device__ __forceinline__ u32x4 draw(uint64_t seed, uint32_t x, uint32_t y,
                                      uint32_t z, uint32_t w) {
    return philox4x32_10(u32x4{x, y, z, w}, (uint32_t)seed, (uint32_t)(seed >> 32));
}

// Alias draw over n buckets: bucket from the high word of r0*n, coin r1
// against the bucket's u32 threshold.  Returns the chosen bucket index.
__device__ __forceinline__ uint32_t alias_bucket(uint32_t r0, uint32_t n) {
    return __umulhi(r0, n);
}

// Maximum resamples of a self-loop at step 0 before the walk is ended
// (the reference resamples without bound, include/walk.hpp:981-983).
constexpr uint32_t kMaxStartAttempts = 64;
// Entries of a walker's visit history kept in LDS; longer walks search
// the rest of their path in global memory.
constexpr int kHistLds = 32;
// Stop threshold on R^2 (include/walk.hpp:36).
constexpr double kRThreshold = 0.999;

// Entropy / correlation state of one walker (include/walk.hpp:272-366),
// restated with exact arithmetic:  A = sum_c T[c] over the visit counts c,
// H = (T[S] - A) / S is the Shannon entropy (bits) of the visit histogram
// after S visits, and the running means follow Correlation::iterateEX.
struct WalkStats {
    double A, EH, ES, EHS, EHH, ESS;
};

__device__ __forceinline__ void walk_stats_init(WalkStats &st) {
    // include/walk.hpp:937-939: (H, S, EH, ES, EHS, EHH, ESS) = (0,1,0,1,0,0,1)
    st.A = 0.0; st.EH = 0.0; st.ES = 1.0; st.EHS = 0.0; st.EHH = 0.0; st.ESS = 1.0;
}

// Enter a vertex previously visited f times; s_new = vertices on the path
// after the move.  Returns true when the stop rule fires at this step.
// Every operation is a correctly rounded IEEE op with contraction disabled,
// so the host replay reproduces each bit.
__device__ __forceinline__ bool walk_stats_step(WalkStats &st, const double *__restrict__ T,
                                                const double *__restrict__ dT, uint32_t f,
                                                uint32_t s_new, uint32_t step, uint32_t min_L) {
#pragma clang fp contract(off)
    st.A = st.A + dT[f];
    const double S = (double)s_new;
    const double H = (T[s_new] - st.A) / S;
    const double N = S;
    st.EH = st.EH + (H - st.EH) / N;
    st.ES = st.ES + (S - st.ES) / N;
    st.EHS = st.EHS + (H * S - st.EHS) / N;
    st.EHH = st.EHH + (H * H - st.EHH) / N;
    st.ESS = st.ESS + (S * S - st.ESS) / N;
    const double DH = st.EHH - st.EH * st.EH;
    const double DS = st.ESS - st.ES * st.ES;
    const double e = st.EHS - st.EH * st.ES;
    const double den = DH * DS;
    double R = 0.0;
    if (e != 0.0 && den > 0.0) R = e / __builtin_sqrt(den);
    return step > min_L && (R * R < kRThreshold || R < 0.0);
}

// ---------------------------------------------------------------------------
// The reference's transition (include/walk.hpp:976-1019): a uniform
// neighbour x of v, accepted with probability tanh(p) where
//   p = (1 / (deg v - cn)) * max(deg v, deg x) / min(deg v, deg x)
// and cn = (int) edge data (the pair's common neighbours); otherwise the
// walker steps to x and back to v (a "trace").  A library tanh differs
// between the host libm and the device, and one ulp can flip a decision, so
// both sides evaluate the restatement below: only + - * / and exact
// scaling, contraction off.
//   e^y - 1 = y (1 + y/2 (1 + y/3 (1 + ...)))   for |y| <= 0.5 (to y^18/18!)
//   e^y     = 2^k (1 + (e^r - 1)),  r = y - k ln2 (fdlibm's two-part ln2)
//   tanh x  = t / (t + 2),  t = e^{2x} - 1;  |x| >= 20 gives +-1.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double expm1_series(double y) {
#pragma clang fp contract(off)
    double p = 1.0;
    for (int i = 18; i >= 2; --i) p = 1.0 + (y / (double)i) * p;
    return y * p;
}

__device__ __forceinline__ double tanh_restated(double x) {
#pragma clang fp contract(off)
    const bool neg = x < 0.0;
    const double a = neg ? -x : x;
    if (!(a < 20.0)) return neg ? -1.0 : 1.0;
    const double y = a + a;
    double t;
    if (y <= 0.5) {
        t = expm1_series(y);
    } else {
        const double k = floor(y * 1.4426950408889634 + 0.5);
        const double r = (y - k * 6.93147180369123816490e-01) - k * 1.90821492927058770002e-10;
        t = ldexp(1.0 + expm1_series(r), (int)k) - 1.0;
    }
    const double th = t / (t + 2.0);
    return neg ? -th : th;
}

// Uniform double in [0, 1) from two draws (53 bits).
__device__ __forceinline__ double uniform53(uint32_t hi, uint32_t lo) {
    return (double)((((uint64_t)hi) << 21) | (lo >> 11)) * 0x1.0p-53;
}

// 1 when the reference's rule traces (tanh(p) <= u) for the step from v
// (degree dv) to x (degree dx) over an edge with data `data`.
__device__ __forceinline__ bool cn_trace(uint32_t dv, uint32_t dx, float data, double u) {
#pragma clang fp contract(off)
    const int64_t cn = (int64_t)(int)data;
    const double hi = (double)(dv > dx ? dv : dx), lo = (double)(dv < dx ? dv : dx);
    const double p = ((1.0 / (double)((int64_t)dv - cn)) * hi) / lo;
    return tanh_restated(p) <= u;
}

// bf16 helpers (round-to-nearest-even is never used for table updates; they
// are stochastically rounded).  gfx950 converts in hardware:
// v_cvt_pk_bf16_f32 rounds two values to nearest even, v_cvt_sr_bf16_f32
// adds the high 16 bits of its random word to the value's bit pattern and
// keeps the upper half.  On finite values both equal the integer forms
// (b + 0x7FFF + ((b >> 16) & 1)) >> 16 and (b + r16) >> 16; NaN comes back
// quiet.
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float bf16_to_f32(uint16_t h) {
    return __uint_as_float(((uint32_t)h) << 16);
}
// (a, b) -> the low and high halves of one word
__device__ __forceinline__ uint32_t f32x2_to_bf16x2_rne(float a, float b) {
    bf16x2_t v;
    v[0] = (__bf16)a;
    v[1] = (__bf16)b;
    return __builtin_bit_cast(uint32_t, v);
}
__device__ __forceinline__ uint16_t f32_to_bf16_rne(float x) { return __builtin_bit_cast(uint16_t, (__bf16)x); }
// a rounded with the high 16 bits of r, b with the low 16 bits
__device__ __forceinline__ uint32_t f32x2_to_bf16x2_sr(float a, float b, uint32_t r) {
    bf16x2_t v = __builtin_bit_cast(bf16x2_t, 0u);
    v = __builtin_amdgcn_cvt_sr_bf16_f32(v, a, r, false);
    v = __builtin_amdgcn_cvt_sr_bf16_f32(v, b, r << 16, true);
    return __builtin_bit_cast(uint32_t, v);
}
// x rounded with the low 16 bits of r16
__device__ __forceinline__ uint16_t f32_to_bf16_sr(float x, uint32_t r16) {
    const bf16x2_t v = __builtin_amdgcn_cvt_sr_bf16_f32(__builtin_bit_cast(bf16x2_t, 0u), x, r16 << 16, false);
    return (uint16_t)__builtin_bit_cast(uint32_t, v);
}

// 32-bit integer hash "lowbias32" (C. Wellons, "Hash function prospector",
// 2018): seeds independent per-thread streams from one draw.
__device__ __forceinline__ uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7FEB352Du;
    x ^= x >> 15;
    x *= 0x846CA68Bu;
    x ^= x >> 16;
    return x;
}

// Chunk plans over long rows: chunk t belongs to the row k with
// cstart[k] <= t < cstart[k + 1] (nh rows, cstart ascending).
__device__ __forceinline__ uint32_t hub_of_chunk(const uint64_t *__restrict__ cstart, uint32_t nh, uint64_t t) {
    uint32_t lo = 0, hi = nh;  // last k with cstart[k] <= t
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (cstart[mid] <= t) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Combine per-wavefront totals of a workgroup of W wavefronts (x is
// wave-uniform; part holds W slots of shared memory) into thread 0, which
// then issues one global atomic per workgroup: atomics on one address
// serialise at the cache, so one per wavefront of a large grid queues tens
// of thousands of them behind the kernel's last wave.
template <int W, typename T, typename Op>
__device__ __forceinline__ T block_combine(T x, T *part, Op op) {
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = x;
    __syncthreads();
    if (threadIdx.x == 0)
        for (int i = 1; i < W; ++i) x = op(x, part[i]);
    return x;
}

}  // namespace felog
