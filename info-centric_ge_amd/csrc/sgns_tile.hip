// sgns_tile.hip - second half of an SGNS training pass: one workgroup
// (4 wavefronts) per tile, on MFMA.
//
// For a tile with distinct input rows r (contexts, syn0) and output rows o
// (centres and shared negatives, syn1neg), with
//   P[r][o] = # (context r, centre o) pairs,
//   Q[r][o] = # (context r, negative o) pairs of those centres
// (rebuilt in LDS from the descriptor), the kernel computes
//   X = Vin * Uout^T                   (MFMA, f32 accumulate)
//   G = alpha (P - (P+Q) sigma(X))     (sigma clamped at +-6)
//   dVin = G * Uout,  dUout = G^T * Vin   (MFMA, f32 accumulate)
// which is the sum of the tile's per-pair word2vec updates
// (src/word2vec.cu:494-534: label 1 for the centre, 0 for negatives,
// negatives equal to the centre skipped, syn0 row = context) with every read
// taken before the tile's writes, i.e. HogBatch's corrM / dOut / dIn
// (include/dsgl.hpp:1241-1310) over deduplicated rows.
//
// Tiles are assigned round-robin to workgroups; each workgroup keeps the
// next tile's descriptor in flight in registers while it computes.  Latency
// of the row gathers is hidden by running several workgroups per CU (the
// LDS footprint is sized for three at d = 128 in f32).
#include <hip/hip_runtime.h>
#include "felog_dev.h"
#include "felog_internal.h"
#include "sgns_common.h"

namespace felog {
namespace sgns {

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

struct Lds {
    uint32_t ldv;  // row pitch of staged rows, in elements
    size_t o_desc, o_v, o_u, o_g, total;
};

__host__ __device__ inline Lds lds_layout(uint32_t dp, uint32_t esize, uint32_t rmax) {
    Lds L;
    L.ldv = dp + (esize == 4 ? 4u : 8u);  // 16-byte skew between rows
    size_t o = 0;
    L.o_desc = o; o += kDescBytes;
    L.o_v = o;    o += (size_t)rmax * L.ldv * esize;
    L.o_u = o;    o += (size_t)OMAX * L.ldv * esize;
    L.o_g = o;    o += (size_t)rmax * GLD * 4;
    L.total = o;
    return L;
}

__device__ __forceinline__ float load_elem(const float *p) { return *p; }
__device__ __forceinline__ float load_elem(const uint16_t *p) { return bf16_to_f32(*p); }

__device__ __forceinline__ float sigmoid_clamped(float x) {
    if (x > kMaxExp) return 1.f;
    if (x < -kMaxExp) return 0.f;
    return 1.f / (1.f + __expf(-x));
}

// bf16 pair += (d0, d1) with stochastic rounding, by compare-and-swap on the
// containing 32-bit word (no delta lost).
__device__ __forceinline__ void bf16_pair_add_sr(uint16_t *p, float d0, float d1, uint32_t rnd) {
    uint32_t *w = reinterpret_cast<uint32_t *>(p);
    uint32_t old = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    while (true) {
        const float lo = bf16_to_f32((uint16_t)(old & 0xFFFFu)) + d0;
        const float hi = bf16_to_f32((uint16_t)(old >> 16)) + d1;
        const uint32_t nw = (uint32_t)f32_to_bf16_sr(lo, rnd) | ((uint32_t)f32_to_bf16_sr(hi, rnd >> 16) << 16);
        const uint32_t prev = atomicCAS(w, old, nw);
        if (prev == old) break;
        old = prev;
    }
}

// bf16 pair <- staged pair + (d0, d1), stochastically rounded, one store.
__device__ __forceinline__ void bf16_pair_store_sr(uint16_t *p, const uint16_t *staged, float d0, float d1,
                                                   uint32_t rnd) {
    const float lo = bf16_to_f32(staged[0]) + d0;
    const float hi = bf16_to_f32(staged[1]) + d1;
    *reinterpret_cast<uint32_t *>(p) =
        (uint32_t)f32_to_bf16_sr(lo, rnd) | ((uint32_t)f32_to_bf16_sr(hi, rnd >> 16) << 16);
}

// Write the 16x16 delta tile held in acc (MFMA output layout: lane l holds
// rows 4*(l/16)+0..3 of column l%16) to rows tok[row0 + ...] of the table.
template <typename T, bool WB>
__device__ __forceinline__ void apply_tile(const Params &a, T *table, const T *staged, uint32_t ldv,
                                           const uint32_t *tok, uint32_t nrows, uint32_t row0, uint32_t col,
                                           uint32_t lane, const f4 &acc, uint32_t rnd_x, uint32_t g0,
                                           uint32_t sid) {
    const uint32_t dp = a.dp;
    const uint32_t grp = lane >> 4;
    if constexpr (sizeof(T) == 4) {
        if (col >= a.d) return;
        for (int jj = 0; jj < 4; ++jj) {
            const uint32_t r = row0 + 4 * grp + jj;
            if (r >= nrows) continue;
            float *dst = (float *)table + (uint64_t)tok[r] * dp + col;
            if constexpr (WB) *dst = ((const float *)staged)[r * ldv + col] + acc[jj];
            else unsafeAtomicAdd(dst, acc[jj]);
        }
    } else {
        // pair adjacent columns across lanes l, l^1: even lanes write rows
        // 0-1 of their 4, odd lanes rows 2-3
        float other[4];
        for (int jj = 0; jj < 4; ++jj) other[jj] = __shfl_xor(acc[jj], 1);
        const u32x4 rnd = draw(a.seed, rnd_x, DOM_ROUND_BF16 | g0, sid, a.round);
        const bool odd = lane & 1;
        const uint32_t c2 = col & ~1u;
        for (int h = 0; h < 2; ++h) {
            const int jj = odd ? 2 + h : h;
            const uint32_t r = row0 + 4 * grp + jj;
            const float lo_v = odd ? other[jj] : acc[jj];
            const float hi_v = odd ? acc[jj] : other[jj];
            if (r >= nrows || c2 >= a.d) continue;
            uint16_t *dst = (uint16_t *)table + (uint64_t)tok[r] * dp + c2;
            const uint32_t bits = h == 0 ? rnd.x : rnd.y;
            if constexpr (WB) bf16_pair_store_sr(dst, (const uint16_t *)staged + r * ldv + c2, lo_v, hi_v, bits);
            else bf16_pair_add_sr(dst, lo_v, hi_v, bits);
        }
    }
}

template <typename T, int RM, bool WB>
__global__ __launch_bounds__(kBlock) void k_sgns_tile(Params a) {
    extern __shared__ __align__(16) unsigned char lds[];
    const Lds L = lds_layout(a.dp, sizeof(T), RM);
    TileDesc &td = *(TileDesc *)(lds + L.o_desc);
    T *Vs = (T *)(lds + L.o_v);
    T *Us = (T *)(lds + L.o_u);
    float *Gs = (float *)(lds + L.o_g);
    uint32_t *cnt = (uint32_t *)(lds + L.o_g);  // counts share the G tile

    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t w = a.window, k = a.negative, B = a.bgroup;
    const uint32_t dp = a.dp, ldv = L.ldv;
    const uint32_t row16 = lane & 15, grp = lane >> 4;
    T *syn0 = (T *)a.syn0;
    T *syn1 = (T *)a.syn1;
    unsigned long long my_pairs = 0, my_tiles = 0;

    uint64_t n_tiles = a.n_sent + a.counters[C_TILES];
    if (n_tiles > a.tiles_cap) n_tiles = a.tiles_cap;
    const uint4 *descs = (const uint4 *)a.tiles;
    uint64_t tile = blockIdx.x;
    uint4 pre = make_uint4(0, 0, 0, 0);
    if (tid < (uint32_t)kDescVec && tile < n_tiles) pre = descs[tile * kDescVec + tid];

    for (; tile < n_tiles; tile += gridDim.x) {
        if (tid < (uint32_t)kDescVec) ((uint4 *)&td)[tid] = pre;
        for (uint32_t i = tid; i < (uint32_t)(RM * GLD); i += kBlock) cnt[i] = 0;
        __syncthreads();
        // the next descriptor is consumed one tile later
        const uint64_t nxt = tile + gridDim.x;
        if (tid < (uint32_t)kDescVec && nxt < n_tiles) pre = descs[nxt * kDescVec + tid];

        const uint32_t c0 = td.c0, nc = td.nc, g0 = td.g0, lo = td.lo, npos = td.npos;
        const uint32_t n_in = td.n_in, n_out = td.n_out;
        if (nc == 0) {  // sentence without a tile
            __syncthreads();
            continue;
        }
        const uint32_t R16 = (n_in + 15) & ~15u, O16 = (n_out + 15) & ~15u;

        // pair counts: P in the low half-word, Q in the high half-word
        const uint32_t span = 2 * w + 1;
        for (uint32_t q = tid; q < nc * span; q += kBlock) {
            const uint32_t i = q / span, aa = q % span;
            const uint32_t b = td.bshrink[i];
            if (aa == w || aa < b || aa >= span - b) continue;
            const int p = (int)(c0 + i) - (int)w + (int)aa - (int)lo;  // input position in the tile
            if (p < 0 || p >= (int)npos) continue;
            const uint32_t rin = td.pos_rin[p];
            const uint32_t ro = td.cen_rout[i];
            atomicAdd(&cnt[rin * GLD + ro], 1u);
            const uint32_t nb = ((c0 + i) / B - g0) * k;  // negatives of this centre's group
            for (uint32_t j = 0; j < k; ++j) {
                const uint32_t no = td.neg_rout[nb + j];
                if (no != ro) atomicAdd(&cnt[rin * GLD + no], 1u << 16);
            }
            ++my_pairs;
        }
        // stage rows (16-byte vectors); padding rows are zero
        {
            constexpr uint32_t epv = 16 / sizeof(T);
            const uint32_t vpr = dp / epv;
            const uint32_t nv = (R16 + O16) * vpr;
            for (uint32_t i = tid; i < nv; i += kBlock) {
                const uint32_t row = i / vpr, c = i % vpr;
                uint4 val = make_uint4(0, 0, 0, 0);
                T *dst;
                if (row < R16) {
                    if (row < n_in) val = *(const uint4 *)(syn0 + (uint64_t)td.rin_tok[row] * dp + c * epv);
                    dst = Vs + row * ldv + c * epv;
                } else {
                    const uint32_t o = row - R16;
                    if (o < n_out) val = *(const uint4 *)(syn1 + (uint64_t)td.rout_tok[o] * dp + c * epv);
                    dst = Us + o * ldv + c * epv;
                }
                *(uint4 *)dst = val;
            }
        }
        __syncthreads();

        const bool dbg_on = a.dbg_max && tile < a.dbg_max;
        if (dbg_on) {
            if (tid == 0) {
                uint32_t *gq = a.dbg_groups + tile * 5;
                gq[0] = td.sid; gq[1] = c0; gq[2] = nc; gq[3] = n_in; gq[4] = n_out;
            }
            uint32_t *rrow = a.dbg_rows + tile * (RMAX + OMAX);
            if (tid < (uint32_t)RMAX) rrow[tid] = tid < n_in ? td.rin_tok[tid] : 0xFFFFFFFFu;
            else if (tid < (uint32_t)(RMAX + OMAX))
                rrow[tid] = (tid - RMAX) < n_out ? td.rout_tok[tid - RMAX] : 0xFFFFFFFFu;
        }

        // ---- X = Vin * Uout^T, G = alpha (P - (P+Q) sigma(X)) in place ----
        const uint32_t tilesX = (R16 / 16) * (O16 / 16);
        for (uint32_t t = wave; t < tilesX; t += 4) {
            const uint32_t rt = t / (O16 / 16), ot = t % (O16 / 16);
            f4 acc = {0.f, 0.f, 0.f, 0.f};
            if constexpr (sizeof(T) == 4) {
                const float *va = (const float *)Vs + (rt * 16 + row16) * ldv + 4 * grp;
                const float *ub = (const float *)Us + (ot * 16 + row16) * ldv + 4 * grp;
                for (uint32_t kb = 0; kb < dp; kb += 16) {
                    const f4 x = *(const f4 *)(va + kb);
                    const f4 y = *(const f4 *)(ub + kb);
                    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(x[0], y[0], acc, 0, 0, 0);
                    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(x[1], y[1], acc, 0, 0, 0);
                    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(x[2], y[2], acc, 0, 0, 0);
                    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(x[3], y[3], acc, 0, 0, 0);
                }
            } else {
                const uint16_t *va = (const uint16_t *)Vs + (rt * 16 + row16) * ldv + 8 * grp;
                const uint16_t *ub = (const uint16_t *)Us + (ot * 16 + row16) * ldv + 8 * grp;
                for (uint32_t kb = 0; kb < dp; kb += 32) {
                    const bf8 x = *(const bf8 *)(va + kb);
                    const bf8 y = *(const bf8 *)(ub + kb);
                    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(x, y, acc, 0, 0, 0);
                }
            }
            const uint32_t o = ot * 16 + row16;
            for (int jj = 0; jj < 4; ++jj) {
                const uint32_t r = rt * 16 + 4 * grp + jj;
                const uint32_t cc = cnt[r * GLD + o];
                const float P = (float)(cc & 0xFFFFu), PQ = (float)((cc & 0xFFFFu) + (cc >> 16));
                Gs[r * GLD + o] = a.alpha * (P - PQ * sigmoid_clamped(acc[jj]));
                if (dbg_on) a.dbg_logits[(tile * RMAX + r) * OMAX + o] = acc[jj];
            }
        }
        __syncthreads();

        // ---- dVin = G * Uout  (R16 x dp, K = O16) ---------------------------
        const uint32_t dtiles = dp / 16;
        const uint32_t tilesV = (R16 / 16) * dtiles;
        for (uint32_t t = wave; t < tilesV; t += 4) {
            const uint32_t rt = t / dtiles, dt = t % dtiles;
            f4 acc = {0.f, 0.f, 0.f, 0.f};
            for (uint32_t kb = 0; kb < O16; kb += 16) {
                const f4 x = *(const f4 *)(Gs + (rt * 16 + row16) * GLD + kb + 4 * grp);
                const T *ub = Us + (kb + 4 * grp) * ldv + dt * 16 + row16;
#pragma unroll
                for (int i4 = 0; i4 < 4; ++i4)
                    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(x[i4], load_elem(ub + i4 * ldv), acc, 0, 0, 0);
            }
            const uint32_t col = dt * 16 + row16;
            if (a.dbg_max) {
                if (dbg_on && col < a.d)
                    for (int jj = 0; jj < 4; ++jj)
                        a.dbg_dv[(tile * RMAX + rt * 16 + 4 * grp + jj) * a.d + col] = acc[jj];
            } else {
                apply_tile<T, WB>(a, syn0, Vs, ldv, td.rin_tok, n_in, rt * 16, col, lane, acc, t * 64 + lane, g0,
                                  td.sid);
            }
        }
        // ---- dUout = G^T * Vin  (O16 x dp, K = R16) ------------------------
        const uint32_t tilesU = (O16 / 16) * dtiles;
        for (uint32_t t = wave; t < tilesU; t += 4) {
            const uint32_t ot = t / dtiles, dt = t % dtiles;
            f4 acc = {0.f, 0.f, 0.f, 0.f};
            for (uint32_t kb = 0; kb < R16; kb += 16) {
                const float *gq = Gs + (kb + 4 * grp) * GLD + ot * 16 + row16;
                const T *vb = Vs + (kb + 4 * grp) * ldv + dt * 16 + row16;
#pragma unroll
                for (int i4 = 0; i4 < 4; ++i4)
                    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(gq[i4 * GLD], load_elem(vb + i4 * ldv), acc, 0, 0, 0);
            }
            const uint32_t col = dt * 16 + row16;
            if (a.dbg_max) {
                if (dbg_on && col < a.d)
                    for (int jj = 0; jj < 4; ++jj)
                        a.dbg_du[(tile * OMAX + ot * 16 + 4 * grp + jj) * a.d + col] = acc[jj];
            } else {
                apply_tile<T, WB>(a, syn1, Us, ldv, td.rout_tok, n_out, ot * 16, col, lane, acc,
                                  (tilesV + t) * 64 + lane, g0, td.sid);
            }
        }
        ++my_tiles;
        __syncthreads();  // the next tile reuses the LDS
    }
    if (tid == 0) atomicAdd(&a.counters[C_GROUPS], my_tiles);
    for (int o = 32; o > 0; o >>= 1) my_pairs += __shfl_xor(my_pairs, o);
    if (lane == 0) atomicAdd(&a.counters[C_PAIRS], my_pairs);
}

template <typename T, int RM, bool WB>
hipError_t launch(felog_ctx *c, const Params &a) {
    const Lds L = lds_layout(c->dp, sizeof(T), RM);
    hipError_t e = hipFuncSetAttribute((const void *)k_sgns_tile<T, RM, WB>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)L.total);
    if (e != hipSuccess) return e;
    int per_cu = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_sgns_tile<T, RM, WB>, kBlock, L.total);
    if (e != hipSuccess) return e;
    if (per_cu < 1) per_cu = 1;
    const unsigned blocks = (unsigned)per_cu * c->cu_count;
    k_sgns_tile<T, RM, WB><<<blocks, kBlock, L.total, c->stream>>>(a);
    return hipGetLastError();
}

template <typename T, int RM>
hipError_t launch_rm(felog_ctx *c, const Params &a) {
    return a.writeback ? launch<T, RM, true>(c, a) : launch<T, RM, false>(c, a);
}

}  // namespace

hipError_t launch_tile(felog_ctx *c, const Params &a) {
    if (c->opt.precision == FELOG_BF16)
        return a.rmax == 48 ? launch_rm<uint16_t, 48>(c, a) : launch_rm<uint16_t, 64>(c, a);
    return a.rmax == 48 ? launch_rm<float, 48>(c, a) : launch_rm<float, 64>(c, a);
}

}  // namespace sgns
}  // namespace felog
