// sgns_kernel.hip - skip-gram negative sampling over a walk corpus, HogBatch
// style, on MFMA.
//
// Semantics restated (reference paths relative to the reference repository):
//   sentence filtering  src/word2vec.cu:1196-1232  (vocabulary lookup,
//                       subsampling keep = min(1, (sqrt(c/sT)+1) sT/c) with
//                       c = degree, 1000-token cut, optional index-0 break)
//   window shrink       src/word2vec.cu:458-465    (b = rand % window)
//   per-pair update     src/word2vec.cu:494-534    (label 1 for the centre,
//                       0 for negatives, negatives equal to the centre are
//                       skipped, logits clamped at +-6; syn0 row = context,
//                       syn1neg row = centre / negative)
//   shared negatives    include/dsgl.hpp:1054-1159, 1241-1310 (HogBatch:
//                       outputs = targets + shared negatives, corrM = out*in^T,
//                       dOut = corrM*in, dIn = corrM^T*out, meta counts)
//
// One workgroup (4 wavefronts) owns one sentence at a time.  The kept tokens
// of a sentence are cut into groups of B consecutive centres; each group
// draws k negatives that its centres share.  Consecutive groups are packed
// into one tile while the tile has at most 32 output slots (centres + k per
// group) and its windows span at most 64 positions; a 22-token walk with
// B = 21 and k = 5 is a single tile.  For a tile, the distinct input rows
// (tokens within the windows) and distinct output rows (centres and
// negatives) are staged in LDS, and with
//   P[r][o] = # (context r, centre o) pairs,
//   Q[r][o] = # (context r, negative o) pairs of those centres,
//   X = Vin * Uout^T                          (MFMA, f32 accumulate)
//   G = alpha * (P - (P+Q) * sigma(X))        (f32, sigma clamped at +-6)
//   dVin = G * Uout,  dUout = G^T * Vin       (MFMA, f32 accumulate)
// which is the sum of the per-pair word2vec updates of the tile with all
// reads taken before the tile's writes.  Rows are then updated Hogwild: f32
// tables by hardware float atomics, bf16 tables by a 32-bit compare-and-swap
// of two elements with stochastic rounding from Philox bits.  Nothing orders
// one tile's writes before another tile's reads (Hogwild).
#include <hip/hip_runtime.h>
#include "felog_dev.h"
#include "felog_internal.h"

namespace felog {

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

constexpr int kBlock = 256;
constexpr int RMAX = 64;       // distinct input rows of a tile
constexpr int OMAX = 32;       // output slots of a tile
constexpr int GLD = OMAX + 4;  // row pitch of the count / G tile (words)
constexpr int MAXKEEP = 1000;  // MAX_SENTENCE_LENGTH, src/word2vec.cu:42
constexpr float kMaxExp = 6.f; // MAX_EXP, src/word2vec.cu:40

struct SgnsArgs {
    const uint64_t *__restrict__ off;
    const uint32_t *__restrict__ tok;
    uint64_t n_sent;
    void *syn0;
    void *syn1;
    uint32_t dp, d;
    const uint8_t *__restrict__ in_vocab;
    const uint32_t *__restrict__ vocab_rank;
    const uint32_t *__restrict__ keep_thr;
    const uint32_t *__restrict__ uni_thr;
    const uint32_t *__restrict__ uni_alias;
    uint32_t V;
    uint32_t window, negative, bgroup;
    int compat_break;
    float alpha;
    uint32_t round;
    uint64_t seed;
    unsigned long long *counters;
    // test hook: when dbg_max > 0 the tables are not touched
    uint32_t dbg_max;
    uint32_t *dbg_groups, *dbg_rows;
    float *dbg_logits, *dbg_dv, *dbg_du;
};

// LDS carve-up (bytes), identical on host and device.
struct Lds {
    uint32_t ldv;  // row pitch of staged rows, in elements
    size_t o_tok, o_v, o_u, o_g, o_small, total;
};

__host__ __device__ inline Lds lds_layout(uint32_t dp, uint32_t esize) {
    Lds L;
    L.ldv = dp + (esize == 4 ? 4u : 8u);  // 16-byte skew between rows
    size_t o = 0;
    L.o_tok = o;   o += (size_t)(MAXKEEP + 24) * 4;
    L.o_v = o;     o += (size_t)RMAX * L.ldv * esize;
    L.o_u = o;     o += (size_t)OMAX * L.ldv * esize;
    L.o_g = o;     o += (size_t)RMAX * GLD * 4;
    L.o_small = o; o += 1024;
    L.total = o;
    return L;
}

struct Small {  // per-tile bookkeeping
    uint32_t rin_tok[RMAX];
    uint32_t rout_tok[OMAX];
    uint32_t neg_tok[OMAX];
    uint8_t pos_rin[RMAX];
    uint8_t cen_rout[OMAX];
    uint8_t neg_rout[OMAX];
    uint8_t bshrink[OMAX];
    uint32_t n_in, n_out, brk, words, wave_cnt[4], dbg_slot;
};
static_assert(sizeof(Small) <= 1024, "small block");

__device__ __forceinline__ float load_elem(const float *p) { return *p; }
__device__ __forceinline__ float load_elem(const uint16_t *p) { return bf16_to_f32(*p); }

__device__ __forceinline__ float sigmoid_clamped(float x) {
    if (x > kMaxExp) return 1.f;
    if (x < -kMaxExp) return 0.f;
    return 1.f / (1.f + __expf(-x));
}

// Apply one f32 delta pair to two adjacent bf16 elements with stochastic
// rounding, by compare-and-swap on the containing 32-bit word.
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

// Number of consecutive groups, starting at group g0, that form one tile.
__device__ __forceinline__ uint32_t tile_groups(uint32_t g0, uint32_t m, uint32_t B, uint32_t w, uint32_t k) {
    const uint32_t c_start = g0 * B;
    const uint32_t lo = c_start >= w ? c_start - w : 0;
    uint32_t ng = 1;
    while (true) {
        const uint32_t g = g0 + ng;
        if (g * B >= m) break;
        const uint32_t ce = (g + 1) * B < m ? (g + 1) * B : m;
        const uint32_t hi = ce + w < m ? ce + w : m;
        if ((ce - c_start) + k * (ng + 1) > (uint32_t)OMAX || hi - lo > (uint32_t)RMAX) break;
        ++ng;
    }
    return ng;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_sgns(SgnsArgs a) {
    extern __shared__ __align__(16) unsigned char lds[];
    const Lds L = lds_layout(a.dp, sizeof(T));
    uint32_t *stok = (uint32_t *)(lds + L.o_tok);
    T *Vs = (T *)(lds + L.o_v);
    T *Us = (T *)(lds + L.o_u);
    float *Gs = (float *)(lds + L.o_g);
    uint32_t *cnt = (uint32_t *)(lds + L.o_g);  // counts share the G tile
    Small &sm = *(Small *)(lds + L.o_small);

    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const uint32_t w = a.window, k = a.negative, B = a.bgroup;
    const uint32_t dp = a.dp, ldv = L.ldv;
    const uint32_t row16 = lane & 15, grp = lane >> 4;
    T *syn0 = (T *)a.syn0;
    T *syn1 = (T *)a.syn1;
    unsigned long long my_pairs = 0, my_tiles = 0, my_words = 0, my_kept = 0;

    for (uint64_t sid = blockIdx.x; sid < a.n_sent; sid += gridDim.x) {
        const uint64_t b0 = a.off[sid];
        const uint32_t n = (uint32_t)(a.off[sid + 1] - b0);

        // ---- filter the sentence into stok[0, m) ----------------------------
        uint32_t m = 0, words = 0;
        for (uint32_t chunk = 0; chunk < n; chunk += kBlock) {
            const uint32_t p = chunk + tid;
            int status = 0;  // 0 none, 1 drop, 2 keep, 3 break
            uint32_t t = 0;
            if (p < n) {
                t = a.tok[b0 + p];
                const bool iv = a.in_vocab[t] != 0;
                const uint32_t thr = a.keep_thr[t];
                if (a.compat_break && (!iv || a.vocab_rank[t] == 0)) {
                    status = 3;
                } else if (!iv) {
                    status = 1;
                } else {
                    status = (thr == 0xFFFFFFFFu ||
                              draw(a.seed, p, DOM_SUBSAMPLE, (uint32_t)sid, a.round).x < thr)
                                 ? 2 : 1;
                }
            }
            if (tid == 0) { sm.brk = 0xFFFFFFFFu; sm.words = 0xFFFFFFFFu; }
            __syncthreads();
            if (status == 3) atomicMin(&sm.brk, p);
            __syncthreads();
            const uint32_t brk = sm.brk;
            const bool keep = status == 2 && p < brk;
            const uint64_t bal = __ballot(keep);
            if (lane == 0) sm.wave_cnt[wave] = (uint32_t)__popcll(bal);
            __syncthreads();
            uint32_t before = 0, total = 0;
            for (uint32_t q = 0; q < kBlock / 64; ++q) {
                if (q < wave) before += sm.wave_cnt[q];
                total += sm.wave_cnt[q];
            }
            const uint32_t idx = m + before + (uint32_t)__popcll(bal & lt_mask);
            if (keep && idx < (uint32_t)MAXKEEP) stok[idx] = t;
            if (keep && idx == (uint32_t)MAXKEEP - 1) sm.words = p + 1;
            __syncthreads();
            // tokens read: to the break, to the token that fills the sentence,
            // or to the end of the chunk
            uint32_t read_end = n < chunk + kBlock ? n : chunk + kBlock;
            if (brk != 0xFFFFFFFFu) read_end = brk + 1;
            if (sm.words != 0xFFFFFFFFu) read_end = sm.words;
            words = read_end;
            m += total;
            __syncthreads();
            if (brk != 0xFFFFFFFFu || m >= (uint32_t)MAXKEEP) break;
        }
        if (m > (uint32_t)MAXKEEP) m = MAXKEEP;
        my_words += words;
        my_kept += m;
        if (m < 2) continue;

        // ---- tiles of consecutive groups ------------------------------------
        for (uint32_t g0 = 0; g0 * B < m;) {
            const uint32_t ng = tile_groups(g0, m, B, w, k);
            const uint32_t c0 = g0 * B;
            const uint32_t c_end = (g0 + ng) * B < m ? (g0 + ng) * B : m;
            const uint32_t nc = c_end - c0;
            const uint32_t nneg = k * ng;
            const uint32_t lo = c0 >= w ? c0 - w : 0;
            const uint32_t hi = c_end + w < m ? c_end + w : m;
            const uint32_t nin = hi - lo;  // <= 64

            if (tid < nc) sm.bshrink[tid] = (uint8_t)(draw(a.seed, c0 + tid, DOM_WINDOW, (uint32_t)sid, a.round).x % w);
            if (tid >= 64 && tid < 64 + nneg) {
                const uint32_t q = tid - 64, gq = g0 + q / k, j = q % k;
                const u32x4 r = draw(a.seed, gq, DOM_NEGATIVE | j, (uint32_t)sid, a.round);
                const uint32_t bkt = alias_bucket(r.x, a.V);
                sm.neg_tok[q] = r.y < a.uni_thr[bkt] ? bkt : a.uni_alias[bkt];
            }
            for (uint32_t i = tid; i < (uint32_t)(RMAX * GLD); i += kBlock) cnt[i] = 0;
            if (a.dbg_max && tid == 0) sm.dbg_slot = (uint32_t)atomicAdd(&a.counters[C_DEBUG], 1ull);
            __syncthreads();
            if (wave == 0) {
                // distinct input rows over positions [lo, hi)
                const bool valid = lane < nin;
                const uint32_t t = valid ? stok[lo + lane] : 0xFFFFFFFFu;
                uint32_t first = lane;
                if (valid)
                    for (uint32_t q = 0; q < lane; ++q)
                        if (stok[lo + q] == t) { first = q; break; }
                const bool isf = valid && first == lane;
                const uint64_t bal = __ballot(isf);
                const uint32_t idx = (uint32_t)__popcll(bal & lt_mask);
                if (isf) sm.rin_tok[idx] = t;
                const uint32_t my = __shfl(idx, (int)first);
                if (valid) sm.pos_rin[lane] = (uint8_t)my;
                if (lane == 0) sm.n_in = (uint32_t)__popcll(bal);
            } else if (wave == 1) {
                // distinct output rows over [centres..., negatives...]
                const uint32_t nl = nc + nneg;
                const bool valid = lane < nl;
                uint32_t t = 0xFFFFFFFFu;
                if (valid) t = lane < nc ? stok[c0 + lane] : sm.neg_tok[lane - nc];
                uint32_t first = lane;
                if (valid)
                    for (uint32_t q = 0; q < lane; ++q) {
                        const uint32_t tq = q < nc ? stok[c0 + q] : sm.neg_tok[q - nc];
                        if (tq == t) { first = q; break; }
                    }
                const bool isf = valid && first == lane;
                const uint64_t bal = __ballot(isf);
                const uint32_t idx = (uint32_t)__popcll(bal & lt_mask);
                if (isf) sm.rout_tok[idx] = t;
                const uint32_t my = __shfl(idx, (int)first);
                if (valid) {
                    if (lane < nc) sm.cen_rout[lane] = (uint8_t)my;
                    else sm.neg_rout[lane - nc] = (uint8_t)my;
                }
                if (lane == 0) sm.n_out = (uint32_t)__popcll(bal);
            }
            __syncthreads();
            const uint32_t n_in = sm.n_in, n_out = sm.n_out;
            const uint32_t R16 = (n_in + 15) & ~15u, O16 = (n_out + 15) & ~15u;

            // pair counts: P in the low half-word, Q in the high half-word
            const uint32_t span = 2 * w + 1;
            for (uint32_t q = tid; q < nc * span; q += kBlock) {
                const uint32_t i = q / span, aa = q % span;
                const uint32_t b = sm.bshrink[i];
                if (aa == w || aa < b || aa >= span - b) continue;
                const int p = (int)(c0 + i) - (int)w + (int)aa;
                if (p < 0 || p >= (int)m) continue;
                const uint32_t rin = sm.pos_rin[p - (int)lo];
                atomicAdd(&cnt[rin * GLD + sm.cen_rout[i]], 1u);
                const uint32_t ct = stok[c0 + i];
                const uint32_t nb = ((c0 + i) / B - g0) * k;  // negatives of this centre's group
                for (uint32_t j = 0; j < k; ++j)
                    if (sm.neg_tok[nb + j] != ct) atomicAdd(&cnt[rin * GLD + sm.neg_rout[nb + j]], 1u << 16);
                ++my_pairs;
            }
            // stage rows (16-byte vectors); padding rows are zero
            {
                const uint32_t vpr = dp * sizeof(T) / 16;  // vectors per row
                const uint32_t nv = (R16 + O16) * vpr;
                for (uint32_t i = tid; i < nv; i += kBlock) {
                    const uint32_t row = i / vpr, c = i % vpr;
                    uint4 val = make_uint4(0, 0, 0, 0);
                    T *dstp;
                    if (row < R16) {
                        if (row < n_in) val = *(const uint4 *)(syn0 + (uint64_t)sm.rin_tok[row] * dp + c * (16 / sizeof(T)));
                        dstp = Vs + row * ldv + c * (16 / sizeof(T));
                    } else {
                        const uint32_t o = row - R16;
                        if (o < n_out) val = *(const uint4 *)(syn1 + (uint64_t)sm.rout_tok[o] * dp + c * (16 / sizeof(T)));
                        dstp = Us + o * ldv + c * (16 / sizeof(T));
                    }
                    *(uint4 *)dstp = val;
                }
            }
            __syncthreads();

            const uint32_t slot_dbg = a.dbg_max ? sm.dbg_slot : 0;
            const bool dbg_on = a.dbg_max && slot_dbg < a.dbg_max;
            if (dbg_on) {
                if (tid == 0) {
                    uint32_t *gq = a.dbg_groups + (uint64_t)slot_dbg * 5;
                    gq[0] = (uint32_t)sid; gq[1] = c0; gq[2] = nc; gq[3] = n_in; gq[4] = n_out;
                }
                uint32_t *rr = a.dbg_rows + (uint64_t)slot_dbg * (RMAX + OMAX);
                if (tid < RMAX) rr[tid] = tid < n_in ? sm.rin_tok[tid] : 0xFFFFFFFFu;
                else if (tid < RMAX + OMAX) rr[tid] = (tid - RMAX) < n_out ? sm.rout_tok[tid - RMAX] : 0xFFFFFFFFu;
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
                    const uint32_t c = cnt[r * GLD + o];
                    const float P = (float)(c & 0xFFFFu), PQ = (float)((c & 0xFFFFu) + (c >> 16));
                    Gs[r * GLD + o] = a.alpha * (P - PQ * sigmoid_clamped(acc[jj]));
                    if (dbg_on) a.dbg_logits[((uint64_t)slot_dbg * RMAX + r) * OMAX + o] = acc[jj];
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
                        for (int jj = 0; jj < 4; ++jj) {
                            const uint32_t r = rt * 16 + 4 * grp + jj;
                            a.dbg_dv[((uint64_t)slot_dbg * RMAX + r) * a.d + col] = acc[jj];
                        }
                } else if constexpr (sizeof(T) == 4) {
                    if (col < a.d)
                        for (int jj = 0; jj < 4; ++jj) {
                            const uint32_t r = rt * 16 + 4 * grp + jj;
                            if (r < n_in) unsafeAtomicAdd((float *)syn0 + (uint64_t)sm.rin_tok[r] * dp + col, acc[jj]);
                        }
                } else {
                    // pair adjacent columns across lanes l, l^1: even lanes
                    // write rows 0-1 of their 4, odd lanes rows 2-3
                    float other[4];
                    for (int jj = 0; jj < 4; ++jj) other[jj] = __shfl_xor(acc[jj], 1);
                    const u32x4 rnd = draw(a.seed, t * 64 + lane, DOM_ROUND_BF16 | g0, (uint32_t)sid, a.round);
                    const bool odd = lane & 1;
                    const uint32_t c2 = col & ~1u;
                    for (int h = 0; h < 2; ++h) {
                        const int jj = odd ? 2 + h : h;
                        const uint32_t r = rt * 16 + 4 * grp + jj;
                        const float lo_v = odd ? other[jj] : acc[jj];
                        const float hi_v = odd ? acc[jj] : other[jj];
                        if (r < n_in && c2 < a.d)
                            bf16_pair_add_sr((uint16_t *)syn0 + (uint64_t)sm.rin_tok[r] * dp + c2, lo_v, hi_v,
                                             h == 0 ? rnd.x : rnd.y);
                    }
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
                        for (int jj = 0; jj < 4; ++jj) {
                            const uint32_t o = ot * 16 + 4 * grp + jj;
                            a.dbg_du[((uint64_t)slot_dbg * OMAX + o) * a.d + col] = acc[jj];
                        }
                } else if constexpr (sizeof(T) == 4) {
                    if (col < a.d)
                        for (int jj = 0; jj < 4; ++jj) {
                            const uint32_t o = ot * 16 + 4 * grp + jj;
                            if (o < n_out) unsafeAtomicAdd((float *)syn1 + (uint64_t)sm.rout_tok[o] * dp + col, acc[jj]);
                        }
                } else {
                    float other[4];
                    for (int jj = 0; jj < 4; ++jj) other[jj] = __shfl_xor(acc[jj], 1);
                    const u32x4 rnd = draw(a.seed, (tilesV + t) * 64 + lane, DOM_ROUND_BF16 | g0, (uint32_t)sid, a.round);
                    const bool odd = lane & 1;
                    const uint32_t c2 = col & ~1u;
                    for (int h = 0; h < 2; ++h) {
                        const int jj = odd ? 2 + h : h;
                        const uint32_t o = ot * 16 + 4 * grp + jj;
                        const float lo_v = odd ? other[jj] : acc[jj];
                        const float hi_v = odd ? acc[jj] : other[jj];
                        if (o < n_out && c2 < a.d)
                            bf16_pair_add_sr((uint16_t *)syn1 + (uint64_t)sm.rout_tok[o] * dp + c2, lo_v, hi_v,
                                             h == 0 ? rnd.x : rnd.y);
                    }
                }
            }
            ++my_tiles;
            g0 += ng;
            __syncthreads();  // the next tile reuses the LDS
        }
    }
    if (tid == 0) {
        atomicAdd(&a.counters[C_GROUPS], my_tiles);
        atomicAdd(&a.counters[C_WORDS], my_words);
        atomicAdd(&a.counters[C_KEPT], my_kept);
    }
    for (int o = 32; o > 0; o >>= 1) my_pairs += __shfl_xor(my_pairs, o);
    if (lane == 0) atomicAdd(&a.counters[C_PAIRS], my_pairs);
}

}  // namespace

uint32_t sgns_group_centres(uint32_t batch, uint32_t window, uint32_t negative) {
    uint32_t b = batch == 0 ? 2 * window + 1 : batch;
    if (b + negative > (uint32_t)OMAX) b = OMAX - negative;
    if (b + 2 * window > (uint32_t)RMAX) b = RMAX - 2 * window;
    return b;
}

#define HIP_TRY(x)                              \
    do {                                        \
        hipError_t e_ = (x);                    \
        if (e_ != hipSuccess) return e_;        \
    } while (0)

hipError_t run_sgns(felog_ctx *c, const uint64_t *d_off, const uint32_t *d_tok, uint64_t n_sent,
                    uint32_t round, float alpha, const SgnsDebug *dbg, float *ms) {
    SgnsArgs a;
    a.off = d_off;
    a.tok = d_tok;
    a.n_sent = n_sent;
    a.syn0 = c->d_syn0;
    a.syn1 = c->d_syn1;
    a.dp = c->dp;
    a.d = c->d;
    a.in_vocab = c->d_in_vocab;
    a.vocab_rank = c->d_vocab_rank;
    a.keep_thr = c->d_keep_thr;
    a.uni_thr = c->d_uni_thr;
    a.uni_alias = c->d_uni_alias;
    a.V = c->V;
    a.window = c->opt.window;
    a.negative = c->opt.negative;
    a.bgroup = sgns_group_centres(c->opt.batch_size, c->opt.window, c->opt.negative);
    a.compat_break = c->opt.compat_vocab_break;
    a.alpha = alpha;
    a.round = round;
    a.seed = c->opt.seed;
    a.counters = (unsigned long long *)c->d_counters;
    a.dbg_max = dbg ? dbg->max_groups : 0;
    a.dbg_groups = dbg ? dbg->groups : nullptr;
    a.dbg_rows = dbg ? dbg->rows : nullptr;
    a.dbg_logits = dbg ? dbg->logits : nullptr;
    a.dbg_dv = dbg ? dbg->dv : nullptr;
    a.dbg_du = dbg ? dbg->du : nullptr;

    const bool bf16 = c->opt.precision == FELOG_BF16;
    const Lds L = lds_layout(c->dp, bf16 ? 2 : 4);
    int per_cu = 0;
    if (bf16) {
        HIP_TRY(hipFuncSetAttribute((const void *)k_sgns<uint16_t>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)L.total));
        HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_sgns<uint16_t>, kBlock, L.total));
    } else {
        HIP_TRY(hipFuncSetAttribute((const void *)k_sgns<float>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)L.total));
        HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_sgns<float>, kBlock, L.total));
    }
    if (per_cu < 1) per_cu = 1;
    uint64_t blocks = (uint64_t)per_cu * c->cu_count;
    if (blocks > n_sent) blocks = n_sent;
    if (blocks < 1) blocks = 1;
    // zero the training counters only
    HIP_TRY(hipMemsetAsync(c->d_counters + C_WORDS, 0, 8 * (C_DEBUG - C_WORDS + 1), c->stream));
    HIP_TRY(hipEventRecord(c->ev0, c->stream));
    if (bf16) k_sgns<uint16_t><<<(unsigned)blocks, kBlock, L.total, c->stream>>>(a);
    else k_sgns<float><<<(unsigned)blocks, kBlock, L.total, c->stream>>>(a);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(c->ev1, c->stream));
    HIP_TRY(hipEventSynchronize(c->ev1));
    HIP_TRY(hipEventElapsedTime(ms, c->ev0, c->ev1));
    HIP_TRY(hipMemcpy(c->h_counters, c->d_counters, kNumCounters * 8, hipMemcpyDeviceToHost));
    return hipSuccess;
}

}  // namespace felog
