// sgns_prep.hip - first half of an SGNS training pass, and the pass driver.
//
// k_sgns_prep runs one wavefront per sentence and does the integer work of
// the skip-gram pass (reference paths relative to the reference repository):
//   sentence filtering  src/word2vec.cu:1196-1232  (vocabulary lookup,
//                       subsampling keep = min(1, (sqrt(c/sT)+1) sT/c) with
//                       c = degree and the draw Philox(seed; position,
//                       SUBSAMPLE, sentence, round), 1000-token cut, optional
//                       break at vocabulary index 0)
//   window shrink       src/word2vec.cu:458-465    (b = rand % window, from
//                       Philox(seed; centre, WINDOW, sentence, round))
//   negatives           src/word2vec.cu:496-507    (unigram^0.75 law, here an
//                       alias draw from Philox(seed; group, NEGATIVE|j, ...))
// and writes, per tile, the distinct input and output rows and the index
// maps the tile kernel needs to rebuild its pair counts.  The dependent
// gathers (tokens -> vocabulary tables -> sampler) run at full occupancy,
// away from the MFMA kernel.
#include <hip/hip_runtime.h>
#include "felog_dev.h"
#include "felog_internal.h"
#include "sgns_common.h"

namespace felog {
namespace sgns {

namespace {

constexpr int kPrepWaves = kBlock / 64;

// Lowest lane whose valid value equals this lane's (the lane itself for
// an invalid lane): one uniform pass per distinct value, each a broadcast
// of the lowest unmatched lane's value and a ballot of the lanes equal to it.
__device__ __forceinline__ uint32_t first_lane_of(uint32_t t, bool valid) {
    uint32_t first = threadIdx.x & 63;
    const uint64_t vmask = __builtin_amdgcn_ballot_w64(valid);
    uint64_t todo = vmask;
    while (todo) {
        const uint32_t q = (uint32_t)__builtin_ctzll(todo);
        const uint32_t tq = __builtin_amdgcn_readlane(t, q);
        const uint64_t eq = __builtin_amdgcn_ballot_w64(t == tq) & vmask;
        if (valid && t == tq) first = q;
        todo &= ~eq;
    }
    return first;
}

// The same through a 64-slot hash table in LDS (tab: 128 words of this
// wavefront's): every valid lane writes its value into slot h(value); the
// lanes whose value holds the slot take the slot's minimum lane id.  All
// lanes of one value share a slot, so a value that lost its slot to another
// loses it on every lane; only those lanes run the ballot loop.
__device__ __forceinline__ uint32_t first_lane_hashed(uint32_t t, bool valid, uint32_t *tab) {
    const uint32_t lane = threadIdx.x & 63;
    uint32_t *key = tab, *low = tab + 64;
    key[lane] = 0xFFFFFFFFu;
    low[lane] = 64u;
    __builtin_amdgcn_wave_barrier();
    const uint32_t h = (t * 0x9E3779B1u) >> 26;
    if (valid) key[h] = t;
    __builtin_amdgcn_wave_barrier();
    const bool held = valid && key[h] == t;
    if (held) atomicMin(&low[h], lane);
    __builtin_amdgcn_wave_barrier();
    uint32_t first = held ? low[h] : lane;
    const bool rest = valid && !held;
    if (__builtin_amdgcn_ballot_w64(rest)) {
        const uint32_t f = first_lane_of(t, rest);
        if (rest) first = f;
    }
    __builtin_amdgcn_wave_barrier();
    return first;
}

__global__ __launch_bounds__(kBlock) void k_sgns_prep(Params a) {
    __shared__ uint32_t s_tok[kPrepWaves][MAXKEEP + 8];
    __shared__ uint32_t s_neg[kPrepWaves][OMAX];
    __shared__ uint32_t s_hash[kPrepWaves][128];
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    uint32_t *stok = s_tok[wv];
    uint32_t *sneg = s_neg[wv];
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const uint32_t w = a.window, k = a.negative, B = a.bgroup;
    unsigned long long my_words = 0, my_kept = 0;
    const uint64_t nwaves = (uint64_t)gridDim.x * kPrepWaves;
    const uint64_t n_sent = a.round_state->n_active;
    const uint32_t round = a.round_state->round;

    // Each wave walks its sentences one ahead: a sentence's offsets and
    // first 64 tokens are loaded while the one before it is processed, and
    // its keep thresholds are loaded together with the negatives of its
    // first groups (so one memory round trip per sentence is exposed).
    uint64_t sid = (uint64_t)blockIdx.x * kPrepWaves + wv;
    uint64_t b0 = 0, b0_nx = 0;
    uint32_t n = 0, n_nx = 0, t_first = 0;
    if (sid < n_sent) {
        b0 = a.off[sid];
        n = (uint32_t)(a.off[sid + 1] - b0);
        if (lane < n) t_first = a.tok[b0 + lane];
    }
    if (sid + nwaves < n_sent) {
        b0_nx = a.off[sid + nwaves];
        n_nx = (uint32_t)(a.off[sid + nwaves + 1] - b0_nx);
    }
    uint64_t b0_nn = 0;
    uint32_t n_nn = 0, t_nx = 0;
    auto advance = [&]() {
        b0 = b0_nx;
        n = n_nx;
        t_first = t_nx;
        b0_nx = b0_nn;
        n_nx = n_nn;
    };
    for (; sid < n_sent; sid += nwaves, advance()) {
        // The sentence's draw for index lane: the subsample coin of raw
        // position lane, the window shrink of kept centre lane and negative
        // lane % k of group lane / k (the first tile's, whose index is the
        // lane).  The alias entries of the negatives the first tile can hold
        // (k per group of at most ceil(n / B) groups) load with the
        // thresholds: each is a random line of the table.
        const u32x4 r0 = draw(a.seed, lane, DOM_SGNS, (uint32_t)sid, round);
        const uint32_t neg_bkt = alias_bucket(r0.z, a.V);
        uint2 neg_e = make_uint2(0, 0);
        if (lane < (uint32_t)OMAX && lane < k * ((n + B - 1) / B)) neg_e = a.uni[neg_bkt];
        const uint32_t thr_first = lane < n ? a.keep_thr[t_first] : 0u;
        // the next sentence's first tokens and the offsets of the one after
        t_nx = 0;
        if (sid + nwaves < n_sent && lane < n_nx) t_nx = a.tok[b0_nx + lane];
        b0_nn = 0;
        n_nn = 0;
        if (sid + 2 * nwaves < n_sent) {
            b0_nn = a.off[sid + 2 * nwaves];
            n_nn = (uint32_t)(a.off[sid + 2 * nwaves + 1] - b0_nn);
        }
        // ---- filter into stok[0, m) ----
        uint32_t m = 0, words = 0;
        for (uint32_t chunk = 0; chunk < n; chunk += 64) {
            const uint32_t p = chunk + lane;
            int status = 0;  // 0 none, 1 drop, 2 keep, 3 break
            uint32_t t = 0;
            if (p < n) {
                const uint32_t u = chunk == 0 ? r0.x : draw(a.seed, p, DOM_SGNS, (uint32_t)sid, round).x;
                t = chunk == 0 ? t_first : a.tok[b0 + p];
                // threshold 0: outside the vocabulary (or never kept)
                const uint32_t thr = chunk == 0 ? thr_first : a.keep_thr[t];
                if (a.compat_break && (a.in_vocab[t] == 0 || a.vocab_rank[t] == 0)) status = 3;
                else if (thr == 0u) status = 1;
                else status = (thr == 0xFFFFFFFFu || u < thr) ? 2 : 1;
            }
            const uint64_t brk_bal = __ballot(status == 3);
            const uint32_t brk = brk_bal ? chunk + (uint32_t)__ffsll((long long)brk_bal) - 1 : 0xFFFFFFFFu;
            const bool keep = status == 2 && p < brk;
            const uint64_t bal = __ballot(keep);
            const uint32_t idx = m + (uint32_t)__popcll(bal & lt_mask);
            if (keep && idx < (uint32_t)MAXKEEP) stok[idx] = t;
            uint32_t read_end = n < chunk + 64 ? n : chunk + 64;
            if (brk != 0xFFFFFFFFu) read_end = brk + 1;
            // the token that fills the sentence ends the read
            const uint64_t fill = __ballot(keep && idx == (uint32_t)MAXKEEP - 1);
            if (fill) read_end = chunk + (uint32_t)__ffsll((long long)fill);
            words = read_end;
            m += (uint32_t)__popcll(bal);
            if (brk != 0xFFFFFFFFu || m >= (uint32_t)MAXKEEP) break;
        }
        if (m > (uint32_t)MAXKEEP) m = MAXKEEP;
        my_words += words;
        my_kept += m;
        if (m < 2) {
            // slot sid stays empty: no centre, nothing to train
            if (lane == 0) {
                TileDesc *td = (TileDesc *)(a.tiles + sid * kDescBytes);
                td->nc = 0;
                td->n_in = 0;
            }
            continue;
        }
        __builtin_amdgcn_wave_barrier();

        // ---- tiles ----
        for (uint32_t g0 = 0; g0 * B < m;) {
            const uint32_t ng = tile_groups(g0, m, B, w, k, a.rmax);
            const uint32_t c0 = g0 * B;
            const uint32_t c_end = (g0 + ng) * B < m ? (g0 + ng) * B : m;
            const uint32_t nc = c_end - c0, nneg = k * ng;
            const uint32_t lo = c0 >= w ? c0 - w : 0;
            const uint32_t hi = c_end + w < m ? c_end + w : m;
            const uint32_t npos = hi - lo;
            // the first tile of sentence sid owns slot sid; further tiles take
            // slots after the n_sent first ones
            uint64_t slot = sid;
            if (g0 > 0) {
                if (lane == 0) slot = n_sent + atomicAdd(&a.counters[C_TILES], 1ull);
                slot = __shfl(slot, 0);
            }
            // negatives of the tile's groups (the first tile's drawn above)
            if (lane < nneg) {
                if (g0 == 0) {
                    sneg[lane] = r0.w < neg_e.x ? neg_bkt : neg_e.y;
                } else {
                    // negative lane % k of group g0 + lane / k: index g0 * k + lane
                    const u32x4 r = draw(a.seed, g0 * k + lane, DOM_SGNS, (uint32_t)sid, round);
                    const uint32_t bkt = alias_bucket(r.z, a.V);
                    const uint2 e = a.uni[bkt];
                    sneg[lane] = r.w < e.x ? bkt : e.y;
                }
            }
            __builtin_amdgcn_wave_barrier();
            if (slot >= a.tiles_cap) {
                g0 += ng;
                continue;  // counted; the host reports the overflow
            }
            TileDesc *td = (TileDesc *)(a.tiles + slot * kDescBytes);
            // distinct input rows over positions [lo, hi)
            uint32_t first_in;
            {
                const bool valid = lane < npos;
                const uint32_t t = valid ? stok[lo + lane] : 0xFFFFFFFFu;
                const uint32_t first = first_lane_hashed(t, valid, s_hash[wv]);
                first_in = first;
                const bool isf = valid && first == lane;
                const uint64_t bal = __ballot(isf);
                const uint32_t idx = (uint32_t)__popcll(bal & lt_mask);
                const uint32_t my = __shfl(idx, (int)first);
                const uint32_t n_in = (uint32_t)__popcll(bal);
                if (lane == 0 && n_in > 32) atomicAdd(&a.counters[C_BIGTILES], 1ull);
                if (isf) td->rin_tok[idx] = t;
                if (lane >= n_in) td->rin_tok[lane] = 0xFFFFFFFFu;
                td->pos_rin[lane] = valid ? (uint8_t)my : 0;
                if (lane == 0) {
                    td->sid = (uint32_t)sid; td->c0 = c0; td->nc = nc; td->g0 = g0;
                    td->lo = lo; td->npos = npos; td->n_in = n_in;
                }
            }
            // distinct output rows over [centres..., negatives...]
            {
                const uint32_t nl = nc + nneg;
                const bool valid = lane < nl;
                uint32_t t = 0xFFFFFFFFu;
                if (valid) t = lane < nc ? stok[c0 + lane] : sneg[lane - nc];
                uint32_t first;
                if (c0 == lo && nc == npos) {
                    // the centres are the input positions (a tile over the
                    // whole window span): their first lanes are the input
                    // pass's; each negative takes the lowest lane holding
                    // its row
                    first = lane < nc ? first_in : lane;
                    for (uint32_t q = nc; q < nl; ++q) {
                        const uint32_t tq = __builtin_amdgcn_readlane(t, q);
                        const uint64_t eq = __builtin_amdgcn_ballot_w64(valid && t == tq);
                        if (lane == q) first = (uint32_t)__builtin_ctzll(eq);
                    }
                } else {
                    first = first_lane_of(t, valid);
                }
                const bool isf = valid && first == lane;
                const uint64_t bal = __ballot(isf);
                const uint32_t idx = (uint32_t)__popcll(bal & lt_mask);
                const uint32_t my = __shfl(idx, (int)first);
                const uint32_t n_out = (uint32_t)__popcll(bal);
                // output row of negative q sits on lane nc + q
                const uint32_t my_neg = __shfl(my, (int)((nc + lane) & 63));
                // negatives of a centre's group equal to the centre (word2vec
                // skips them): its group is lane / B within the tile (c0 is a
                // group boundary; the float quotient of integers below 64 is
                // exact after truncation), its negatives sit on lanes
                // nc + group * k + j
                uint32_t coll = 0;
                {
                    const uint32_t grp = (uint32_t)((float)lane / (float)B);
                    const uint32_t src = nc + grp * k;
                    for (uint32_t j = 0; j < k; ++j) coll += __shfl(t, (int)((src + j) & 63)) == t;
                }
                if (lane < OMAX) {
                    if (isf) td->rout_tok[idx] = t;
                    if (lane >= n_out) td->rout_tok[lane] = 0xFFFFFFFFu;
                    td->cen_rout[lane] = (valid && lane < nc) ? (uint8_t)my : 0;
                    td->neg_rout[lane] = lane < nneg ? (uint8_t)my_neg : 0;
                    // window shrink of kept centre c0 + lane (the first tile's: r0)
                    uint32_t by = r0.y;
                    if (c0 != 0) by = draw(a.seed, c0 + lane, DOM_SGNS, (uint32_t)sid, round).y;
                    td->bshrink[lane] = lane < nc ? (uint8_t)(by % w) : 0;
                    td->cen_coll[lane] = lane < nc ? (uint8_t)coll : 0;
                }
                if (lane == 0) td->n_out = n_out;
            }
            g0 += ng;
            __builtin_amdgcn_wave_barrier();
        }
    }
    // my_words and my_kept are wave-uniform; one atomic per workgroup and
    // counter (atomics on one address serialise)
    __shared__ unsigned long long part[2][kPrepWaves];
    if (lane == 0) {
        part[0][threadIdx.x >> 6] = my_words;
        part[1][threadIdx.x >> 6] = my_kept;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < kPrepWaves; ++i) {
            my_words += part[0][i];
            my_kept += part[1][i];
        }
        atomicAdd(&a.counters[C_WORDS], my_words);
        atomicAdd(&a.counters[C_KEPT], my_kept);
    }
}

}  // namespace
}  // namespace sgns

uint32_t sgns_group_centres(uint32_t batch, uint32_t window, uint32_t negative) {
    uint32_t b = batch == 0 ? 2 * window + 1 : batch;
    if (b + negative > (uint32_t)sgns::OMAX) b = sgns::OMAX - negative;
    if (b + 2 * window > (uint32_t)sgns::RMAX) b = sgns::RMAX - 2 * window;
    return b;
}

uint32_t sgns_tile_rows(uint32_t centres, uint32_t window) {
    return centres + 2 * window <= 48 ? 48u : 64u;
}

#define HIP_TRY(x)                              \
    do {                                        \
        hipError_t e_ = (x);                    \
        if (e_ != hipSuccess) return e_;        \
    } while (0)

hipError_t sgns_reserve(felog_ctx *c, uint64_t n_sent, uint64_t n_tok) {
    using namespace sgns;
    const uint64_t need = sgns_tiles_needed(c, n_sent, n_tok);
    if (need <= c->tiles_cap) return hipSuccess;
    drop_round_graph(c);
    dfree(c->d_tiles);
    c->d_tiles = nullptr;
    c->tiles_cap = 0;
    const uint64_t cap = need + need / 4;
    HIP_TRY(dalloc(c, &c->d_tiles, cap * kDescBytes));
    c->tiles_cap = cap;
    return hipSuccess;
}

hipError_t sgns_enqueue(felog_ctx *c, const uint64_t *d_off, const uint32_t *d_tok, const SgnsDebug *dbg,
                        hipEvent_t mid) {
    return sgns_enqueue_into(c, d_off, d_tok, dbg, mid, c->d_tiles, c->tiles_cap);
}

hipError_t sgns_enqueue_into(felog_ctx *c, const uint64_t *d_off, const uint32_t *d_tok, const SgnsDebug *dbg,
                             hipEvent_t mid, void *tiles, uint64_t tiles_cap) {
    using namespace sgns;
    Params a;
    a.off = d_off;
    a.tok = d_tok;
    a.round_state = c->d_rs;
    a.syn0 = c->d_syn0;
    a.syn1 = c->d_syn1;
    a.dp = c->dp;
    a.d = c->d;
    a.in_vocab = c->d_in_vocab;
    a.vocab_rank = c->d_vocab_rank;
    a.keep_thr = c->d_keep_thr;
    a.uni = c->d_uni;
    a.V = c->V;
    a.window = c->opt.window;
    a.negative = c->opt.negative;
    a.bgroup = sgns_group_centres(c->opt.batch_size, c->opt.window, c->opt.negative);
    // tiles are packed for 48 or 64 rows (a 32-row shape splits them by size)
    a.rmax = c->tile_rows >= 48 ? c->tile_rows : sgns_tile_rows(a.bgroup, c->opt.window);
    {
        // windows per wavefront 1, 2, 4 or 8: 64, 32, 16 or 8 lanes per window
        const uint32_t wpw = c->tile_wpw ? c->tile_wpw : 8u;
        a.win_shift = wpw >= 8 ? 3u : wpw >= 4 ? 4u : wpw >= 2 ? 5u : 6u;
        a.col_parts = c->tile_parts ? c->tile_parts : 1u;
    }
    a.compat_break = c->opt.compat_vocab_break;
    a.writeback = c->opt.update_mode != FELOG_UPDATE_ATOMIC;
    a.seed = c->opt.seed;
    a.counters = (unsigned long long *)c->d_counters;
    a.dbg_max = dbg ? dbg->max_groups : 0;
    a.dbg_groups = dbg ? dbg->groups : nullptr;
    a.dbg_rows = dbg ? dbg->rows : nullptr;
    a.dbg_logits = dbg ? dbg->logits : nullptr;
    a.dbg_dv = dbg ? dbg->dv : nullptr;
    a.dbg_du = dbg ? dbg->du : nullptr;
    a.tiles = (unsigned char *)tiles;
    a.tiles_cap = tiles_cap;
    // lost updates are a property of write-back tiles that train the tables
    a.row_ver = (dbg || !a.writeback) ? nullptr : c->d_row_ver;
    // the sentence count is read on the device: a fixed grid strides over it
    k_sgns_prep<<<c->cu_count * 16, kBlock, 0, c->stream>>>(a);
    HIP_TRY(hipGetLastError());
    if (mid) HIP_TRY(hipEventRecord(mid, c->stream));
    return launch_tile(c, a);
}

static_assert(kSgnsDescBytes == (uint64_t)sgns::kDescBytes, "tile descriptor size");

uint64_t sgns_tiles_needed(const felog_ctx *c, uint64_t n_sent, uint64_t n_tok) {
    // every tile holds at least one centre and every sentence ends one tile
    const uint32_t B = sgns_group_centres(c->opt.batch_size, c->opt.window, c->opt.negative);
    return n_tok / B + n_sent + 64;
}

hipError_t run_sgns(felog_ctx *c, const uint64_t *d_off, const uint32_t *d_tok, uint64_t n_sent, uint64_t n_tok,
                    const SgnsDebug *dbg, float *ms) {
    HIP_TRY(sgns_reserve(c, n_sent, n_tok));
    return run_sgns_into(c, d_off, d_tok, dbg, ms, c->d_tiles, c->tiles_cap);
}

hipError_t run_sgns_into(felog_ctx *c, const uint64_t *d_off, const uint32_t *d_tok, const SgnsDebug *dbg,
                         float *ms, void *tiles, uint64_t tiles_cap) {
    HIP_TRY(hipMemsetAsync(c->d_counters + C_WORDS, 0, 8 * (C_BIGTILES - C_WORDS + 1), c->stream));
    HIP_TRY(hipMemsetAsync(c->d_counters + C_ROWS_WB, 0, 8 * 2, c->stream));
    HIP_TRY(hipEventRecord(c->ev0, c->stream));
    HIP_TRY(sgns_enqueue_into(c, d_off, d_tok, dbg, c->ev1, tiles, tiles_cap));
    HIP_TRY(hipEventRecord(c->ev2, c->stream));
    HIP_TRY(hipEventSynchronize(c->ev2));
    HIP_TRY(hipEventElapsedTime(&ms[0], c->ev0, c->ev2));
    HIP_TRY(hipEventElapsedTime(&ms[1], c->ev0, c->ev1));
    HIP_TRY(hipEventElapsedTime(&ms[2], c->ev1, c->ev2));
    HIP_TRY(hipMemcpy(c->h_counters, c->d_counters, kNumCounters * 8, hipMemcpyDeviceToHost));
    return hipSuccess;
}

}  // namespace felog
