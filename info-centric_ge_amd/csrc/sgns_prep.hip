// sgns_prep.hip - first half of an SGNS training pass, and the pass driver.
//
// k_sgns_prep runs one sentence per wavefront, or two of at most 32 tokens
// per wavefront (one per half), and does the integer work of the
// skip-gram pass (reference paths relative to the reference repository):
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
// Workgroups per CU of the two-sentence launch (6 are resident at 76 VGPRs):
// at 2.24M sentences a grid of 16 per CU took 1.32 ms, 48 1.25, 96 1.23 and
// 192 1.42 ms; above that the time grows with the workgroup count.
constexpr int kPairGridPerCu = 96;

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

__device__ __forceinline__ uint64_t readlane64(uint64_t x, uint32_t lane) {
    const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)x, lane);
    const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(x >> 32), lane);
    return ((uint64_t)hi << 32) | lo;
}

// Pair mode: lanes 0-31 hold one sentence, lanes 32-63 the next.  The
// lowest lane of this lane's half whose valid value equals this lane's
// (the lane itself for an invalid lane), as first_lane_of with every
// ballot cut to the half of the value it looks for.
__device__ __forceinline__ uint32_t first_lane_of_pair(uint32_t t, bool valid) {
    const uint32_t lane = threadIdx.x & 63;
    uint32_t first = lane;
    const uint64_t vmask = __builtin_amdgcn_ballot_w64(valid);
    uint64_t todo = vmask;
    while (todo) {
        const uint32_t q = (uint32_t)__builtin_ctzll(todo);
        const uint32_t tq = __builtin_amdgcn_readlane(t, q);
        const uint64_t qh = q < 32 ? 0x00000000FFFFFFFFull : 0xFFFFFFFF00000000ull;
        const uint64_t eq = __builtin_amdgcn_ballot_w64(t == tq) & vmask & qh;
        if (valid && t == tq && (lane >> 5) == (q >> 5)) first = q;
        todo &= ~eq;
    }
    return first;
}

// The same through a 32-slot hash table per half in LDS (tab: the
// wavefront's 128 words, half h using words 64h .. 64h + 63), as
// first_lane_hashed.
__device__ __forceinline__ uint32_t first_lane_hashed_pair(uint32_t t, bool valid, uint32_t *tab) {
    const uint32_t lane = threadIdx.x & 63, hl = lane & 31;
    uint32_t *key = tab + (lane >> 5) * 64, *low = key + 32;
    key[hl] = 0xFFFFFFFFu;
    low[hl] = 64u;
    __builtin_amdgcn_wave_barrier();
    const uint32_t h = (t * 0x9E3779B1u) >> 27;
    if (valid) key[h] = t;
    __builtin_amdgcn_wave_barrier();
    const bool held = valid && key[h] == t;
    if (held) atomicMin(&low[h], lane);
    __builtin_amdgcn_wave_barrier();
    uint32_t first = held ? low[h] : lane;
    const bool rest = valid && !held;
    if (__builtin_amdgcn_ballot_w64(rest)) {
        const uint32_t f = first_lane_of_pair(t, rest);
        if (rest) first = f;
    }
    __builtin_amdgcn_wave_barrier();
    return first;
}

// kPairs: the pairs of sentences of at most 32 tokens each; else every
// sentence of the other pairs, one at a time.  Two launches, one of each,
// cover a round: kept apart, each path gets its own register allocation
// (about 76 VGPRs each, 100 together).
template <bool kPairs>
__global__ __launch_bounds__(kBlock) void k_sgns_prep(Params a) {
    __shared__ uint32_t s_tok[kPrepWaves][MAXKEEP + 8];
    __shared__ uint32_t s_neg[kPrepWaves][OMAX];
    __shared__ uint32_t s_hash[kPrepWaves][128];
    // no sentence over 32 tokens: nothing for the whole-wavefront launch
    if (!kPairs && a.max_sent_len && *a.max_sent_len <= 32) return;
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    uint32_t *stok = s_tok[wv];
    uint32_t *sneg = s_neg[wv];
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const uint32_t w = a.window, k = a.negative, B = a.bgroup;
    unsigned long long my_words = 0, my_kept = 0;
    const uint64_t nwaves = (uint64_t)gridDim.x * kPrepWaves;
    const uint64_t n_sent = a.round_state->n_active;
    const uint32_t round = a.round_state->round;

    // One sentence of any length on the whole wavefront.
    auto one = [&](uint64_t sid, uint64_t b0, uint32_t n) {
        uint32_t t_first = 0;
        if (lane < n) t_first = a.tok[b0 + lane];
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
            return;
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
    };

    // Two sentences of at most 32 tokens, one per half of the wavefront
    // (lanes 0-31: sentence s0, lanes 32-63: s0 + 1): the steps of `one`
    // with lane hl = lane & 31 in place of lane, every ballot cut to the
    // lane's half and every broadcast taken from the half's own lanes.  Each
    // lane holds token hl of its half's sentence in t_first.  A half past
    // the last sentence has no tokens and no slot.
    auto pair = [&](uint64_t s0, uint32_t nA, uint32_t nB, uint32_t t_first) {
        const uint32_t hl = lane & 31, hb = lane & 32;
        const uint64_t hm = hb ? 0xFFFFFFFF00000000ull : 0x00000000FFFFFFFFull;
        const uint64_t lt_half = lt_mask & hm;
        const uint64_t sid = s0 + (hb >> 5);
        const bool present = sid < n_sent;
        const uint32_t n = present ? (hb ? nB : nA) : 0u;
        uint32_t *ptok = stok + 2 * hb;    // kept tokens: 64 words per half
        uint32_t *pneg = stok + 128 + hb;  // negatives: 32 words per half
        const u32x4 r0 = draw(a.seed, hl, DOM_SGNS, (uint32_t)sid, round);
        const uint32_t neg_bkt = alias_bucket(r0.z, a.V);
        uint2 neg_e = make_uint2(0, 0);
        if (hl < k * ((n + B - 1) / B)) neg_e = a.uni[neg_bkt];
        const uint32_t thr = hl < n ? a.keep_thr[t_first] : 0u;
        // ---- filter (one chunk) ----
        const uint32_t t = t_first;
        int status = 0;  // 0 none, 1 drop, 2 keep, 3 break
        if (hl < n) {
            if (a.compat_break && (a.in_vocab[t] == 0 || a.vocab_rank[t] == 0)) status = 3;
            else if (thr == 0u) status = 1;
            else status = (thr == 0xFFFFFFFFu || r0.x < thr) ? 2 : 1;
        }
        const uint64_t brk_bal = __ballot(status == 3) & hm;
        const uint32_t brk = brk_bal ? (uint32_t)__builtin_ctzll(brk_bal) - hb : 0xFFFFFFFFu;
        const bool keep = status == 2 && hl < brk;
        const uint64_t bal = __ballot(keep) & hm;
        if (keep) ptok[__popcll(bal & lt_half)] = t;
        const uint32_t m = (uint32_t)__popcll(bal);
        const uint32_t words = brk != 0xFFFFFFFFu ? brk + 1 : n;
        my_words += __builtin_amdgcn_readlane(words, 0) + __builtin_amdgcn_readlane(words, 32);
        my_kept += __builtin_amdgcn_readlane(m, 0) + __builtin_amdgcn_readlane(m, 32);
        if (present && m < 2 && hl == 0) {
            // slot sid stays empty: no centre, nothing to train
            TileDesc *td = (TileDesc *)(a.tiles + sid * kDescBytes);
            td->nc = 0;
            td->n_in = 0;
        }
        __builtin_amdgcn_wave_barrier();

        // ---- tiles: each half runs its own, finished halves idle ----
        uint32_t g0 = 0;
        bool more = present && m >= 2;
        while (__ballot(more)) {
            uint32_t ng = 0, c0 = 0, nc = 0, nneg = 0, lo = 0, npos = 0;
            if (more) {
                ng = tile_groups(g0, m, B, w, k, a.rmax);
                c0 = g0 * B;
                const uint32_t c_end = (g0 + ng) * B < m ? (g0 + ng) * B : m;
                nc = c_end - c0;
                nneg = k * ng;
                lo = c0 >= w ? c0 - w : 0;
                const uint32_t hi = c_end + w < m ? c_end + w : m;
                npos = hi - lo;
            }
            uint64_t slot = sid;
            if (more && g0 > 0 && hl == 0) slot = n_sent + atomicAdd(&a.counters[C_TILES], 1ull);
            slot = __shfl(slot, (int)hb);
            if (more && hl < nneg) {
                if (g0 == 0) {
                    pneg[hl] = r0.w < neg_e.x ? neg_bkt : neg_e.y;
                } else {
                    const u32x4 r = draw(a.seed, g0 * k + hl, DOM_SGNS, (uint32_t)sid, round);
                    const uint32_t bkt = alias_bucket(r.z, a.V);
                    const uint2 e = a.uni[bkt];
                    pneg[hl] = r.w < e.x ? bkt : e.y;
                }
            }
            __builtin_amdgcn_wave_barrier();
            // a slot past the buffer is counted; the host reports the overflow
            const bool live = more && slot < a.tiles_cap;
            TileDesc *td = (TileDesc *)(a.tiles + slot * kDescBytes);
            // distinct input rows over positions [lo, lo + npos)
            uint32_t first_in;
            {
                const bool valid = live && hl < npos;
                const uint32_t ti = valid ? ptok[lo + hl] : 0xFFFFFFFFu;
                const uint32_t first = first_lane_hashed_pair(ti, valid, s_hash[wv]);
                first_in = first;
                const bool isf = valid && first == lane;
                const uint64_t bi = __ballot(isf) & hm;
                const uint32_t idx = (uint32_t)__popcll(bi & lt_half);
                const uint32_t my = __shfl(idx, (int)first);
                const uint32_t n_in = (uint32_t)__popcll(bi);
                if (live) {
                    if (isf) td->rin_tok[idx] = ti;
                    if (hl >= n_in) td->rin_tok[hl] = 0xFFFFFFFFu;
                    td->rin_tok[32 + hl] = 0xFFFFFFFFu;
                    td->pos_rin[hl] = valid ? (uint8_t)my : 0;
                    td->pos_rin[32 + hl] = 0;
                    if (hl == 0) {
                        td->sid = (uint32_t)sid; td->c0 = c0; td->nc = nc; td->g0 = g0;
                        td->lo = lo; td->npos = npos; td->n_in = n_in;
                    }
                }
            }
            // distinct output rows over [centres..., negatives...]
            {
                const uint32_t nl = nc + nneg;
                const bool valid = live && hl < nl;
                uint32_t to = 0xFFFFFFFFu;
                if (valid) to = hl < nc ? ptok[c0 + hl] : pneg[hl - nc];
                // a tile over its whole window span: the centres' first lanes
                // are the input pass's, each negative takes the lowest lane of
                // its half holding its row; other tiles dedupe from scratch
                const bool span = c0 == lo && nc == npos;
                uint32_t first = (span && hl < nc) ? first_in : lane;
                {
                    const uint32_t f = first_lane_of_pair(to, valid && !span);
                    if (valid && !span) first = f;
                }
                const uint32_t jA = __builtin_amdgcn_readlane((live && span) ? nneg : 0u, 0);
                const uint32_t jB = __builtin_amdgcn_readlane((live && span) ? nneg : 0u, 32);
                for (uint32_t j = 0; j < (jA > jB ? jA : jB); ++j) {
                    const uint32_t src = hb + ((nc + j) & 31);
                    const uint32_t tq = __shfl(to, (int)src);
                    const uint64_t eq = __ballot(valid && to == tq) & hm;
                    if (span && j < nneg && lane == src) first = (uint32_t)__builtin_ctzll(eq);
                }
                const bool isf = valid && first == lane;
                const uint64_t bo = __ballot(isf) & hm;
                const uint32_t idx = (uint32_t)__popcll(bo & lt_half);
                const uint32_t my = __shfl(idx, (int)first);
                const uint32_t n_out = (uint32_t)__popcll(bo);
                // output row of negative q sits on lane nc + q of the half
                const uint32_t my_neg = __shfl(my, (int)(hb + ((nc + hl) & 31)));
                // negatives of a centre's group equal to the centre (see one)
                uint32_t coll = 0;
                {
                    const uint32_t grp = (uint32_t)((float)hl / (float)B);
                    const uint32_t src = nc + grp * k;
                    for (uint32_t j = 0; j < k; ++j) coll += __shfl(to, (int)(hb + ((src + j) & 31))) == to;
                }
                if (live) {
                    if (isf) td->rout_tok[idx] = to;
                    if (hl >= n_out) td->rout_tok[hl] = 0xFFFFFFFFu;
                    td->cen_rout[hl] = (valid && hl < nc) ? (uint8_t)my : 0;
                    td->neg_rout[hl] = hl < nneg ? (uint8_t)my_neg : 0;
                    // window shrink of kept centre c0 + hl (the first tile's: r0)
                    uint32_t by = r0.y;
                    if (c0 != 0) by = draw(a.seed, c0 + hl, DOM_SGNS, (uint32_t)sid, round).y;
                    td->bshrink[hl] = hl < nc ? (uint8_t)(by % w) : 0;
                    td->cen_coll[hl] = hl < nc ? (uint8_t)coll : 0;
                    if (hl == 0) td->n_out = n_out;
                }
            }
            if (more) {
                g0 += ng;
                more = g0 * B < m;
            }
            __builtin_amdgcn_wave_barrier();
        }
    };

    // Sentences go to wavefronts in pairs (2p, 2p + 1).  A pair of
    // sentences of at most 32 tokens each is done in one pass, one sentence
    // per half of the wavefront; otherwise each takes the whole wavefront in
    // turn.  A pair's offsets and tokens are loaded while the pair before it
    // is processed.
    auto offsets = [&](uint64_t p, uint64_t &bA, uint32_t &nA, uint64_t &bB, uint32_t &nB) {
        bA = bB = 0;
        nA = nB = 0;
        if (2 * p < n_sent) {
            bA = a.off[2 * p];
            bB = a.off[2 * p + 1];
            nA = (uint32_t)(bB - bA);
            if (2 * p + 1 < n_sent) nB = (uint32_t)(a.off[2 * p + 2] - bB);
        }
    };
    // token lane & 31 of the lane's half's sentence, for a pair of short ones
    auto pair_tokens = [&](uint64_t bA, uint32_t nA, uint64_t bB, uint32_t nB) -> uint32_t {
        const uint32_t hl = lane & 31;
        const uint32_t nh = lane < 32 ? nA : nB;
        if (nA > 32 || nB > 32 || hl >= nh) return 0u;
        return a.tok[(lane < 32 ? bA : bB) + hl];
    };
    const uint64_t wave_id = (uint64_t)blockIdx.x * kPrepWaves + wv;
    if constexpr (kPairs) {
        uint64_t pr = wave_id;
        uint64_t bA, bB, bA_nx, bB_nx;
        uint32_t nA, nB, nA_nx, nB_nx;
        offsets(pr, bA, nA, bB, nB);
        uint32_t t_first = pair_tokens(bA, nA, bB, nB);
        offsets(pr + nwaves, bA_nx, nA_nx, bB_nx, nB_nx);
        for (; 2 * pr < n_sent; pr += nwaves) {
            // the next pair's tokens and the offsets of the pair after it
            const uint32_t t_nx = pair_tokens(bA_nx, nA_nx, bB_nx, nB_nx);
            uint64_t bA_nn, bB_nn;
            uint32_t nA_nn, nB_nn;
            offsets(pr + 2 * nwaves, bA_nn, nA_nn, bB_nn, nB_nn);
            if (nA <= 32 && nB <= 32) pair(2 * pr, nA, nB, t_first);
            bA = bA_nx;
            nA = nA_nx;
            bB = bB_nx;
            nB = nB_nx;
            t_first = t_nx;
            bA_nx = bA_nn;
            nA_nx = nA_nn;
            bB_nx = bB_nn;
            nB_nx = nB_nn;
        }
    } else {
        // the other pairs: 64 pairs tested at a time, lane j pair base + j,
        // and only those holding a sentence of more than 32 tokens are run
        for (uint64_t base = wave_id * 64; 2 * base < n_sent; base += nwaves * 64) {
            const uint64_t p = base + lane;
            uint64_t o0 = 0, o1 = 0, o2 = 0;
            bool big = false;
            if (2 * p < n_sent) {
                o0 = a.off[2 * p];
                o1 = a.off[2 * p + 1];
                o2 = 2 * p + 1 < n_sent ? a.off[2 * p + 2] : o1;
                big = o1 - o0 > 32 || o2 - o1 > 32;
            }
            for (uint64_t bm = __ballot(big); bm; bm &= bm - 1) {
                const uint32_t j = (uint32_t)__builtin_ctzll(bm);
                const uint64_t q = base + j;
                const uint64_t bA = readlane64(o0, j), bB = readlane64(o1, j), eB = readlane64(o2, j);
                one(2 * q, bA, (uint32_t)(bB - bA));
                if (2 * q + 1 < n_sent) one(2 * q + 1, bB, (uint32_t)(eB - bB));
            }
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
        // a workgroup with no sentence adds nothing
        if (my_words) atomicAdd(&a.counters[C_WORDS], my_words);
        if (my_kept) atomicAdd(&a.counters[C_KEPT], my_kept);
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
    // the library's corpus buffers hold the last walk round's paths, whose
    // longest length the walk counted; other corpora are scanned
    a.max_sent_len = (d_off == c->d_off && d_tok == c->d_tok) ? a.counters + C_MAXLEN : nullptr;
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
    k_sgns_prep<true><<<c->cu_count * kPairGridPerCu, kBlock, 0, c->stream>>>(a);
    k_sgns_prep<false><<<c->cu_count * 16, kBlock, 0, c->stream>>>(a);
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
