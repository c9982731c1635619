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
// Work per tile and workgroup (its rows were gathered into registers while
// the previous tile trained):
//   1. count tile cleared, pair counts built with LDS atomics;
//   2. gathered rows stored to LDS, the next descriptor to its slot;
//   3. the next tile's row gathers issued into registers;
//   4. X and G (G^T overwrites the counts in place);
//   5. dVin and dUout, each 16x16 output tile written straight to the table.
// Every loop index is advanced incrementally (no integer division in the
// tile loop) and each lane's table row addresses are formed once per 16-row
// block.  Latency of the gathers is hidden by the pipeline and by several
// workgroups per CU (at d = 128 in f32 the LDS allows four of the 32-row
// instance, three of the 48-row one).  Tiles with at most 32 distinct input
// rows run on the 32-row instance, the rest on a 48/64-row instance
// launched after it.
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <type_traits>
#include "felog_dev.h"
#include "felog_internal.h"
#include "sgns_common.h"

namespace felog {
namespace sgns {

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

static_assert(offsetof(TileDesc, rout_tok) == offsetof(TileDesc, rin_tok) + 4 * RMAX,
              "output row ids follow the input row ids");

typedef short s4 __attribute__((ext_vector_type(4)));
typedef short s8 __attribute__((ext_vector_type(8)));
typedef unsigned int u4v __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) s4 lds_s4;
typedef __attribute__((address_space(3))) unsigned char lds_u8;

// Pitch (words) of the OMAX-row count / G^T tile.  bf16 tables contract
// over 32 input rows per MFMA, so the rows are padded to a multiple of 32.
__host__ __device__ constexpr uint32_t gt_pitch(uint32_t esize, uint32_t rmax) {
    return esize == 4 ? rmax + 4 : (rmax + 31) / 32 * 32 + 4;
}

// Groups of one tile: each has at least one centre and k output slots of
// negatives, and the first group always fits, so at most
// max(1, OMAX / (k + 1)) (5 at k = 5).
__host__ __device__ inline uint32_t tile_wrows(uint32_t k) {
    const uint32_t g = OMAX / (k + 1);
    return g ? g : 1;
}

// LDS: two descriptor slots (this tile's, the next one's) | RM input rows
// then OMAX output rows (pitch ldv) | the OMAX-row count / G^T tile, then
// one row of context counts per group (same pitch) | per-output-row masks
// of negative slots | the staged rows' versions (lost-update tracking) |
// the tile cursor's broadcast word.  At k = 5 and
// d = 128 the f32 32-row layout (40.5 KB)
// fits four workgroups in a CU's 160 KB, the bf16 one six.
struct Lds {
    uint32_t ldv;    // row pitch of staged rows, in elements
    uint32_t wrows;  // rows of per-group context counts
    size_t o_desc, o_rows, o_g, o_mask, o_ver, o_grab, total;
};

__host__ __device__ inline Lds lds_layout(uint32_t dp, uint32_t esize, uint32_t rmax, uint32_t k, bool trk = false) {
    Lds L;
    // skew between rows: 16 bytes for f32; 32 bytes for bf16, whose
    // transposing loads conflict at 16 (the bf16 tile runs 21% faster)
    L.ldv = dp + (esize == 2 ? 16u : 4u);
    L.wrows = tile_wrows(k);
    size_t o = 0;
    L.o_desc = o; o += 2 * kDescBytes;
    L.o_rows = o; o += (size_t)(rmax + OMAX) * L.ldv * esize;
    L.o_g = o;    o += (size_t)(OMAX + L.wrows) * gt_pitch(esize, rmax) * 4;
    L.o_mask = o; o += (size_t)OMAX * 4;
    L.o_ver = o;  o += trk ? (size_t)(rmax + OMAX) * 4 : 0;
    L.o_grab = o; o += 8;
    L.total = o;
    return L;
}

// ds_read_b64_tr_b16: within each 16-lane group, lane 4j+q supplies the
// address of four consecutive 16-bit elements (chunk C[4j+q]) and lane i
// receives element i%4 of chunks C[i/4], C[4+i/4], C[8+i/4], C[12+i/4];
// i.e. the group reads a 4 x 16 block whose row j is C[4j..4j+3] and lane i
// gets its column i.  Pointing the lanes of row j at row k_j of a row-major
// matrix (columns n0..n0+15) hands lane i the four values M[k_j][n0+i].
__device__ __forceinline__ s4 tr_load(lds_u8 *base, uint32_t byte_off) {
    return __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_s4 *)(base + byte_off));
}

__device__ __forceinline__ bf8 join8(s4 a, s4 b) {
    return __builtin_bit_cast(bf8, __builtin_shufflevector(a, b, 0, 1, 2, 3, 4, 5, 6, 7));
}

__device__ __forceinline__ float load_elem(const float *p) { return *p; }
__device__ __forceinline__ float load_elem(const uint16_t *p) { return bf16_to_f32(*p); }

// s_waitcnt immediate for vmcnt(0) with expcnt and lgkmcnt left at their
// maxima (gfx9 encoding: vmcnt[3:0] | expcnt[6:4] | lgkmcnt[11:8] |
// vmcnt[15:14])
constexpr unsigned kWaitVm0 = 0x0F70u;

// descriptor slots per cursor request of the 32-row launch (the 48/64-row
// launch takes 64: most of its slots hold small tiles)
constexpr uint32_t kChunk32 = 16;

// sigma(x) with word2vec's clamp (0 below -MAX_EXP, 1 above), branch-free
__device__ __forceinline__ float sigmoid_clamped(float x) {
    const float s = __builtin_amdgcn_rcpf(1.f + __expf(-x));
    return x > kMaxExp ? 1.f : (x < -kMaxExp ? 0.f : s);
}

// alpha (P - (P+Q) sigma(x)) for a packed count word (P low, Q high half)
__device__ __forceinline__ float pair_grad(uint32_t cc, float x, float alpha) {
    const float P = (float)(cc & 0xFFFFu);
    const float PQ = (float)((cc & 0xFFFFu) + (cc >> 16));
    return alpha * (P - PQ * sigmoid_clamped(x));
}

// 32 rounding bits: Marsaglia's xorshift32 (13, 17, 5), seeded by one
// Philox draw per lane and tile (a zero seed is replaced)
__device__ __forceinline__ uint32_t round_bits(uint32_t &s) {
    s ^= s << 13;
    s ^= s >> 17;
    s ^= s << 5;
    return s;
}

// bf16 pair += (d0, d1) with stochastic rounding, by compare-and-swap on the
// containing 32-bit word (no delta lost).
__device__ __forceinline__ void bf16_pair_add_sr(uint16_t *p, float d0, float d1, uint32_t rnd) {
    uint32_t *w = reinterpret_cast<uint32_t *>(p);
    uint32_t old = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    while (true) {
        const float lo = bf16_to_f32((uint16_t)(old & 0xFFFFu)) + d0;
        const float hi = bf16_to_f32((uint16_t)(old >> 16)) + d1;
        const uint32_t nw = f32x2_to_bf16x2_sr(lo, hi, rnd);
        const uint32_t prev = atomicCAS(w, old, nw);
        if (prev == old) break;
        old = prev;
    }
}

// dVin and dUout are formed transposed (operands swapped in the MFMA), so
// lane l holds row 16*blk + l%16 of the output and its four consecutive
// columns col0 = 16*dt + 4*(l/16) .. col0+3: one table row per lane and one
// contiguous store per 16x16 tile.
template <typename T>
struct LaneRow {
    T *p;     // table row
    bool ok;  // row exists in the tile
};

template <typename T>
__device__ __forceinline__ LaneRow<T> lane_row(T *table, const uint32_t *ids, uint32_t r, uint32_t n,
                                               uint32_t dp) {
    LaneRow<T> R;
    R.ok = r < n;
    R.p = table + (uint64_t)ids[R.ok ? r : 0] * dp;
    return R;
}

// f32 table: the lane's four values (columns col0..col0+3 of its row;
// padding columns past d hold zero deltas on zero rows, so whole vectors are
// written).  Write-back: acc already holds the staged value plus the delta
// (the products accumulate onto the staged row).  Atomic: acc is the delta.
template <typename T, bool WB>
__device__ __forceinline__ void store_row(const LaneRow<T> &R, uint32_t col0, const f4 &acc) {
    if (!R.ok) return;
    float *dst = (float *)R.p + col0;
    if constexpr (WB) {
        *(f4 *)dst = acc;
    } else {
#pragma unroll
        for (int j = 0; j < 4; ++j) unsafeAtomicAdd(dst + j, acc[j]);
    }
}

// The staged bf16 values of columns c0..c0+7 (the accumulators' start for a
// write-back update)
__device__ __forceinline__ void staged_bf16x8(const uint16_t *staged, uint32_t c0, f4 &lo, f4 &hi) {
    const uint4 v = *(const uint4 *)(staged + c0);
    lo = f4{__uint_as_float(v.x << 16), __uint_as_float(v.x & 0xFFFF0000u), __uint_as_float(v.y << 16),
            __uint_as_float(v.y & 0xFFFF0000u)};
    hi = f4{__uint_as_float(v.z << 16), __uint_as_float(v.z & 0xFFFF0000u), __uint_as_float(v.w << 16),
            __uint_as_float(v.w & 0xFFFF0000u)};
}

// bf16 table: the lane's eight consecutive columns c0..c0+7 of its row.
// Write-back: lo/hi hold staged value + delta, rounded stochastically (16
// random bits per element) and stored as one 16-byte vector.  Atomic: the
// deltas are added by compare-and-swap pairs.  The rounding stream advances
// on every lane, live row or not.
template <bool WB>
__device__ __forceinline__ void store_row8(const LaneRow<uint16_t> &R, uint32_t c0, const f4 &lo, const f4 &hi,
                                           uint32_t &rs) {
    uint16_t *dst = R.p + c0;
    const uint32_t b0 = round_bits(rs), b1 = round_bits(rs), b2 = round_bits(rs), b3 = round_bits(rs);
    if constexpr (WB) {
        uint4 nw;
        nw.x = f32x2_to_bf16x2_sr(lo[0], lo[1], b0);
        nw.y = f32x2_to_bf16x2_sr(lo[2], lo[3], b1);
        nw.z = f32x2_to_bf16x2_sr(hi[0], hi[1], b2);
        nw.w = f32x2_to_bf16x2_sr(hi[2], hi[3], b3);
        if (R.ok) *(uint4 *)dst = nw;
    } else {
        if (!R.ok) return;
        bf16_pair_add_sr(dst, lo[0], lo[1], b0);
        bf16_pair_add_sr(dst + 2, lo[2], lo[3], b1);
        bf16_pair_add_sr(dst + 4, hi[0], hi[1], b2);
        bf16_pair_add_sr(dst + 6, hi[2], hi[3], b3);
    }
}

// WPE: waves per SIMD the register budget is cut for (the LDS of one
// workgroup limits the occupancy first at d = 128).
// DP: the row pitch when known at compile time (128 for d = 128, so every
// LDS offset is an immediate and the column loops unroll), 0: from a.dp.
// DBG: the test hook (products diverted to the debug buffers, tables left
// alone); the production kernels carry none of its code.
// TRK: lost-update tracking (a.row_ver), likewise compiled only into the
// kernels that run with it.
template <typename T, int RM, bool WB, int WPE, int DP, bool DBG, bool TRK>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(WPE))) void k_sgns_tile(Params a) {
    constexpr uint32_t GT = gt_pitch(sizeof(T), RM);  // pitch of the count / G^T tile (words)
    constexpr uint32_t EPV = 16 / sizeof(T);          // elements per 16-byte vector
    // gathered vectors held per thread: all of a 32 + 32-row bf16 tile at
    // d = 128 (4), else the first NB of the tile's vectors
    constexpr int NB = sizeof(T) == 4 ? 8 : ((DP == 128 && RM == 32) ? 4 : 5);
    extern __shared__ __align__(16) unsigned char lds[];
    const Lds L = lds_layout(DP ? (uint32_t)DP : a.dp, sizeof(T), RM, a.negative, TRK);
    const uint32_t ldv = L.ldv;
    T *Vs = (T *)(lds + L.o_rows);
    T *Us = Vs + RM * ldv;
    float *Gt = (float *)(lds + L.o_g);
    uint32_t *cnt = (uint32_t *)Gt;  // counts, overwritten in place by G^T
    uint32_t *wcnt = cnt + OMAX * GT;  // per-group context counts W[g][r]
    uint32_t *negmask = (uint32_t *)(lds + L.o_mask);
    uint32_t *verl = (uint32_t *)(lds + L.o_ver);  // staged row -> version at its gather

    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t l16 = lane & 15, g4 = 4 * (lane >> 4);
    const uint32_t w = a.window, k = a.negative, B = a.bgroup;
    const uint32_t dp = DP ? (uint32_t)DP : a.dp, d = a.d, dtiles = dp >> 4;
    T *syn0 = (T *)a.syn0;
    T *syn1 = (T *)a.syn1;
    unsigned long long my_pairs = 0, my_tiles = 0;
    const uint32_t round = a.round_state->round;
    const float alpha = a.round_state->alpha;

    // row gathers: 16-byte vectors per row, each thread's first one
    const uint32_t vpr = dp / EPV;
    const uint32_t st_r0 = tid / vpr, st_c0 = tid % vpr;
    const uint32_t st_dr = kBlock / vpr, st_dc = kBlock % vpr;
    // pair counts: 2^win_shift lanes per centre (64 >> win_shift windows
    // per wavefront), strided over its window
    const uint32_t wsh = a.win_shift, lpw = 1u << wsh, cs = tid & (lpw - 1), span = 2 * w + 1;
    // update products: items (row block, column part) dealt round-robin to
    // the four wavefronts, input row blocks first, then output row blocks
    const uint32_t parts = a.col_parts;
    // negative slot q of a tile belongs to its group q / k (q < 32: the
    // rounded float quotient is exact)
    const float inv_k = k ? 1.f / (float)k : 0.f;
    // write-back tiles start the products from the staged rows (acc = old +
    // delta); the test hook takes the deltas alone
    constexpr bool acc_old = WB && !DBG;
    // stochastic-rounding stream of this thread (bf16 tables), seeded per tile
    uint32_t rs = 0;
    // lost-update tracking: rows written back, and of those the ones whose
    // version moved since this tile gathered them (another tile wrote the
    // row in between, and this write-back replaces that update).  Lanes
    // 0-15 take the 16 rows of a row block, once per row (column part 0).
    uint64_t rows_wb = 0, rows_lost = 0;
    auto track_rows = [&](bool live, uint32_t vidx, uint32_t lrow) {
        const bool mine = live && lane < 16;
        bool lost = false;
        if (mine) lost = atomicAdd(&a.row_ver[vidx], 1u) != verl[lrow];
        rows_wb += (uint64_t)__popcll(__ballot(mine));
        rows_lost += (uint64_t)__popcll(__ballot(lost));
    };

    // bf16 products contract over whole 32-row blocks, reading staged rows
    // past a tile's own (rows of earlier tiles, finite); clear them once
    if constexpr (sizeof(T) == 2)
        for (uint32_t i = tid; i < (RM + OMAX) * ldv * 2 / 16; i += kBlock) ((uint4 *)Vs)[i] = make_uint4(0, 0, 0, 0);

    // the 48/64-row launch of a split finds nothing to do when the prep
    // counted no tile over 32 input rows
    if (a.nin_lo > 32 && a.counters[C_BIGTILES] == 0) return;
    uint64_t n_tiles = (uint64_t)a.round_state->n_active + a.counters[C_TILES];
    if (n_tiles > a.tiles_cap) n_tiles = a.tiles_cap;
    const uint4 *descs = (const uint4 *)a.tiles;
    const uint32_t *desc_words = (const uint32_t *)a.tiles;

    // Tiles come from a cursor shared by the launch: a workgroup takes
    // chunks of consecutive descriptor slots, so faster workgroups take
    // more tiles (a workgroup's speed depends on its placement).  Every wave
    // reads the nc and n_in words of the chunk's slots and ballots, so all
    // waves walk the same live tiles with no barrier spent on empty slots.
    // The chunk after the current one is requested ahead by thread 0 and
    // reaches every wave through LDS at the staging barrier; a chunk with
    // no tile left before that waits for it at a barrier of its own.  The
    // last workgroup to finish returns the cursor to zero for the next
    // launch.
    const uint32_t chunk = a.nin_lo > 32 ? 64u : kChunk32;
    unsigned long long *cursor = &a.counters[C_CURSOR];
    unsigned long long *grab_lds = (unsigned long long *)(lds + L.o_grab);
    uint64_t win = 0, win_nx = 0, live = 0;
    unsigned long long grab = 0;  // thread 0: the requested chunk
    bool pending = false;         // a request not yet seen by every wave
    // thread 0's request to every wave (a uniform call: two barriers)
    auto publish = [&]() -> uint64_t {
        if (tid == 0) *grab_lds = grab;
        __syncthreads();
        const uint64_t g = *grab_lds;
        __syncthreads();
        return g;
    };
    auto scan = [&]() {
        const uint64_t t = win + lane;
        bool ok = false;
        if (lane < chunk && t < n_tiles) {
            // both words loaded before either is tested (one round trip)
            const uint32_t *dw = desc_words + t * (kDescBytes / 4);
            const uint32_t nc = dw[2], nin = dw[6];
            ok = (nc != 0u) & (nin >= a.nin_lo) & (nin <= a.nin_hi);
        }
        live = __ballot(ok);
    };
    auto next_live = [&]() -> uint64_t {
        while (live == 0) {
            win = pending ? publish() : win_nx;
            pending = true;
            if (tid == 0) grab = atomicAdd(cursor, (unsigned long long)chunk);
            if (win >= n_tiles) return n_tiles;
            scan();
        }
        const uint64_t t = win + (uint64_t)__builtin_ctzll(live);
        live &= live - 1;
        return t;
    };
    if (tid == 0) grab = atomicAdd(cursor, (unsigned long long)chunk);
    win_nx = publish();

    // Software pipeline over the workgroup's tiles: tile i's descriptor is
    // in slot i & 1, its rows arrive in registers while tile i - 1 trains,
    // and the descriptor of tile i + 1 arrives with them.  The row gathers
    // of tile i + 1 thus read the tables before tile i writes them back
    // (Hogwild, as between concurrent workgroups).
    auto slot = [&](uint32_t s) -> TileDesc & { return *(TileDesc *)(lds + L.o_desc + s * kDescBytes); };
    // gathers: vector v = tid + kBlock*j of a tile covers row v / vpr and
    // 16-byte column v % vpr; (row, column) advance by fixed steps.  Staged
    // row s < R16 is input row s, s >= R16 output row s - R16 (LDS row
    // RM + s - R16, table row V + id in the joint allocation).
    auto advance = [&](uint32_t &r, uint32_t &cc) {
        r += st_dr;
        cc += st_dc;
        if (cc >= vpr) { cc -= vpr; ++r; }
    };
    // where this thread's vectors past its NB register ones start (the same
    // for every tile): larger tiles load those directly when staging
    uint32_t row_x = st_r0, c_x = st_c0;
#pragma unroll
    for (int j = 0; j < NB; ++j) advance(row_x, c_x);
    uint4 buf[NB];
    uint32_t my_ver = 0;  // version of the staged row this thread tracks (row_ver)
    // this thread's first NB vectors of the tile in slot s (padding rows 0):
    // slots and LDS reads of the row ids first, then the loads
    auto gather = [&](uint32_t s) {
        const TileDesc &t = slot(s);
        const uint32_t *tk = t.rin_tok;  // output row ids follow at RMAX
        const uint32_t n_in = t.n_in, n_out = t.n_out, R16 = (n_in + 15) & ~15u;
        uint32_t tix[NB], ce[NB], grow[NB];
        uint32_t row = st_r0, c = st_c0;
        // select-only index math (no short-circuit branches): rr wraps for
        // input rows, so rr < n_out holds exactly for live output rows
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            const uint32_t rr = row - R16;
            const bool live = (row < n_in) | (rr < n_out);
            tix[j] = live ? (row < R16 ? row : rr + (uint32_t)RMAX) : 0xFFFFFFFFu;
            ce[j] = c * EPV;
            advance(row, c);
        }
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            const bool live = tix[j] != 0xFFFFFFFFu;
            const uint32_t id = tk[live ? tix[j] : 0u];
            grow[j] = live ? id + (tix[j] >= (uint32_t)RMAX ? a.V : 0u) : 0u;
        }
        // lost-update tracking: thread s reads the version of staged row s
        if constexpr (TRK) {
            const uint32_t rr = tid - R16;
            my_ver = 0xFFFFFFFFu;
            if (tid < n_in) my_ver = a.row_ver[tk[tid]];
            else if (tid >= R16 && rr < n_out) my_ver = a.row_ver[a.V + tk[RMAX + rr]];
        }
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            buf[j] = make_uint4(0, 0, 0, 0);
            if (tix[j] != 0xFFFFFFFFu) buf[j] = *(const uint4 *)(syn0 + ((uint64_t)grow[j] * dp + ce[j]));
        }
    };

    uint64_t tile = next_live();
    if (tid < (uint32_t)kDescVec && tile < n_tiles) ((uint4 *)&slot(0))[tid] = descs[tile * kDescVec + tid];
    uint64_t nxt = tile < n_tiles ? next_live() : n_tiles, nxt2;
    uint4 pre = make_uint4(0, 0, 0, 0);  // the descriptor of tile nxt
    if (tid < (uint32_t)kDescVec && nxt < n_tiles) pre = descs[nxt * kDescVec + tid];
    __syncthreads();
    if (tile < n_tiles) gather(0);

    for (uint32_t s = 0; tile < n_tiles; tile = nxt, nxt = nxt2, s ^= 1) {
        TileDesc &td = slot(s);
        const uint32_t *toks = td.rin_tok;
        nxt2 = nxt < n_tiles ? next_live() : n_tiles;
        if constexpr (sizeof(T) == 2) {
            // a tile's rounding bits depend on (round, slot, thread) only,
            // not on the workgroup that takes it: one draw per tile (scalar
            // unit), hashed with the thread index
            const uint32_t ts = __builtin_amdgcn_readfirstlane((uint32_t)tile);
            const uint32_t base = draw(a.seed, 0, DOM_ROUND_BF16, ts, round).x;
            rs = mix32(base + tid * 0x9E3779B9u);
            if (rs == 0) rs = 0x9E3779B9u;
        }
        for (uint32_t i = tid; i < ((OMAX + L.wrows) * GT + OMAX) / 4; i += kBlock)
            ((uint4 *)cnt)[i] = make_uint4(0, 0, 0, 0);  // counts, W and the masks
        __syncthreads();

        const uint32_t c0 = td.c0, nc = td.nc, g0 = td.g0, lo = td.lo, npos = td.npos;
        const uint32_t n_in = td.n_in, n_out = td.n_out;
        const uint32_t R16 = (n_in + 15) & ~15u, O16 = (n_out + 15) & ~15u;
        const uint32_t nrows = R16 + O16;

        // ---- 1. pair counts (the rows are in flight): P in the low
        //         half-word, Q in the high ---------------------------------
        // Q[o][r] = sum over the negative slots q holding row o of W[q/k][r]
        // (pairs of q's group with context row r), less the pairs whose centre
        // row is o itself (word2vec skips a negative equal to the centre).
        // Per pair: P += 1 and the skipped negatives subtracted at [ro][r]
        // (one atomic), W += 1 (one atomic); Q is completed from W and the
        // slot masks when G is formed.
        for (uint32_t ci = tid >> wsh; ci < nc; ci += kBlock >> wsh) {
            const uint32_t b = td.bshrink[ci];
            const uint32_t ro = td.cen_rout[ci];
            const uint32_t grp = (c0 + ci) / B - g0;  // this centre's group within the tile
            const uint32_t coll = td.cen_coll[ci];     // its group's negatives equal to it
            const uint32_t pinc = 1u - (coll << 16);
            uint32_t *wrow = wcnt + grp * GT;
            for (uint32_t aa = cs; aa < span; aa += lpw) {
                if (aa == w || aa < b || aa >= span - b) continue;
                const int p = (int)(c0 + ci) - (int)w + (int)aa - (int)lo;  // input position in the tile
                if (p < 0 || p >= (int)npos) continue;
                const uint32_t rin = td.pos_rin[p];
                atomicAdd(&cnt[ro * GT + rin], pinc);
                if (k) atomicAdd(&wrow[rin], 1u);
                ++my_pairs;
            }
        }
        {
            const uint32_t nslots = ((c0 + nc - 1) / B - g0 + 1) * k;  // negative slots of the tile
            if (tid < nslots) atomicOr(&negmask[td.neg_rout[tid]], 1u << tid);
        }

        // ---- 2. gathered vectors to LDS, any further vectors directly; the
        //         next tile's descriptor to the other slot ------------------
        // One wait for every vector-memory operation in flight: the row
        // loads, the descriptor, and the previous tile's stores.
        __builtin_amdgcn_s_waitcnt(kWaitVm0);
        {
            uint32_t r = st_r0, cc = st_c0;
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                if (r < nrows) *(uint4 *)(Vs + (r < R16 ? r : r + (RM - R16)) * ldv + cc * EPV) = buf[j];
                advance(r, cc);
            }
        }
        for (uint32_t row = row_x, c = c_x; row < nrows; advance(row, c)) {
            const bool in = row < R16;
            const uint32_t r = in ? row : row - R16;
            uint4 val = make_uint4(0, 0, 0, 0);
            if (r < (in ? n_in : n_out))
                val = *(const uint4 *)((in ? syn0 : syn1) + ((uint64_t)toks[in ? r : RMAX + r] * dp + c * EPV));
            *(uint4 *)(Vs + (in ? row : RM + r) * ldv + c * EPV) = val;
        }
        if (TRK && tid < nrows) verl[tid < R16 ? tid : RM + tid - R16] = my_ver;
        if (tid < (uint32_t)kDescVec) ((uint4 *)&slot(s ^ 1))[tid] = pre;
        if (pending && tid == 0) *grab_lds = grab;
        __syncthreads();
        if (pending) {
            win_nx = *grab_lds;
            pending = false;
        }

        // ---- 3. the next tile's rows and the descriptor after it, in flight
        //         while this tile trains ------------------------------------
        if (nxt < n_tiles) gather(s ^ 1);
        pre = make_uint4(0, 0, 0, 0);
        if (tid < (uint32_t)kDescVec && nxt2 < n_tiles) pre = descs[nxt2 * kDescVec + tid];

        const bool dbg_on = DBG && tile < a.dbg_max;
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

        // ---- 4. X = Vin * Uout^T; G^T = alpha (P - (P+Q) sigma(X))^T in place ----
        {
            const uint32_t two = O16 > 16;
            const uint32_t nX = (R16 >> 4) << two;
            for (uint32_t t = wave; t < nX; t += 4) {
                const uint32_t rt = t >> two, ot = t & two;
                f4 acc = {0.f, 0.f, 0.f, 0.f};
                if constexpr (sizeof(T) == 4) {
                    // dp is a multiple of 32: two 16-deep blocks per step, both
                    // blocks' operands read before their MFMAs
                    const float *va = (const float *)Vs + (rt * 16 + l16) * ldv + g4;
                    const float *ub = (const float *)Us + (ot * 16 + l16) * ldv + g4;
                    for (uint32_t kb = 0; kb < dp; kb += 32) {
                        const f4 x0 = *(const f4 *)(va + kb), y0 = *(const f4 *)(ub + kb);
                        const f4 x1 = *(const f4 *)(va + kb + 16), y1 = *(const f4 *)(ub + kb + 16);
#pragma unroll
                        for (int i = 0; i < 4; ++i) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(x0[i], y0[i], acc, 0, 0, 0);
#pragma unroll
                        for (int i = 0; i < 4; ++i) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(x1[i], y1[i], acc, 0, 0, 0);
                    }
                } else {
                    const uint16_t *va = (const uint16_t *)Vs + (rt * 16 + l16) * ldv + 2 * g4;
                    const uint16_t *ub = (const uint16_t *)Us + (ot * 16 + l16) * ldv + 2 * g4;
                    bf8 x = *(const bf8 *)va, y = *(const bf8 *)ub;
                    for (uint32_t kb = 32; kb < dp; kb += 32) {
                        const bf8 xn = *(const bf8 *)(va + kb), yn = *(const bf8 *)(ub + kb);
                        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(x, y, acc, 0, 0, 0);
                        x = xn;
                        y = yn;
                    }
                    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(x, y, acc, 0, 0, 0);
                }
                const uint32_t o = ot * 16 + l16, r0 = rt * 16 + g4;
                uint32_t *cp = cnt + o * GT + r0;
                uint4 cc = *(const uint4 *)cp;
                for (uint32_t m = negmask[o]; m; m &= m - 1) {  // negative slots holding row o
                    const uint32_t g = (uint32_t)(((float)__builtin_ctz(m) + 0.5f) * inv_k);
                    const uint4 wv = *(const uint4 *)(wcnt + g * GT + r0);
                    cc.x += wv.x << 16;
                    cc.y += wv.y << 16;
                    cc.z += wv.z << 16;
                    cc.w += wv.w << 16;
                }
                f4 gv;
                gv[0] = pair_grad(cc.x, acc[0], alpha);
                gv[1] = pair_grad(cc.y, acc[1], alpha);
                gv[2] = pair_grad(cc.z, acc[2], alpha);
                gv[3] = pair_grad(cc.w, acc[3], alpha);
                if constexpr (sizeof(T) == 4) {
                    *(f4 *)cp = gv;
                } else {
                    // G = hi + lo in bf16; each 4-row block of G^T holds its
                    // four hi values, then its four lo values
                    const uint32_t h01 = f32x2_to_bf16x2_rne(gv[0], gv[1]);
                    const uint32_t h23 = f32x2_to_bf16x2_rne(gv[2], gv[3]);
                    const uint32_t l01 = f32x2_to_bf16x2_rne(gv[0] - __uint_as_float(h01 << 16),
                                                             gv[1] - __uint_as_float(h01 & 0xFFFF0000u));
                    const uint32_t l23 = f32x2_to_bf16x2_rne(gv[2] - __uint_as_float(h23 << 16),
                                                             gv[3] - __uint_as_float(h23 & 0xFFFF0000u));
                    *(uint4 *)cp = make_uint4(h01, h23, l01, l23);
                }
                if (dbg_on)
                    for (int j = 0; j < 4; ++j) a.dbg_logits[(tile * RMAX + r0 + j) * OMAX + o] = acc[j];
            }
        }
        __syncthreads();

        if constexpr (sizeof(T) == 2) {
            // ---- 5 on bf16 MFMA (16x16x32): G = hi + lo, two products each ----
            // Operands are swapped (first = table rows, second = G) so the
            // products come out transposed: lane = one row.  Columns go in
            // pairs of 16-column blocks: the transposing loads point chunk q
            // of block h of pair p at columns 32p + 8q + 4h, so after both
            // blocks lane group g holds columns 32p + 8g .. 32p + 8g + 7 of
            // its row, stored as one 16-byte vector (64 contiguous bytes per
            // row and store instruction instead of 32).
            lds_u8 *l3 = (lds_u8 *)lds;
            const uint32_t grp = lane >> 4, jq = l16 >> 2, qq = l16 & 3;
            const uint32_t gb = (uint32_t)L.o_g, rb = (uint32_t)L.o_rows, rowb = ldv * 2;
            const uint32_t npairs = dtiles >> 1, cq = 16 * qq, cg = 8 * grp;
            // dVin = G * Uout, K = all OMAX output rows (zero G past n_out)
            const uint32_t ndv = (R16 >> 4) * parts;
            for (uint32_t it = wave; it < ndv; it += 4) {
                const uint32_t rt = it / parts, part = it - rt * parts;
                const uint32_t r = rt * 16 + l16;
                const LaneRow<T> row = lane_row(syn0, td.rin_tok, r, n_in, dp);
                // G[r'][o], r' = 16*rt + lane%16, o = 8*grp + s: transposed from G^T
                const uint32_t a0 = gb + (8 * grp + jq) * GT * 4 + (rt * 4 + qq) * 16;
                const bf8 ghi = join8(tr_load(l3, a0), tr_load(l3, a0 + 16 * GT));
                const bf8 glo = join8(tr_load(l3, a0 + 8), tr_load(l3, a0 + 16 * GT + 8));
                for (uint32_t pp = part; pp < npairs; pp += parts) {
                    const uint32_t c0 = 32 * pp + cg;
                    // Uout[o][32p + 8q + 4h + i], transposed from the staged rows
                    const uint32_t b0 = rb + (RM + 8 * grp + jq) * rowb + 64 * pp + cq;
                    const bf8 u0 = join8(tr_load(l3, b0), tr_load(l3, b0 + 4 * rowb));
                    const bf8 u1 = join8(tr_load(l3, b0 + 8), tr_load(l3, b0 + 4 * rowb + 8));
                    f4 x0 = {0.f, 0.f, 0.f, 0.f}, x1 = {0.f, 0.f, 0.f, 0.f};
                    if constexpr (acc_old) staged_bf16x8((const uint16_t *)Vs + r * ldv, c0, x0, x1);
                    x0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(u0, ghi, x0, 0, 0, 0);
                    x1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(u1, ghi, x1, 0, 0, 0);
                    x0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(u0, glo, x0, 0, 0, 0);
                    x1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(u1, glo, x1, 0, 0, 0);
                    if constexpr (DBG) {
                        if (dbg_on)
                            for (int j = 0; j < 4; ++j) {
                                float *dv = a.dbg_dv + (tile * RMAX + r) * d + c0;
                                if (c0 + j < d) dv[j] = x0[j];
                                if (c0 + 4 + j < d) dv[4 + j] = x1[j];
                            }
                    } else {
                        store_row8<WB>(row, c0, x0, x1, rs);
                    }
                }
                if (TRK && part == 0) track_rows(r < n_in, td.rin_tok[r < n_in ? r : 0], r);
            }
            // dUout = G^T * Vin, K = R16 rounded up to 32 (zero G past n_in;
            // input rows past RM are output rows, finite)
            const uint32_t nk32 = (R16 + 31) >> 5;
            const uint32_t ndu = (O16 >> 4) * parts;
            for (uint32_t it = (wave + 4 - (ndv & 3)) & 3; it < ndu; it += 4) {
                const uint32_t ot = it / parts, part = it - ot * parts;
                const uint32_t o = ot * 16 + l16;
                const LaneRow<T> row = lane_row(syn1, td.rout_tok, o, n_out, dp);
                for (uint32_t pp = part; pp < npairs; pp += parts) {
                    const uint32_t c0 = 32 * pp + cg;
                    f4 x0 = {0.f, 0.f, 0.f, 0.f}, x1 = {0.f, 0.f, 0.f, 0.f};
                    if constexpr (acc_old) staged_bf16x8((const uint16_t *)Us + o * ldv, c0, x0, x1);
#pragma unroll
                    for (int kb = 0; kb < 2; ++kb) {
                        if ((uint32_t)kb >= nk32) break;
                        // G^T[o][r], r = 32*kb + 8*grp + s: two 4-row blocks
                        // (read per column pair: fewer live registers)
                        const uint4 *gp = (const uint4 *)(lds + gb + (o * GT + 32 * kb + 8 * grp) * 4);
                        const uint4 x = gp[0], y = gp[1];
                        const bf8 ghi = __builtin_bit_cast(bf8, u4v{x.x, x.y, y.x, y.y});
                        const bf8 glo = __builtin_bit_cast(bf8, u4v{x.z, x.w, y.z, y.w});
                        // Vin[r][32p + 8q + 4h + i], transposed from the staged rows
                        const uint32_t b0 = rb + (32 * kb + 8 * grp + jq) * rowb + 64 * pp + cq;
                        const bf8 v0 = join8(tr_load(l3, b0), tr_load(l3, b0 + 4 * rowb));
                        const bf8 v1 = join8(tr_load(l3, b0 + 8), tr_load(l3, b0 + 4 * rowb + 8));
                        x0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(v0, ghi, x0, 0, 0, 0);
                        x1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(v1, ghi, x1, 0, 0, 0);
                        x0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(v0, glo, x0, 0, 0, 0);
                        x1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(v1, glo, x1, 0, 0, 0);
                    }
                    if constexpr (DBG) {
                        if (dbg_on)
                            for (int j = 0; j < 4; ++j) {
                                float *du = a.dbg_du + (tile * OMAX + o) * d + c0;
                                if (c0 + j < d) du[j] = x0[j];
                                if (c0 + 4 + j < d) du[4 + j] = x1[j];
                            }
                    } else {
                        store_row8<WB>(row, c0, x0, x1, rs);
                    }
                }
                if (TRK && part == 0)
                    track_rows(o < n_out, a.V + td.rout_tok[o < n_out ? o : 0], RM + o);
            }
            ++my_tiles;
            __syncthreads();  // the next tile reuses the LDS
            continue;
        }

        // ---- 5 on f32 MFMA (16x16x4), operands swapped as above -------------
        // dVin = G * Uout  (R16 x dp, K = O16: KQ blocks of 16 output rows)
        const uint32_t ndv = (R16 >> 4) * parts;
        auto dv_rows = [&](auto kq) {
            constexpr int KQ = decltype(kq)::value;
            for (uint32_t it = wave; it < ndv; it += 4) {
                const uint32_t rt = it / parts, part = it - rt * parts;
                const uint32_t r = rt * 16 + l16;
                const LaneRow<T> row = lane_row(syn0, td.rin_tok, r, n_in, dp);
                // G[r'][k], k = 16*q + 4*(lane/16) + i: one column of G^T per
                // k, the same for every column block of this row block
                float gx[4 * KQ];
                const float *ga = Gt + g4 * GT + rt * 16 + l16;
#pragma unroll
                for (int q = 0; q < KQ; ++q)
#pragma unroll
                    for (int i = 0; i < 4; ++i) gx[4 * q + i] = ga[(16 * q + i) * GT];
                for (uint32_t dt = part; dt < dtiles; dt += parts) {
                    const uint32_t col0 = dt * 16 + g4;
                    const T *ub = Us + g4 * ldv + dt * 16 + l16;
                    float ux[4 * KQ];
#pragma unroll
                    for (int q = 0; q < KQ; ++q)
#pragma unroll
                        for (int i = 0; i < 4; ++i) ux[4 * q + i] = load_elem(ub + (16 * q + i) * ldv);
                    f4 acc = (acc_old && sizeof(T) == 4) ? *(const f4 *)(Vs + r * ldv + col0) : f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                    for (int i = 0; i < 4 * KQ; ++i) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(ux[i], gx[i], acc, 0, 0, 0);
                    if constexpr (DBG) {
                        if (dbg_on)
                            for (int j = 0; j < 4; ++j)
                                if (col0 + j < d) a.dbg_dv[(tile * RMAX + r) * d + col0 + j] = acc[j];
                    } else {
                        store_row<T, WB>(row, col0, acc);
                    }
                }
                if (TRK && part == 0) track_rows(r < n_in, td.rin_tok[r < n_in ? r : 0], r);
            }
        };
        if (O16 > 16) dv_rows(std::integral_constant<int, 2>{});
        else dv_rows(std::integral_constant<int, 1>{});
        // dUout = G^T * Vin  (O16 x dp, K = R16: KQ blocks of 16 input rows)
        auto du_rows = [&](auto kq) {
            constexpr int KQ = decltype(kq)::value;
            const uint32_t ndu = (O16 >> 4) * parts;
            for (uint32_t it = (wave + 4 - (ndv & 3)) & 3; it < ndu; it += 4) {
                const uint32_t ot = it / parts, part = it - ot * parts;
                const uint32_t o = ot * 16 + l16;
                const LaneRow<T> row = lane_row(syn1, td.rout_tok, o, n_out, dp);
                // G^T[o][k]: four consecutive words per 16-deep block, the
                // same for every column block of this row block
                f4 gx[KQ];
                const float *ga = Gt + o * GT + g4;
#pragma unroll
                for (int q = 0; q < KQ; ++q) gx[q] = *(const f4 *)(ga + 16 * q);
                for (uint32_t dt = part; dt < dtiles; dt += parts) {
                    const uint32_t col0 = dt * 16 + g4;
                    const T *vb = Vs + g4 * ldv + dt * 16 + l16;
                    float vx[4 * KQ];
#pragma unroll
                    for (int q = 0; q < KQ; ++q)
#pragma unroll
                        for (int i = 0; i < 4; ++i) vx[4 * q + i] = load_elem(vb + (16 * q + i) * ldv);
                    f4 acc = (acc_old && sizeof(T) == 4) ? *(const f4 *)(Us + o * ldv + col0) : f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                    for (int q = 0; q < KQ; ++q)
#pragma unroll
                        for (int i = 0; i < 4; ++i)
                            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(vx[4 * q + i], gx[q][i], acc, 0, 0, 0);
                    if constexpr (DBG) {
                        if (dbg_on)
                            for (int j = 0; j < 4; ++j)
                                if (col0 + j < d) a.dbg_du[(tile * OMAX + o) * d + col0 + j] = acc[j];
                    } else {
                        store_row<T, WB>(row, col0, acc);
                    }
                }
                if (TRK && part == 0)
                    track_rows(o < n_out, a.V + td.rout_tok[o < n_out ? o : 0], RM + o);
            }
        };
        switch (R16 >> 4) {
            case 1: du_rows(std::integral_constant<int, 1>{}); break;
            case 2: du_rows(std::integral_constant<int, 2>{}); break;
            case 3: du_rows(std::integral_constant<int, 3>{}); break;
            default:
                if constexpr (RM >= 64) du_rows(std::integral_constant<int, 4>{});
                break;
        }
        ++my_tiles;
        __syncthreads();  // the next tile reuses the LDS
    }
    // the pair count goes out once per workgroup (atomics on one address
    // serialise), combined in the descriptor slots, which no tile reads now
    for (int o = 32; o > 0; o >>= 1) my_pairs += __shfl_xor(my_pairs, o);
    unsigned long long *part = (unsigned long long *)(lds + L.o_desc);
    if (lane == 0) part[wave] = my_pairs;
    __syncthreads();
    if (tid == 0) {
        for (uint32_t i = 1; i < kBlock / 64; ++i) my_pairs += part[i];
        atomicAdd(&a.counters[C_PAIRS], my_pairs);
        atomicAdd(&a.counters[C_GROUPS], my_tiles);
        // every request of this workgroup is complete before it counts as done
        __threadfence();
        if (atomicAdd(&a.counters[C_DONE], 1ull) == gridDim.x - 1ull) {
            atomicExch(cursor, 0ull);
            atomicExch(&a.counters[C_DONE], 0ull);
        }
    }
    if (TRK && lane == 0) {
        atomicAdd(&a.counters[C_ROWS_WB], (unsigned long long)rows_wb);
        atomicAdd(&a.counters[C_ROWS_LOST], (unsigned long long)rows_lost);
    }
}

template <typename T> constexpr int kWpeLow = sizeof(T) == 4 ? 2 : 3;

template <typename T, int RM, bool WB, int WPE, int DP, bool DBG = false, bool TRK = false>
hipError_t occupancy_dp(felog_ctx *c, int *per_cu) {
    const Lds L = lds_layout(c->dp, sizeof(T), RM, c->opt.negative, TRK);
    hipError_t e = hipFuncSetAttribute((const void *)k_sgns_tile<T, RM, WB, WPE, DP, DBG, TRK>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)L.total);
    if (e != hipSuccess) return e;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, k_sgns_tile<T, RM, WB, WPE, DP, DBG, TRK>, kBlock,
                                                     L.total);
    if (*per_cu < 1) *per_cu = 1;
    return e;
}

template <typename T, int RM, bool WB, int WPE>
hipError_t occupancy(felog_ctx *c, int *per_cu) {
    return c->dp == 128 ? occupancy_dp<T, RM, WB, WPE, 128>(c, per_cu) : occupancy_dp<T, RM, WB, WPE, 0>(c, per_cu);
}

// Persistent grid of gpc workgroups per CU (0 or more than fit: as many as
// fit); *blocks receives the grid.
template <typename T, int RM, bool WB, int WPE, int DP, bool DBG, bool TRK = false>
hipError_t launch_dp(felog_ctx *c, const Params &a, uint32_t gpc, unsigned *blocks) {
    int per_cu = 0;
    hipError_t e = occupancy_dp<T, RM, WB, WPE, DP, DBG, TRK>(c, &per_cu);
    if (e != hipSuccess) return e;
    if (gpc && (int)gpc < per_cu) per_cu = (int)gpc;
    *blocks = (unsigned)per_cu * c->cu_count;
    const Lds L = lds_layout(c->dp, sizeof(T), RM, c->opt.negative, TRK);
    k_sgns_tile<T, RM, WB, WPE, DP, DBG, TRK><<<*blocks, kBlock, L.total, c->stream>>>(a);
    return hipGetLastError();
}

template <typename T, int RM, bool WB, int WPE>
hipError_t launch(felog_ctx *c, const Params &a, uint32_t gpc, unsigned *blocks) {
    if (a.dbg_max)  // the test hook: one register budget per table type
        return c->dp == 128 ? launch_dp<T, RM, WB, kWpeLow<T> + 1, 128, true>(c, a, gpc, blocks)
                            : launch_dp<T, RM, WB, kWpeLow<T> + 1, 0, true>(c, a, gpc, blocks);
    if constexpr (WB) {
        if (a.row_ver)  // lost-update tracking (write-back only)
            return c->dp == 128 ? launch_dp<T, RM, WB, WPE, 128, false, true>(c, a, gpc, blocks)
                                : launch_dp<T, RM, WB, WPE, 0, false, true>(c, a, gpc, blocks);
    }
    return c->dp == 128 ? launch_dp<T, RM, WB, WPE, 128, false>(c, a, gpc, blocks)
                        : launch_dp<T, RM, WB, WPE, 0, false>(c, a, gpc, blocks);
}

template <typename T, int RM, bool WB>
hipError_t launch_wpe(felog_ctx *c, const Params &a, uint32_t gpc, unsigned *blocks) {
    constexpr int lo = kWpeLow<T>;
    return c->tile_wpe == (uint32_t)lo ? launch<T, RM, WB, lo>(c, a, gpc, blocks)
                                       : launch<T, RM, WB, lo + 1>(c, a, gpc, blocks);
}

template <typename T, int RM>
hipError_t launch_rm(felog_ctx *c, const Params &a, uint32_t gpc, unsigned *blocks) {
    return a.writeback ? launch_wpe<T, RM, true>(c, a, gpc, blocks) : launch_wpe<T, RM, false>(c, a, gpc, blocks);
}

// The 48/64-row kernel (a.rmax rows), and the 32-row one.
template <typename T>
hipError_t launch_big(felog_ctx *c, const Params &a, uint32_t gpc, unsigned *blocks) {
    return a.rmax == 48 ? launch_rm<T, 48>(c, a, gpc, blocks) : launch_rm<T, 64>(c, a, gpc, blocks);
}

template <typename T>
hipError_t launch_split(felog_ctx *c, Params a) {
    // tiles of at most 32 input rows on the 32-row kernel (a smaller LDS
    // footprint, more workgroups per CU), the rest on the 48/64-row one
    unsigned blocks = 0;
    a.nin_lo = 0;
    a.nin_hi = 32;
    hipError_t e = launch_rm<T, 32>(c, a, c->tile_gpc, &blocks);
    if (e != hipSuccess) return e;
    a.nin_lo = 33;
    a.nin_hi = RMAX;
    return launch_big<T>(c, a, 0, &blocks);
}

template <typename T>
hipError_t occupancy_t(felog_ctx *c, uint32_t rows, uint32_t wpe, int *per_cu) {
    constexpr int lo = kWpeLow<T>;
    const bool wb = c->opt.update_mode != FELOG_UPDATE_ATOMIC;
    if (rows == 32) {
        if (wpe == (uint32_t)lo)
            return wb ? occupancy<T, 32, true, lo>(c, per_cu) : occupancy<T, 32, false, lo>(c, per_cu);
        return wb ? occupancy<T, 32, true, lo + 1>(c, per_cu) : occupancy<T, 32, false, lo + 1>(c, per_cu);
    }
    if (rows == 48) {
        if (wpe == (uint32_t)lo)
            return wb ? occupancy<T, 48, true, lo>(c, per_cu) : occupancy<T, 48, false, lo>(c, per_cu);
        return wb ? occupancy<T, 48, true, lo + 1>(c, per_cu) : occupancy<T, 48, false, lo + 1>(c, per_cu);
    }
    if (wpe == (uint32_t)lo)
        return wb ? occupancy<T, 64, true, lo>(c, per_cu) : occupancy<T, 64, false, lo>(c, per_cu);
    return wb ? occupancy<T, 64, true, lo + 1>(c, per_cu) : occupancy<T, 64, false, lo + 1>(c, per_cu);
}

}  // namespace

hipError_t launch_tile(felog_ctx *c, const Params &a0) {
    const bool bf = c->opt.precision == FELOG_BF16;
    // default: the split (tiles of walks up to 32 tokens on the 32-row kernel)
    if (c->tile_rows == 0 || c->tile_rows == 32)
        return bf ? launch_split<uint16_t>(c, a0) : launch_split<float>(c, a0);
    Params a = a0;
    a.nin_lo = 0;
    a.nin_hi = RMAX;
    unsigned blocks = 0;
    return bf ? launch_big<uint16_t>(c, a, c->tile_gpc, &blocks) : launch_big<float>(c, a, c->tile_gpc, &blocks);
}

}  // namespace sgns

uint32_t tile_wpe_low(const felog_ctx *c) { return c->opt.precision == FELOG_BF16 ? 3u : 2u; }

hipError_t tile_occupancy(felog_ctx *c, uint32_t rows, uint32_t wpe, int *per_cu) {
    if (c->opt.precision == FELOG_BF16) return sgns::occupancy_t<uint16_t>(c, rows, wpe, per_cu);
    return sgns::occupancy_t<float>(c, rows, wpe, per_cu);
}

}  // namespace felog
