// sgns_common.h - what the two SGNS kernels share: tile geometry, the tile
// descriptor the prep kernel writes and the tile kernel reads, and the
// kernel parameters.
//
// The kept tokens of a sentence are cut into groups of B consecutive
// centres; each group draws k negatives that its centres share (HogBatch,
// include/dsgl.hpp:1054-1159).  Consecutive groups are packed into one tile
// while the tile has at most 32 output slots (centres + k per group) and its
// windows span at most RMAX positions (48 or 64); a 22-token walk with
// B = 21, k = 5 is one tile.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "felog_internal.h"

namespace felog {
namespace sgns {

constexpr int kBlock = 256;
constexpr int RMAX = 64;        // input positions of a tile, upper limit
constexpr int OMAX = 32;        // output slots of a tile
constexpr int GLD = OMAX + 4;   // row pitch of the count / G tile (words)
constexpr int MAXKEEP = 1000;   // MAX_SENTENCE_LENGTH, src/word2vec.cu:42
constexpr float kMaxExp = 6.f;  // MAX_EXP, src/word2vec.cu:40

// Tile descriptor.  Slot s < n_sent (the round's sentences) holds the first tile of sentence s
// (nc = 0 when the sentence keeps fewer than two tokens); further tiles of
// long sentences take slots from n_sent on.
struct TileDesc {
    uint32_t sid, c0, nc, g0, lo, npos, n_in, n_out;
    uint32_t pad0[8];
    uint32_t rin_tok[RMAX];   // distinct input rows (vertex ids)
    uint32_t rout_tok[OMAX];  // distinct output rows
    uint8_t pos_rin[RMAX];    // input position lo+i -> input row
    uint8_t cen_rout[OMAX];   // centre c0+i -> output row
    uint8_t neg_rout[OMAX];   // negative q of the tile -> output row
    uint8_t bshrink[OMAX];    // window shrink b of centre c0+i
    uint8_t cen_coll[OMAX];   // negatives of centre c0+i's group equal to it
};
static_assert(sizeof(TileDesc) == 640, "descriptor layout");
constexpr int kDescBytes = 640;
constexpr int kDescVec = kDescBytes / 16;

struct Params {
    const uint64_t *__restrict__ off;
    const uint32_t *__restrict__ tok;
    // round, learning rate and sentence count (n_active) of the pass
    const RoundState *__restrict__ round_state;
    void *syn0;
    void *syn1;
    uint32_t dp, d;
    const uint8_t *__restrict__ in_vocab;
    const uint32_t *__restrict__ vocab_rank;
    const uint32_t *__restrict__ keep_thr;
    const uint2 *__restrict__ uni;  // {threshold, alias} per vocabulary bucket
    uint32_t V;
    uint32_t window, negative, bgroup, rmax;
    int compat_break;
    int writeback;  // 1: store staged row + delta; 0: atomic updates
    uint64_t seed;
    unsigned char *tiles;
    uint64_t tiles_cap;
    // tiles of this launch: n_in in [nin_lo, nin_hi] (a 32-row kernel takes
    // the small tiles, a 48/64-row one the rest)
    uint32_t nin_lo, nin_hi;
    // lost-update tracking (null: off): one version per row of the joint
    // table allocation (syn0 rows, then syn1neg rows), read when a tile
    // gathers the row and incremented when it writes the row back
    uint32_t *row_ver;
    // tile shape: log2 of the lanes that count one centre's window (3: eight
    // windows per wavefront), and the column parts of a row block in the
    // update products (a wavefront's item is one row block's columns
    // part, part + parts, ...: dp / parts columns)
    uint32_t win_shift, col_parts;
    // the longest sentence of the corpus when known on the device (the walk
    // round's maximum length for the library's own corpus), else null
    const unsigned long long *max_sent_len;
    unsigned long long *counters;
    // test hook: when dbg_max > 0 the tables are not touched
    uint32_t dbg_max;
    uint32_t *dbg_groups, *dbg_rows;
    float *dbg_logits, *dbg_dv, *dbg_du;
};

// Number of consecutive groups, starting at group g0, that form one tile.
__device__ __forceinline__ uint32_t tile_groups(uint32_t g0, uint32_t m, uint32_t B, uint32_t w, uint32_t k,
                                                uint32_t rmax) {
    const uint32_t c_start = g0 * B;
    const uint32_t lo = c_start >= w ? c_start - w : 0;
    uint32_t ng = 1;
    while (true) {
        const uint32_t g = g0 + ng;
        if (g * B >= m) break;
        const uint32_t ce = (g + 1) * B < m ? (g + 1) * B : m;
        const uint32_t hi = ce + w < m ? ce + w : m;
        if ((ce - c_start) + k * (ng + 1) > (uint32_t)OMAX || hi - lo > rmax) break;
        ++ng;
    }
    return ng;
}

// Launch the tile kernel for table element type (f32 or bf16 bits) and
// tile rows; defined in sgns_tile.hip.
hipError_t launch_tile(felog_ctx *c, const Params &a);

}  // namespace sgns
}  // namespace felog
