// felog_internal.h - the context behind the C ABI and the host-side launch
// functions of each kernel family.  Not part of the public interface.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>
#include "../../include/felog.h"

namespace felog {

// Device record of one CSR vertex: first CSR entry and degree.
struct VertexRec { uint32_t begin, deg; };
// In a walk's vertex records the degree carries this bit when the degree is
// 2 and the first neighbour is the vertex itself (such a vertex ends a walk,
// include/walk.hpp:977-979).
constexpr uint32_t kSelfFirst = 0x80000000u;
// Device record of one CSR entry j of vertex v: the alias threshold of
// column j, the neighbour in column j, the neighbour in column alias(j),
// alias(j) itself, and both neighbours' walk vertex records.  One 32-byte
// gather (one cache line) decides a walk step and hands the next step its
// vertex.
struct EdgeRec {
    uint32_t thresh, col, alias_col, alias_idx;
    VertexRec col_vr, alias_vr;
};
static_assert(sizeof(EdgeRec) == 32, "one 32-byte gather per walk step");

// Rows longer than kHubChunk CSR entries ("hubs") are processed in chunks of
// kHubChunk entries by one wavefront each; hub k (vertex hub[k]) owns chunks
// cstart[k] .. cstart[k + 1] - 1 of the plan.
constexpr uint32_t kHubChunk = 4096;
struct HubPlan {
    uint32_t *hub = nullptr;     // n
    uint64_t *cstart = nullptr;  // n + 1 (device)
    uint32_t n = 0;
    uint64_t chunks = 0;
};

struct Corpus {
    uint32_t round = 0;
    uint64_t n_walks = 0, n_tokens = 0, n_steps = 0;
    uint32_t max_len = 0;
    bool valid = false;
    bool from_walk = false;
};

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

// Per-round values in device memory.  Every kernel of a round reads them
// there, so a captured round (hipGraph) replays unchanged: the round-start
// kernel advances the round and its learning rate, and the evaluation
// writes the stop decision back.
struct RoundState {
    uint32_t round;        // 1-based round being run
    float alpha;           // learning rate of the round
    uint32_t n_active;     // sentences of the round: walkers started (written by the walker select)
    uint32_t stop;         // 1 once the plain rule's stop ratio fired
    uint32_t last_eva;     // unconverged count of the previous evaluation
    uint32_t eval_from;    // first round that evaluates (min_R)
    uint32_t budget;       // 1: rounds follow a learned walk budget
    uint32_t evaluated;    // 1: this round ran the evaluation
    uint64_t n_tokens;     // corpus tokens of the round
};

// Whether a captured round evaluates: from round min_R on, while the plain
// rule has not stopped, and (under a budget) only after a round that walked
// (src/word2vec.cu:2016-2080; the budget loop stops before training when
// no walker starts).
__host__ __device__ inline bool round_evaluates(const RoundState &s) {
    return s.round >= s.eval_from && !s.stop && !(s.budget && s.n_active == 0);
}

}  // namespace felog

struct felog_ctx {
    felog_options opt{};
    std::string err;
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr;
    uint32_t cu_count = 0;

    // sizes
    uint32_t V = 0;
    uint64_t E = 0;         // directed CSR entries
    uint32_t d = 0;         // -size
    uint32_t dp = 0;        // row pitch (d rounded up to 32)
    uint64_t vocab_size = 0, train_words = 0;
    uint32_t slot = 0;      // per-walker fixed walk slot (vertices)

    // graph input (directed edge list before the CSR is built; the CSR's
    // columns and weights take over d_edst's and d_ew's blocks)
    uint32_t *d_esrc = nullptr, *d_edst = nullptr;
    float *d_ew = nullptr;
    // whole device blocks the build has finished with, kept for the round
    // buffers and tile descriptors that follow instead of being freed and
    // allocated again: {base, bytes}; freed with the context
    std::vector<std::pair<void *, size_t>> spare;
    bool have_edges = false, built = false;

    // CSR + alias
    uint64_t *d_row_ptr = nullptr;       // V+1
    uint32_t *d_col = nullptr;           // E
    float *d_w = nullptr;                // E
    felog::VertexRec *d_vrec = nullptr;  // V
    felog::VertexRec *d_vwalk = nullptr; // V: vrec with the kSelfFirst bit
    felog::EdgeRec *d_erec = nullptr;    // E
    uint32_t *d_deg = nullptr;           // V
    felog::HubPlan hubs;                 // rows of more than kHubChunk entries

    // vocabulary / sampling tables
    uint32_t *d_vocab_order = nullptr;   // V, rank -> vertex
    uint32_t *d_vocab_rank = nullptr;    // V, vertex -> rank
    uint8_t *d_in_vocab = nullptr;       // V
    uint2 *d_uni = nullptr;              // V: {threshold, alias} of the unigram^0.75 sampler
    uint32_t *d_keep_thr = nullptr;      // V: subsampling keep threshold, 0 outside the vocabulary
    double *d_T = nullptr, *d_dT = nullptr;  // n log2 n table and its differences
    std::vector<double> h_T;             // their host copy (T, then dT)

    // embedding tables (f32 or bf16 rows of dp elements, indexed by vertex
    // id); one allocation owned by d_syn0, d_syn1 = d_syn0 + V rows
    void *d_syn0 = nullptr, *d_syn1 = nullptr;

    // convergence state
    uint8_t *d_flags = nullptr;          // V
    float *d_scores = nullptr;           // V
    uint32_t last_eva_num = 0;
    uint32_t n_unconverged = 0;
    uint32_t rounds_done = 0;
    bool stop = false;

    // walk buffers
    uint32_t walkers = 0;
    uint32_t *d_active = nullptr;        // walker ids of this round
    uint32_t *d_sel = nullptr;           // walkers: inclusive sum of the start flags
    uint32_t *d_slots = nullptr;         // walkers * slot
    uint32_t *d_lens = nullptr;          // walkers
    uint64_t *d_ovf_base = nullptr;      // walkers
    uint32_t *d_ovf = nullptr;           // overflow pool
    uint64_t ovf_cap = 0;
    uint64_t *d_counters = nullptr;      // small device counters
    uint64_t *h_counters = nullptr;      // pinned mirror
    // flat corpus
    uint64_t *d_off = nullptr;           // n_walks + 1
    uint32_t *d_tok = nullptr;
    uint32_t *d_sent_id = nullptr;       // walker id per sentence
    uint64_t tok_cap = 0, off_cap = 0;
    felog::Corpus corpus;

    // SGNS tile descriptors (written by the prep kernel, read by the tile kernel)
    void *d_tiles = nullptr;
    uint64_t tiles_cap = 0;
    // SGNS tile shape (0: default): input rows 48/64, waves per SIMD the
    // kernel is compiled for, persistent workgroups per CU, centre windows
    // counted per wavefront (8), column parts of a row block in the update
    // products (1)
    uint32_t tile_rows = 0, tile_wpe = 0, tile_gpc = 0, tile_wpw = 0, tile_parts = 0;

    // scratch for library primitives
    void *d_tmp = nullptr;
    size_t tmp_bytes = 0;

    // device round state and the captured round
    felog::RoundState *d_rs = nullptr;
    felog::RoundState *h_rs = nullptr;   // pinned mirror
    hipGraphExec_t round_exec = nullptr;
    bool round_exec_budget = false;      // the captured round uses the budget
    bool round_exec_visits = false;      // the captured round adds corpus visits

    // learned walk budget: per-vertex walk count and length (0: plain rule)
    bool budget_on = false;
    uint16_t *d_budget_cnt = nullptr, *d_budget_len = nullptr;  // V
    felog_budget_model *d_model = nullptr;
    float *d_feat = nullptr;             // V x FELOG_BUDGET_FEATURES
    float *d_feat_part = nullptr;        // hub chunk partials of the features
    uint32_t *d_budget_sum = nullptr;    // prediction summary (4 words)
    uint32_t budget_rounds = 0;          // largest predicted count
    bool budget_plain_left = false;      // some vertex keeps the plain rule (count 0)
    bool budget_recheck = false;         // score budgeted vertices once at min_R
    // corpus visit histogram (cumulative over walk rounds while tracked)
    uint32_t *d_visits = nullptr;        // V, indexed by vocabulary rank
    uint32_t *d_row_ver = nullptr;       // 2V row versions (lost-update tracking; null: off)
    uint64_t visits_total = 0;
    double *d_vstats = nullptr;          // 3 doubles

    // per-pair word2vec trainers: filtered corpus, batch indices, the
    // unigram table and the sigmoid table
    uint32_t *d_w2v_ftok = nullptr, *d_w2v_kcnt = nullptr, *d_w2v_blk = nullptr;
    uint64_t w2v_cap_sent = 0, w2v_cap_tok = 0;
    int32_t *d_w2v_table = nullptr;
    uint32_t w2v_table_size = 0;
    float *d_w2v_exp = nullptr;
    // common neighbours: the CSR sorted by neighbour, and per CSR entry the
    // count the reference's transition reads (null: the edge data)
    uint32_t *d_scol = nullptr;
    float *d_cn = nullptr;

    felog_stats st{};
};

namespace felog {
// Counter slots in ctx->d_counters.
enum : int {
    C_WORK = 0,        // persistent-kernel work cursor
    C_STEPS = 1,       // walk steps taken
    C_OVF_NEXT = 2,    // overflow pool cursor
    C_OVF_FAIL = 3,    // walks that did not get overflow storage
    C_MAXLEN = 4,      // longest walk
    C_BADEDGE = 5,     // edges with out-of-range ids
    C_WORDS = 6,       // SGNS tokens read
    C_KEPT = 7,        // SGNS tokens kept
    C_PAIRS = 8,       // SGNS (centre, context) pairs
    C_GROUPS = 9,      // SGNS tiles trained
    C_TILES = 10,      // SGNS tile descriptors written
    C_EVALNUM = 11,    // vertices evaluated and not converged
    C_EVALCONV = 12,   // vertices flagged this evaluation
    C_DEBUG = 13,      // debug group cursor
    C_BIGTILES = 14,   // SGNS tiles with more than 32 input rows
    C_CURSOR = 15,     // SGNS tile cursor (zero between launches)
    C_DONE = 16,       // SGNS tile workgroups finished (zero between launches)
    C_ROWS_WB = 17,    // SGNS rows written back (lost-update tracking on)
    C_ROWS_LOST = 18,  // of those, writes over another tile's update to the row
    kNumCounters = 19
};

// Device memory for the build's temporaries: one block taken at the start of
// felog_build and handed out in 256-byte aligned pieces, so the build does
// not map and unmap gigabytes of temporaries stage by stage.  Pieces are
// released in stack order (release(mark)); a request the block cannot hold
// falls back to hipMalloc and is freed by done().  Every user of a piece
// runs on the context's stream, so a released piece is reused only by
// later work on that stream.
struct Scratch {
    char *base = nullptr;
    size_t cap = 0, top = 0;
    std::vector<void *> extra;
    hipError_t reserve(size_t bytes);
    hipError_t take(void **p, size_t bytes);
    template <typename T>
    hipError_t take(T **p, size_t bytes) {
        void *q = nullptr;
        const hipError_t e = take(&q, bytes);
        *p = (T *)q;
        return e;
    }
    size_t mark() const { return top; }
    void release(size_t m) { top = m; }
    // free everything; or (c set) keep the main block in c's spare blocks
    void done(felog_ctx *c = nullptr);
};

// graph_kernels.hip
hipError_t launch_undirect(felog_ctx *c, const uint32_t *src, const uint32_t *dst, const float *w,
                           uint64_t m, bool undirected);
// Bytes of scratch the build takes at its peak (edge list of c->E entries).
size_t build_scratch_bytes(felog_ctx *c);
// CSR; *edge_src receives the vertex of every CSR entry (device, E
// entries, in sc or in the context's edge input) for the alias build.
hipError_t build_csr(felog_ctx *c, Scratch &sc, uint32_t **edge_src);
hipError_t build_alias(felog_ctx *c, Scratch &sc, const uint32_t *edge_src);
hipError_t build_vocab_and_tables(felog_ctx *c, Scratch &sc);
// walk_kernel.hip
// Round state: eager calls set the round, its learning rate (alpha < 0: the
// schedule's) and, unless ~0u, the sentence count; captured rounds advance
// it on the device.
hipError_t round_set(felog_ctx *c, uint32_t round, float alpha, uint32_t n_sent);
hipError_t round_next(felog_ctx *c);
hipError_t round_end(felog_ctx *c);
hipError_t round_buffers(felog_ctx *c);
hipError_t walk_enqueue(felog_ctx *c);
hipError_t compact_enqueue(felog_ctx *c);
hipError_t run_walks(felog_ctx *c, float *walk_ms);
hipError_t compact_corpus(felog_ctx *c, float *ms);
// Schedule of the learning rate: base and per-round decay.
float alpha_base(const felog_options &o);
constexpr float kAlphaDecay = 0.1f;
// sgns_prep.hip (driver) + sgns_tile.hip
struct SgnsDebug {
    uint32_t max_groups;
    uint32_t *groups, *rows;
    float *logits, *dv, *du;
};
// Tile descriptor capacity for corpora of at most n_sent sentences and
// n_tok tokens (grows the buffer, which drops a captured round).
hipError_t sgns_reserve(felog_ctx *c, uint64_t n_sent, uint64_t n_tok);
// The pass over the corpus whose sentence count the round state holds,
// without a host wait (the tile descriptors must be reserved).
// mid (when set) is recorded between the prep and tile kernels.
hipError_t sgns_enqueue(felog_ctx *c, const uint64_t *d_off, const uint32_t *d_tok, const SgnsDebug *dbg,
                        hipEvent_t mid);
// ms[0] = whole training pass, ms[1] = prep kernel, ms[2] = tile kernel
hipError_t run_sgns(felog_ctx *c, const uint64_t *d_off, const uint32_t *d_tok, uint64_t n_sent,
                    uint64_t n_tok, const SgnsDebug *dbg, float *ms);
// The same with the tile descriptors in a caller-provided device buffer of
// tiles_cap descriptors (at least sgns_tiles_needed()); allocates nothing.
uint64_t sgns_tiles_needed(const felog_ctx *c, uint64_t n_sent, uint64_t n_tok);
constexpr uint64_t kSgnsDescBytes = 640;
hipError_t sgns_enqueue_into(felog_ctx *c, const uint64_t *d_off, const uint32_t *d_tok, const SgnsDebug *dbg,
                             hipEvent_t mid, void *tiles, uint64_t tiles_cap);
hipError_t run_sgns_into(felog_ctx *c, const uint64_t *d_off, const uint32_t *d_tok, const SgnsDebug *dbg,
                         float *ms, void *tiles, uint64_t tiles_cap);
// w2v_kernels.hip: the reference's per-pair trainers (opt.trainer != 0)
hipError_t w2v_prepare(felog_ctx *c);
hipError_t w2v_reserve(felog_ctx *c, uint64_t n_sent, uint64_t n_tok);
hipError_t w2v_enqueue(felog_ctx *c, const uint64_t *d_off, const uint32_t *d_tok);
hipError_t run_w2v(felog_ctx *c, const uint64_t *d_off, const uint32_t *d_tok, uint64_t n_sent, uint64_t n_tok,
                   float *ms);
// graph_tools.hip: common neighbours of n pairs (device arrays), of every
// CSR entry (into d_cn), and the held-out link-prediction AUC of n pairs
// (2n vertex ids, labels 0/1) from the cosine of their syn0 rows
hipError_t common_neighbours_pairs(felog_ctx *c, const uint32_t *d_u, const uint32_t *d_v, uint64_t n,
                                   int32_t *d_out);
hipError_t common_neighbours_csr(felog_ctx *c);
hipError_t link_auc(felog_ctx *c, const uint32_t *d_pairs, const uint8_t *d_lab, uint64_t n, float *d_score,
                    double *auc);
// k-means classes of the vocabulary's syn0 rows (FP32 tables) into d_cl
// (vocabulary order)
hipError_t kmeans_classes(felog_ctx *c, uint32_t K, uint32_t iters, int32_t *d_cl);
// eval_kernel.hip
// gated: evaluate only when the round state says so (captured rounds)
hipError_t eval_enqueue(felog_ctx *c, bool gated);
// the budget recheck at min_R (gated: only when the round state's round is
// min_R and the round walked)
hipError_t recheck_enqueue(felog_ctx *c, bool gated);
hipError_t run_eval(felog_ctx *c, float *ms);
// budget_kernels.hip
constexpr uint32_t kBudgetMaxCount = 255;  // rounds a budget may assign
hipError_t budget_features(felog_ctx *c, float *d_out);
hipError_t budget_grad(felog_ctx *c, const float *d_x, const float *d_t, uint64_t n, uint64_t sample_seed,
                       uint32_t step, uint32_t blocks, const float *d_p, float *d_grad, double *d_loss, float inv_n);
hipError_t budget_adam(felog_ctx *c, float *d_p, float *d_grad, float *d_m, float *d_v, float lr, uint32_t t);
// Counts and lengths of every vertex; d_sum (4 words, zeroed here) receives
// the largest count, the vertices left to the plain rule (count 0) and those
// of them in the vocabulary.
hipError_t budget_predict(felog_ctx *c, const float *d_feat, const felog_budget_model *d_model, uint32_t *d_sum);
// start flags (0/1 per walker) into d_ok
hipError_t budget_walk_ok(felog_ctx *c, uint32_t *d_ok);
hipError_t add_visits(felog_ctx *c, const uint32_t *d_tok, uint64_t n_tok);
// the round's corpus (token count from the round state)
hipError_t add_round_visits(felog_ctx *c);
hipError_t visit_stats(felog_ctx *c, double *d_out);
hipError_t init_tables(felog_ctx *c, Scratch *sc);

uint32_t sgns_group_centres(uint32_t batch, uint32_t window, uint32_t negative);
// Drop the captured round (its buffers are about to change).
void drop_round_graph(felog_ctx *c);
// Input rows of an SGNS tile: 48 when a group's windows fit, else 64.
uint32_t sgns_tile_rows(uint32_t centres, uint32_t window);
// Tile kernel variants: waves per SIMD lo and lo + 1, lo = 2 (FP32) or 3
// (BF16); workgroups of a shape that fit on one CU.
uint32_t tile_wpe_low(const felog_ctx *c);
hipError_t tile_occupancy(felog_ctx *c, uint32_t rows, uint32_t wpe, int *per_cu);

// Release device memory; a failure here has nothing left to report to.
inline void dfree(void *p) {
    if (p) (void)hipFree(p);
}
// Keep a whole block (base pointer of its allocation, its size) for reuse.
inline void dkeep(felog_ctx *c, void *p, size_t bytes) {
    if (p) c->spare.emplace_back(p, bytes);
}
// At least `bytes` of device memory: the smallest spare block that holds it
// and is at most twice its size, else a new allocation.  Either way the
// block is a whole allocation, released with dfree.
hipError_t dalloc(felog_ctx *c, void **p, size_t bytes);
template <typename T>
hipError_t dalloc(felog_ctx *c, T **p, size_t bytes) {
    void *q = nullptr;
    const hipError_t e = dalloc(c, &q, bytes);
    *p = (T *)q;
    return e;
}
}  // namespace felog
