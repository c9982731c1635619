// felog_api.cpp - the C ABI of libfelog_hip.so (declared in include/felog.h).
//
// Host orchestration of one FeLoG run on one GPU: graph input, device build
// (CSR, alias tables, vocabulary, tables), walk rounds, SGNS training,
// convergence evaluation, read-back and the reference's file formats.
// The round structure follows TrainModel (src/word2vec.cu:1922-2180) and
// WalkEngine::internal_random_walk (include/walk.hpp:630-840), made serial
// and deterministic: round r walks every unconverged vertex, trains on that
// corpus with alpha_r, and from round min_R on evaluates and flags vertices.
#include <hip/hip_runtime.h>
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cctype>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include "felog_internal.h"

using namespace felog;

namespace {

felog_status hip_fail(felog_ctx *c, hipError_t e, const char *what) {
    c->err = std::string(what) + ": " + hipGetErrorString(e);
    return e == hipErrorOutOfMemory ? FELOG_ERR_OOM : FELOG_ERR_HIP;
}

felog_status fail(felog_ctx *c, felog_status code, const std::string &msg) {
    if (c) c->err = msg;
    return code;
}

#define TRY(c, x)                                               \
    do {                                                        \
        hipError_t e_ = (x);                                    \
        if (e_ != hipSuccess) return hip_fail((c), e_, #x);     \
    } while (0)

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

namespace felog {
// Base of the learning-rate schedule (the rate of round r is
// base * e^{-0.1 (r-1)}, computed on the device).  The reference always uses
// base 0.025; the default here honours -alpha as the base.  A tile sums all
// of its pair gradients at one point (the reference's sequential per-pair
// updates let the sigmoid saturate between pairs), so the default base is
// 0.0025: on the 1.1M R-MAT and 2.24M Chung-Lu graphs 0.025 and 0.005 drive
// held-out AUC down after one or two rounds, 0.0025 keeps it at or above
// word2vec's.
float alpha_base(const felog_options &o) { return o.compat_lr ? 0.025f : o.alpha; }

void drop_round_graph(felog_ctx *c) {
    if (c->round_exec) (void)hipGraphExecDestroy(c->round_exec);
    c->round_exec = nullptr;
}
}  // namespace felog

namespace {

void free_all(felog_ctx *c) {
    void *ptrs[] = {c->d_esrc, c->d_edst, c->d_ew, c->d_row_ptr, c->d_col, c->d_w, c->d_vrec, c->d_vwalk, c->d_erec,
                    c->d_deg, c->d_vocab_order, c->d_vocab_rank, c->d_in_vocab, c->d_uni,
                    c->d_keep_thr, c->d_T, c->d_dT, c->d_syn0, c->d_flags,
                    c->d_scores, c->d_active, c->d_sel, c->d_slots, c->d_lens, c->d_ovf_base, c->d_ovf,
                    c->d_counters, c->d_off, c->d_tok, c->d_sent_id, c->d_tiles, c->d_tmp,
                    c->d_budget_cnt, c->d_budget_len, c->d_model, c->d_feat, c->d_visits, c->d_vstats,
                    c->d_w2v_ftok, c->d_w2v_kcnt, c->d_w2v_blk, c->d_w2v_table, c->d_w2v_exp, c->d_scol,
                    c->d_cn, c->hubs.hub, c->hubs.cstart, c->d_feat_part, c->d_budget_sum, c->d_row_ver};
    for (void *p : ptrs) dfree(p);
    for (auto &b : c->spare) dfree(b.first);
    c->spare.clear();
    drop_round_graph(c);
    dfree(c->d_rs);
    if (c->h_rs) (void)hipHostFree(c->h_rs);
    if (c->h_counters) (void)hipHostFree(c->h_counters);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->ev2) (void)hipEventDestroy(c->ev2);
    if (c->stream) (void)hipStreamDestroy(c->stream);
}

felog_status need_built(felog_ctx *c) {
    if (!c) return FELOG_ERR_ARG;
    if (!c->built) return fail(c, FELOG_ERR_STATE, "felog_build has not run");
    TRY(c, hipSetDevice(c->device));
    return FELOG_OK;
}

}  // namespace

extern "C" {

void felog_default_options(felog_options *o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->min_L = 20;
    o->min_R = 5;
    o->eval_batch = 16384;
    o->size = 100;
    o->window = 5;
    o->negative = 5;
    o->batch_size = 0;
    o->sample = 1e-3f;
    o->alpha = 0.0025f;  // word2vec's 0.025 is too large a step for tile-summed gradients
    o->min_count = 5;
    o->precision = FELOG_FP32;
    o->max_walk_len = 1000;
    o->overflow_capacity = 1ull << 24;
    o->seed = 0x5EEDF00DULL;
}

felog_status felog_create(felog_ctx **out, const felog_options *opt) {
    if (!out || !opt) return FELOG_ERR_ARG;
    *out = nullptr;
    felog_ctx *c = new felog_ctx();
    c->opt = *opt;
    felog_options &o = c->opt;
    if (o.min_L == 0) o.min_L = 20;   // include/walk.hpp:511-514
    if (o.min_R == 0) o.min_R = 5;    // include/walk.hpp:506-509
    if (o.max_walk_len == 0) o.max_walk_len = 1000;
    if (o.overflow_capacity == 0) o.overflow_capacity = 1ull << 24;
    std::string bad;
    if (o.size == 0 || o.size > 512) bad = "-size must be in [1, 512]";
    else if (o.window == 0 || o.window > 26) bad = "-window must be in [1, 26]";
    else if (o.negative > 24) bad = "-negative must be <= 24";
    else if (o.precision != FELOG_FP32 && o.precision != FELOG_BF16) bad = "precision must be FP32 or BF16";
    else if (o.max_walk_len < 2 || o.max_walk_len > (1u << 20)) bad = "max_walk_len must be in [2, 2^20]";
    else if (!(o.alpha > 0.f)) bad = "-alpha must be > 0";
    else if (o.update_mode < FELOG_UPDATE_DEFAULT || o.update_mode > FELOG_UPDATE_WRITEBACK) bad = "bad update_mode";
    else if (o.trainer < FELOG_TRAIN_HOGBATCH || o.trainer > FELOG_TRAIN_CBOW) bad = "bad trainer";
    else if (o.trainer != FELOG_TRAIN_HOGBATCH && o.precision != FELOG_FP32)
        bad = "the per-pair trainers update FP32 tables";
    if (bad.empty() && sgns_group_centres(o.batch_size, o.window, o.negative) < 1)
        bad = "-window/-negative leave no room for a centre in an SGNS group";
    if (!bad.empty()) {
        delete c;
        return FELOG_ERR_ARG;
    }
    c->d = o.size;
    c->dp = (o.size + 31) / 32 * 32;
    if (o.precision == FELOG_FP32 && c->dp > 320) {
        delete c;
        return FELOG_ERR_ARG;  // f32 rows beyond 320 do not fit the SGNS LDS tile
    }
    c->device = o.device;
    hipError_t e = hipSetDevice(c->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&c->ev0);
    if (e == hipSuccess) e = hipEventCreate(&c->ev1);
    if (e == hipSuccess) e = hipEventCreate(&c->ev2);
    if (e == hipSuccess) e = hipMalloc(&c->d_counters, kNumCounters * 8);
    if (e == hipSuccess) e = hipHostMalloc(&c->h_counters, kNumCounters * 8, hipHostMallocDefault);
    if (e == hipSuccess) e = hipMalloc(&c->d_rs, sizeof(RoundState));
    if (e == hipSuccess) e = hipHostMalloc(&c->h_rs, sizeof(RoundState), hipHostMallocDefault);
    if (e == hipSuccess) {
        std::memset(c->h_rs, 0, sizeof(RoundState));
        c->h_rs->eval_from = o.min_R;
        e = hipMemcpy(c->d_rs, c->h_rs, sizeof(RoundState), hipMemcpyHostToDevice);
    }
    int cus = 0;
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device);
    if (e != hipSuccess) {
        free_all(c);
        delete c;
        return e == hipErrorOutOfMemory ? FELOG_ERR_OOM : FELOG_ERR_HIP;
    }
    c->cu_count = (uint32_t)cus;
    *out = c;
    return FELOG_OK;
}

felog_status felog_destroy(felog_ctx *c) {
    if (!c) return FELOG_ERR_ARG;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    free_all(c);
    delete c;
    return FELOG_OK;
}

const char *felog_last_error(const felog_ctx *c) {
    return c ? c->err.c_str() : "null context";
}

felog_status felog_set_edges(felog_ctx *c, const uint32_t *src, const uint32_t *dst, const float *w, uint64_t m,
                             uint32_t v_num, int make_undirected) {
    if (!c) return FELOG_ERR_ARG;
    if (c->have_edges || c->d_row_ptr) return fail(c, FELOG_ERR_STATE, "edges already set on this context");
    if (v_num == 0) return fail(c, FELOG_ERR_ARG, "v_num must be > 0");
    if (m > 0 && (!src || !dst || !w)) return fail(c, FELOG_ERR_ARG, "null edge array");
    const uint64_t E = make_undirected ? 2 * m : m;
    if (E >= (1ull << 32)) return fail(c, FELOG_ERR_ARG, "directed edge count must be < 2^32");
    TRY(c, hipSetDevice(c->device));
    c->V = v_num;
    if (m == 0) {
        c->E = 0;
    } else {
        TRY(c, launch_undirect(c, src, dst, w, m, make_undirected != 0));
        uint64_t bad = 0;
        TRY(c, hipMemcpy(&bad, c->d_counters + C_BADEDGE, 8, hipMemcpyDeviceToHost));
        if (bad) {
            dfree(c->d_esrc); dfree(c->d_edst); dfree(c->d_ew);
            c->d_esrc = c->d_edst = nullptr; c->d_ew = nullptr;
            return fail(c, FELOG_ERR_ARG, std::to_string(bad) + " edges have a vertex id >= v_num");
        }
    }
    c->have_edges = true;
    return FELOG_OK;
}

felog_status felog_load_edges(felog_ctx *c, const char *path, uint32_t v_num, int make_undirected) {
    if (!c || !path) return FELOG_ERR_ARG;
    const int fd = open(path, O_RDONLY);
    if (fd < 0) return fail(c, FELOG_ERR_IO, std::string("cannot open ") + path);
    struct stat sb;
    if (fstat(fd, &sb) != 0) {
        close(fd);
        return fail(c, FELOG_ERR_IO, std::string("cannot stat ") + path);
    }
    const uint64_t bytes = (uint64_t)sb.st_size;
    if (bytes % 12 != 0) {
        close(fd);
        return fail(c, FELOG_ERR_IO, std::string(path) + ": size is not a multiple of 12-byte edge records");
    }
    const uint64_t m = bytes / 12;
    std::vector<uint32_t> src(m), dst(m);
    std::vector<float> w(m);
    if (m > 0) {
        void *map = mmap(nullptr, bytes, PROT_READ, MAP_PRIVATE, fd, 0);
        if (map == MAP_FAILED) {
            close(fd);
            return fail(c, FELOG_ERR_IO, std::string("cannot map ") + path);
        }
        const unsigned char *b = (const unsigned char *)map;
        for (uint64_t i = 0; i < m; ++i) {
            std::memcpy(&src[i], b + 12 * i, 4);
            std::memcpy(&dst[i], b + 12 * i + 4, 4);
            std::memcpy(&w[i], b + 12 * i + 8, 4);
        }
        munmap(map, bytes);
    }
    close(fd);
    return felog_set_edges(c, src.data(), dst.data(), w.data(), m, v_num, make_undirected);
}

felog_status felog_build(felog_ctx *c) {
    if (!c) return FELOG_ERR_ARG;
    if (!c->have_edges) return fail(c, FELOG_ERR_STATE, "no edges: call felog_set_edges or felog_load_edges");
    if (c->built) return fail(c, FELOG_ERR_STATE, "already built");
    TRY(c, hipSetDevice(c->device));
    const double t0 = now_ms();
    if (c->E == 0) {
        // a graph without edges still gets its per-vertex arrays
        TRY(c, hipMalloc(&c->d_esrc, 4));
        TRY(c, hipMalloc(&c->d_edst, 4));
        TRY(c, hipMalloc(&c->d_ew, 4));
    }
    // the build's temporaries come from one block (falls back to separate
    // allocations past it)
    Scratch sc;
    hipError_t be = sc.reserve(build_scratch_bytes(c));
    uint32_t *edge_src = nullptr;
    if (be == hipSuccess) be = build_csr(c, sc, &edge_src);
    if (be == hipSuccess) be = build_vocab_and_tables(c, sc);
    if (be == hipSuccess) be = build_alias(c, sc, edge_src);
    if (be == hipSuccess) be = init_tables(c, &sc);
    // the sorted sources were either in the scratch or the edge input's
    // source block: both blocks are kept for the buffers that follow
    dkeep(c, c->d_esrc, c->E * 4 + 4);
    c->d_esrc = nullptr;
    sc.done(c);
    // the edge input is consumed either way: a failed build cannot be
    // retried on this context
    if (be != hipSuccess) c->have_edges = false;
    TRY(c, be);
    const uint32_t V = c->V;
    TRY(c, dalloc(c, &c->d_flags, V));
    TRY(c, hipMemsetAsync(c->d_flags, 0, V, c->stream));
    TRY(c, dalloc(c, &c->d_scores, (uint64_t)V * 4));
    TRY(c, hipMemsetAsync(c->d_scores, 0, (uint64_t)V * 4, c->stream));
    // walk buffers
    c->walkers = c->opt.walkers ? c->opt.walkers : V;
    uint32_t slot = (c->opt.min_L + 2 + 3) / 4 * 4;
    if (slot > c->opt.max_walk_len) slot = c->opt.max_walk_len;
    c->slot = slot;
    const uint64_t W = c->walkers;
    TRY(c, dalloc(c, &c->d_slots, W * slot * 4));
    TRY(c, dalloc(c, &c->d_active, W * 4));
    TRY(c, dalloc(c, &c->d_sel, W * 4));
    TRY(c, dalloc(c, &c->d_lens, W * 4));
    TRY(c, dalloc(c, &c->d_ovf_base, W * 8));
    c->ovf_cap = c->opt.max_walk_len > slot ? c->opt.overflow_capacity : 0;
    TRY(c, dalloc(c, &c->d_ovf, (c->ovf_cap ? c->ovf_cap : 1) * 4));
    TRY(c, round_buffers(c));
    // the build is done when its last copy and clear are
    TRY(c, hipStreamSynchronize(c->stream));
    c->last_eva_num = (uint32_t)c->vocab_size;
    c->n_unconverged = (uint32_t)c->vocab_size;
    c->st.build_ms = now_ms() - t0;
    c->st.v_num = V;
    c->st.e_num = c->E;
    c->st.vocab_size = c->vocab_size;
    c->built = true;
    return FELOG_OK;
}

felog_status felog_walk_round(felog_ctx *c, uint32_t round, felog_corpus *out) {
    felog_status s = need_built(c);
    if (s) return s;
    if (round == 0) return fail(c, FELOG_ERR_ARG, "rounds are numbered from 1");
    float wms = 0.f, cms = 0.f;
    c->corpus = Corpus();
    TRY(c, round_set(c, round, -1.f, ~0u));
    TRY(c, run_walks(c, &wms));
    const uint64_t fail_n = c->h_counters[C_OVF_FAIL];
    const uint64_t steps = c->h_counters[C_STEPS];
    const uint64_t maxlen = c->h_counters[C_MAXLEN];
    c->st.overflow_used = c->h_counters[C_OVF_NEXT];
    TRY(c, compact_corpus(c, &cms));
    if (c->d_visits) TRY(c, add_visits(c, c->d_tok, c->corpus.n_tokens));
    c->corpus.round = round;
    c->corpus.n_steps = steps;
    c->corpus.max_len = (uint32_t)maxlen;
    c->corpus.valid = true;
    c->corpus.from_walk = true;
    c->st.walk_steps = steps;
    c->st.walk_tokens = c->corpus.n_tokens;
    c->st.walk_tokens_total += c->corpus.n_tokens;
    c->st.walk_kernel_ms = wms;
    c->st.compact_kernel_ms = cms;
    c->st.walk_ms_total += wms + cms;
    if (out) {
        out->round = round;
        out->max_len = c->corpus.max_len;
        out->n_walks = c->corpus.n_walks;
        out->n_tokens = c->corpus.n_tokens;
        out->n_steps = steps;
    }
    if (fail_n)
        return fail(c, FELOG_ERR_CAPACITY,
                    std::to_string(fail_n) + " walks outgrew the overflow pool; raise overflow_capacity");
    return FELOG_OK;
}

// alpha < 0: the schedule's rate for the round
static felog_status train_device_corpus(felog_ctx *c, const uint64_t *d_off, const uint32_t *d_tok, uint64_t n,
                                        uint64_t n_tok, uint32_t round, float alpha) {
    if (n >= 0xFFFFFFFFull) return fail(c, FELOG_ERR_ARG, "a corpus holds fewer than 2^32 - 1 sentences");
    float ms[3] = {0.f, 0.f, 0.f};
    TRY(c, round_set(c, round, alpha, (uint32_t)n));
    if (c->opt.trainer != FELOG_TRAIN_HOGBATCH) TRY(c, run_w2v(c, d_off, d_tok, n, n_tok, ms));
    else TRY(c, run_sgns(c, d_off, d_tok, n, n_tok, nullptr, ms));
    c->st.train_words = c->h_counters[C_WORDS];
    c->st.train_kept = c->h_counters[C_KEPT];
    c->st.train_pairs = c->h_counters[C_PAIRS];
    c->st.train_groups = c->h_counters[C_GROUPS];
    c->st.train_rows_written = c->h_counters[C_ROWS_WB];
    c->st.train_rows_lost = c->h_counters[C_ROWS_LOST];
    c->st.train_kernel_ms = ms[0];
    c->st.sgns_prep_ms = ms[1];
    c->st.sgns_tile_ms = ms[2];
    c->st.train_ms_total += ms[0];
    if (c->opt.trainer == FELOG_TRAIN_HOGBATCH && n + c->h_counters[C_TILES] > c->tiles_cap)
        return fail(c, FELOG_ERR_INVARIANT, "SGNS tile descriptors exceeded their bound");
    return FELOG_OK;
}

felog_status felog_train_round(felog_ctx *c, const felog_corpus *corpus) {
    felog_status s = need_built(c);
    if (s) return s;
    if (!c->corpus.valid || !c->corpus.from_walk)
        return fail(c, FELOG_ERR_STATE, "no walk corpus: call felog_walk_round first");
    if (corpus && corpus->round != c->corpus.round)
        return fail(c, FELOG_ERR_ARG, "corpus handle is not the context's current corpus");
    const uint32_t round = c->corpus.round;
    s = train_device_corpus(c, c->d_off, c->d_tok, c->corpus.n_walks, c->corpus.n_tokens, round, -1.f);
    if (s) return s;
    c->rounds_done = round;
    return FELOG_OK;
}

felog_status felog_train_corpus(felog_ctx *c, const uint64_t *offsets, const uint32_t *tokens, uint64_t n,
                                uint32_t round) {
    felog_status s = need_built(c);
    if (s) return s;
    if (round == 0) return fail(c, FELOG_ERR_ARG, "rounds are numbered from 1");
    if (n > 0 && !offsets) return fail(c, FELOG_ERR_ARG, "null offsets");
    if (n == 0) return FELOG_OK;
    const uint64_t nt = offsets[n];
    if (offsets[0] != 0) return fail(c, FELOG_ERR_ARG, "offsets[0] must be 0");
    for (uint64_t i = 0; i < n; ++i)
        if (offsets[i + 1] < offsets[i]) return fail(c, FELOG_ERR_ARG, "offsets must be non-decreasing");
    if (nt > 0 && !tokens) return fail(c, FELOG_ERR_ARG, "null tokens");
    for (uint64_t i = 0; i < nt; ++i)
        if (tokens[i] >= c->V) return fail(c, FELOG_ERR_ARG, "token id >= v_num");
    uint64_t *d_off = nullptr;
    uint32_t *d_tok = nullptr;
    TRY(c, hipMalloc(&d_off, (n + 1) * 8));
    TRY(c, hipMalloc(&d_tok, (nt ? nt : 1) * 4));
    TRY(c, hipMemcpy(d_off, offsets, (n + 1) * 8, hipMemcpyHostToDevice));
    if (nt) TRY(c, hipMemcpy(d_tok, tokens, nt * 4, hipMemcpyHostToDevice));
    s = train_device_corpus(c, d_off, d_tok, n, nt, round, -1.f);
    dfree(d_off);
    dfree(d_tok);
    return s;
}

felog_status felog_debug_sgns(felog_ctx *c, const uint64_t *offsets, const uint32_t *tokens, uint64_t n,
                              uint32_t round, float alpha, uint32_t max_groups, uint32_t *n_groups,
                              uint32_t *groups_out, uint32_t *rows_out, float *logits_out, float *dv_out,
                              float *du_out) {
    felog_status s = need_built(c);
    if (s) return s;
    if (!offsets || !n_groups || !groups_out || !rows_out || !logits_out || !dv_out || !du_out || max_groups == 0)
        return fail(c, FELOG_ERR_ARG, "null debug buffer");
    if (c->opt.trainer != FELOG_TRAIN_HOGBATCH) return fail(c, FELOG_ERR_STATE, "the tile trainer is not selected");
    if (n >= 0xFFFFFFFFull) return fail(c, FELOG_ERR_ARG, "a corpus holds fewer than 2^32 - 1 sentences");
    const uint64_t nt = offsets[n];
    for (uint64_t i = 0; i < nt; ++i)
        if (tokens[i] >= c->V) return fail(c, FELOG_ERR_ARG, "token id >= v_num");
    uint64_t *d_off = nullptr;
    uint32_t *d_tok = nullptr;
    SgnsDebug dbg;
    dbg.max_groups = max_groups;
    TRY(c, hipMalloc(&d_off, (n + 1) * 8));
    TRY(c, hipMalloc(&d_tok, (nt ? nt : 1) * 4));
    TRY(c, hipMemcpy(d_off, offsets, (n + 1) * 8, hipMemcpyHostToDevice));
    if (nt) TRY(c, hipMemcpy(d_tok, tokens, nt * 4, hipMemcpyHostToDevice));
    const uint64_t G = max_groups;
    TRY(c, hipMalloc(&dbg.groups, G * 5 * 4));
    TRY(c, hipMalloc(&dbg.rows, G * 96 * 4));
    TRY(c, hipMalloc(&dbg.logits, G * 64 * 32 * 4));
    TRY(c, hipMalloc(&dbg.dv, G * 64 * c->d * 4));
    TRY(c, hipMalloc(&dbg.du, G * 32 * c->d * 4));
    TRY(c, hipMemset(dbg.groups, 0, G * 5 * 4));
    TRY(c, hipMemset(dbg.logits, 0, G * 64 * 32 * 4));
    TRY(c, hipMemset(dbg.dv, 0, G * 64 * c->d * 4));
    TRY(c, hipMemset(dbg.du, 0, G * 32 * c->d * 4));
    float ms[3];
    TRY(c, round_set(c, round, alpha, (uint32_t)n));
    TRY(c, run_sgns(c, d_off, d_tok, n, nt, &dbg, ms));
    // slots: one per sentence (empty when it has no centre) plus extra tiles
    const uint64_t ng = n + c->h_counters[C_TILES];
    *n_groups = (uint32_t)(ng < G ? ng : G);
    const uint64_t k = *n_groups;
    TRY(c, hipMemcpy(groups_out, dbg.groups, k * 5 * 4, hipMemcpyDeviceToHost));
    TRY(c, hipMemcpy(rows_out, dbg.rows, k * 96 * 4, hipMemcpyDeviceToHost));
    TRY(c, hipMemcpy(logits_out, dbg.logits, k * 64 * 32 * 4, hipMemcpyDeviceToHost));
    TRY(c, hipMemcpy(dv_out, dbg.dv, k * 64 * c->d * 4, hipMemcpyDeviceToHost));
    TRY(c, hipMemcpy(du_out, dbg.du, k * 32 * c->d * 4, hipMemcpyDeviceToHost));
    dfree(d_off); dfree(d_tok);
    dfree(dbg.groups); dfree(dbg.rows); dfree(dbg.logits); dfree(dbg.dv); dfree(dbg.du);
    if (ng > G) return fail(c, FELOG_ERR_CAPACITY, "more groups than max_groups");
    return FELOG_OK;
}

felog_status felog_evaluate(felog_ctx *c, uint32_t *n_unconverged, int *stop) {
    felog_status s = need_built(c);
    if (s) return s;
    float ms = 0.f;
    TRY(c, round_set(c, c->rounds_done, -1.f, ~0u));
    TRY(c, run_eval(c, &ms));
    const uint32_t eva = (uint32_t)c->h_counters[C_EVALNUM];
    // src/word2vec.cu:2073-2078: stop once the unconverged count shrinks by
    // less than 30% in a round.
    const float ratio = c->last_eva_num ? (float)eva / (float)c->last_eva_num : 0.f;
    const bool st = c->last_eva_num != 0 && ratio > 0.7f;
    c->last_eva_num = eva;
    c->n_unconverged = eva;
    c->stop = st;
    c->st.n_unconverged = eva;
    c->st.eval_kernel_ms = ms;
    c->st.eval_ms_total += ms;
    if (n_unconverged) *n_unconverged = eva;
    if (stop) *stop = st ? 1 : 0;
    return FELOG_OK;
}

felog_status felog_run(felog_ctx *c) {
    felog_status s = need_built(c);
    if (s) return s;
    for (uint32_t r = c->rounds_done + 1;; ++r) {
        if (c->opt.max_rounds && r > c->opt.max_rounds) break;
        felog_corpus corpus;
        if ((s = felog_walk_round(c, r, &corpus))) return s;
        if ((s = felog_train_round(c, &corpus))) return s;
        if (r >= c->opt.min_R) {
            uint32_t nu = 0;
            int stop = 0;
            if ((s = felog_evaluate(c, &nu, &stop))) return s;
            if (stop) break;
        }
        if (corpus.n_walks == 0) break;
    }
    return FELOG_OK;
}

// ---- learned walk budget ------------------------------------------------

felog_status felog_budget_features(felog_ctx *c, float *features) {
    felog_status s = need_built(c);
    if (s) return s;
    if (!features) return fail(c, FELOG_ERR_ARG, "null features buffer");
    const uint64_t bytes = (uint64_t)c->V * FELOG_BUDGET_FEATURES * 4;
    if (!c->d_feat) TRY(c, hipMalloc(&c->d_feat, bytes));
    TRY(c, budget_features(c, c->d_feat));
    TRY(c, hipMemcpyAsync(features, c->d_feat, bytes, hipMemcpyDeviceToHost, c->stream));
    TRY(c, hipStreamSynchronize(c->stream));
    return FELOG_OK;
}

namespace {
// The last round a budgeted vertex can walk: the largest count, or min_R + 1
// when the recheck may extend counts that end at min_R.
uint32_t last_budget_round(const felog_ctx *c) {
    const uint32_t r = c->budget_rounds;
    return c->budget_recheck && c->opt.min_R + 1 > r ? c->opt.min_R + 1 : r;
}

// Device copies of a training set and the optimiser state; freed on return.
struct MlpBufs {
    float *x = nullptr, *t = nullptr, *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
    double *loss = nullptr;
    ~MlpBufs() {
        for (void *q : {(void *)x, (void *)t, (void *)p, (void *)g, (void *)m, (void *)v, (void *)loss}) dfree(q);
    }
};
}  // namespace

felog_status felog_budget_grad(felog_ctx *c, const float *X, const float *T, uint64_t n, const float *params,
                               float *grad, double *loss) {
    if (!c) return FELOG_ERR_ARG;
    if (!X || !T || !params || !grad || n == 0) return fail(c, FELOG_ERR_ARG, "budget_grad: bad arguments");
    TRY(c, hipSetDevice(c->device));
    MlpBufs b;
    const size_t P = FELOG_BUDGET_PARAMS * 4;
    TRY(c, hipMalloc(&b.x, n * FELOG_BUDGET_FEATURES * 4));
    TRY(c, hipMalloc(&b.t, n * 2 * 4));
    TRY(c, hipMalloc(&b.p, P));
    TRY(c, hipMalloc(&b.g, P));
    TRY(c, hipMalloc(&b.loss, 8));
    TRY(c, hipMemcpyAsync(b.x, X, n * FELOG_BUDGET_FEATURES * 4, hipMemcpyHostToDevice, c->stream));
    TRY(c, hipMemcpyAsync(b.t, T, n * 2 * 4, hipMemcpyHostToDevice, c->stream));
    TRY(c, hipMemcpyAsync(b.p, params, P, hipMemcpyHostToDevice, c->stream));
    TRY(c, hipMemsetAsync(b.g, 0, P, c->stream));
    TRY(c, hipMemsetAsync(b.loss, 0, 8, c->stream));
    const uint64_t blocks = (n + 63) / 64;
    if (blocks > 0x7FFFFFFFull) return fail(c, FELOG_ERR_ARG, "budget_grad: too many samples");
    TRY(c, budget_grad(c, b.x, b.t, n, 0, 0, (uint32_t)blocks, b.p, b.g, b.loss, 1.f / (float)n));
    double l = 0.0;
    TRY(c, hipMemcpyAsync(grad, b.g, P, hipMemcpyDeviceToHost, c->stream));
    TRY(c, hipMemcpyAsync(&l, b.loss, 8, hipMemcpyDeviceToHost, c->stream));
    TRY(c, hipStreamSynchronize(c->stream));
    if (loss) *loss = l / (double)n;
    return FELOG_OK;
}

felog_status felog_budget_train(felog_ctx *c, const float *X, const float *T, uint64_t n, uint32_t epochs,
                                uint32_t batch, float lr, uint64_t seed, float *params, double *loss_per_epoch) {
    if (!c) return FELOG_ERR_ARG;
    if (!X || !T || !params || n == 0 || epochs == 0 || batch == 0 || batch % 64 || !(lr > 0.f))
        return fail(c, FELOG_ERR_ARG, "budget_train: bad arguments (batch must be a positive multiple of 64)");
    TRY(c, hipSetDevice(c->device));
    MlpBufs b;
    const size_t P = FELOG_BUDGET_PARAMS * 4;
    TRY(c, hipMalloc(&b.x, n * FELOG_BUDGET_FEATURES * 4));
    TRY(c, hipMalloc(&b.t, n * 2 * 4));
    TRY(c, hipMalloc(&b.p, P));
    TRY(c, hipMalloc(&b.g, P));
    TRY(c, hipMalloc(&b.m, P));
    TRY(c, hipMalloc(&b.v, P));
    TRY(c, hipMalloc(&b.loss, 8));
    TRY(c, hipMemcpyAsync(b.x, X, n * FELOG_BUDGET_FEATURES * 4, hipMemcpyHostToDevice, c->stream));
    TRY(c, hipMemcpyAsync(b.t, T, n * 2 * 4, hipMemcpyHostToDevice, c->stream));
    TRY(c, hipMemcpyAsync(b.p, params, P, hipMemcpyHostToDevice, c->stream));
    TRY(c, hipMemsetAsync(b.g, 0, P, c->stream));
    TRY(c, hipMemsetAsync(b.m, 0, P, c->stream));
    TRY(c, hipMemsetAsync(b.v, 0, P, c->stream));
    const uint64_t steps = (n + batch - 1) / batch;
    const uint64_t key = seed ? seed : 1;
    uint32_t t = 0;
    for (uint32_t e = 0; e < epochs; ++e) {
        TRY(c, hipMemsetAsync(b.loss, 0, 8, c->stream));
        for (uint64_t k = 0; k < steps; ++k) {
            ++t;
            TRY(c, budget_grad(c, b.x, b.t, n, key, t, batch / 64, b.p, b.g, b.loss, 1.f / (float)batch));
            TRY(c, budget_adam(c, b.p, b.g, b.m, b.v, lr, t));
        }
        if (loss_per_epoch) {
            double l = 0.0;
            TRY(c, hipMemcpyAsync(&l, b.loss, 8, hipMemcpyDeviceToHost, c->stream));
            TRY(c, hipStreamSynchronize(c->stream));
            loss_per_epoch[e] = l / ((double)steps * batch);
        }
    }
    TRY(c, hipMemcpyAsync(params, b.p, P, hipMemcpyDeviceToHost, c->stream));
    TRY(c, hipStreamSynchronize(c->stream));
    return FELOG_OK;
}

felog_status felog_budget_set(felog_ctx *c, const felog_budget_model *model) {
    felog_status s = need_built(c);
    if (s) return s;
    if (!model) {
        c->budget_on = false;
        c->budget_rounds = 0;
        return FELOG_OK;
    }
    const uint64_t fbytes = (uint64_t)c->V * FELOG_BUDGET_FEATURES * 4;
    if (!c->d_feat) TRY(c, hipMalloc(&c->d_feat, fbytes));
    if (!c->d_model) TRY(c, hipMalloc(&c->d_model, sizeof(felog_budget_model)));
    if (!c->d_budget_cnt) TRY(c, hipMalloc(&c->d_budget_cnt, (uint64_t)c->V * 2));
    if (!c->d_budget_len) TRY(c, hipMalloc(&c->d_budget_len, (uint64_t)c->V * 2));
    if (!c->d_budget_sum) TRY(c, hipMalloc(&c->d_budget_sum, 16));
    TRY(c, hipMemcpyAsync(c->d_model, model, sizeof(felog_budget_model), hipMemcpyHostToDevice, c->stream));
    TRY(c, budget_features(c, c->d_feat));
    TRY(c, budget_predict(c, c->d_feat, c->d_model, c->d_budget_sum));
    uint32_t sum[4] = {0, 0, 0, 0};
    TRY(c, hipMemcpyAsync(sum, c->d_budget_sum, 16, hipMemcpyDeviceToHost, c->stream));
    TRY(c, hipStreamSynchronize(c->stream));
    c->budget_rounds = sum[0];
    c->budget_plain_left = sum[1] != 0;
    c->budget_on = true;
    // the plain rule's stop ratio now runs over the plain-rule vertices only:
    // its first denominator is their vocabulary count (src/word2vec.cu:1980)
    c->last_eva_num = sum[2];
    return FELOG_OK;
}

felog_status felog_budget_recheck(felog_ctx *c, int on) {
    if (!c) return FELOG_ERR_ARG;
    const bool v = on != 0;
    // the captured round holds the recheck launch or not
    if (v != c->budget_recheck) drop_round_graph(c);
    c->budget_recheck = v;
    return FELOG_OK;
}

felog_status felog_budget_copy_out(felog_ctx *c, uint16_t *count, uint16_t *length, uint32_t *rounds) {
    felog_status s = need_built(c);
    if (s) return s;
    if (!c->budget_on) return fail(c, FELOG_ERR_STATE, "no budget installed");
    if (length) TRY(c, hipMemcpy(length, c->d_budget_len, (uint64_t)c->V * 2, hipMemcpyDeviceToHost));
    if (count || rounds) {
        // the recheck may have raised counts since the model was installed
        std::vector<uint16_t> tmp;
        uint16_t *cnt = count;
        if (!cnt) {
            tmp.resize(c->V);
            cnt = tmp.data();
        }
        TRY(c, hipMemcpy(cnt, c->d_budget_cnt, (uint64_t)c->V * 2, hipMemcpyDeviceToHost));
        uint32_t mx = 0;
        for (uint32_t v = 0; v < c->V; ++v) mx = cnt[v] > mx ? cnt[v] : mx;
        if (rounds) *rounds = mx;
    }
    return FELOG_OK;
}

felog_status felog_run_budget(felog_ctx *c) {
    felog_status s = need_built(c);
    if (s) return s;
    if (!c->budget_on) return fail(c, FELOG_ERR_STATE, "no budget installed");
    // plain-rule vertices: those with count 0
    bool plain_stopped = !c->budget_plain_left;
    const uint32_t last = last_budget_round(c);
    for (uint32_t r = c->rounds_done + 1;; ++r) {
        if (c->opt.max_rounds && r > c->opt.max_rounds) break;
        if (r > last && plain_stopped) break;
        felog_corpus corpus;
        if ((s = felog_walk_round(c, r, &corpus))) return s;
        if (corpus.n_walks == 0) break;
        if ((s = felog_train_round(c, &corpus))) return s;
        if (c->budget_recheck && r == c->opt.min_R) {
            TRY(c, recheck_enqueue(c, false));
            TRY(c, hipStreamSynchronize(c->stream));
        }
        if (!plain_stopped && r >= c->opt.min_R) {
            uint32_t nu = 0;
            int stop = 0;
            if ((s = felog_evaluate(c, &nu, &stop))) return s;
            if (stop) plain_stopped = true;
        }
    }
    return FELOG_OK;
}

// ---- captured rounds -------------------------------------------------------

}  // extern "C"

namespace {

// One round as a graph: round advance (counters cleared by a kernel),
// walker select and walks, corpus, visits, SGNS pass, gated evaluation and
// the stop decision.  Every launch reads its per-round values from the
// device round state; no memset or host copy is recorded.
hipError_t capture_round(felog_ctx *c, bool visits) {
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed);
    if (e != hipSuccess) return e;
    e = round_next(c);
    if (e == hipSuccess) e = walk_enqueue(c);
    if (e == hipSuccess) e = compact_enqueue(c);
    if (e == hipSuccess && visits) e = add_round_visits(c);
    if (e == hipSuccess)
        e = c->opt.trainer != FELOG_TRAIN_HOGBATCH ? w2v_enqueue(c, c->d_off, c->d_tok)
                                                   : sgns_enqueue(c, c->d_off, c->d_tok, nullptr, nullptr);
    if (e == hipSuccess) e = eval_enqueue(c, true);
    if (e == hipSuccess && c->budget_on && c->budget_recheck) e = recheck_enqueue(c, true);
    if (e == hipSuccess) e = round_end(c);
    const hipError_t e2 = hipStreamEndCapture(c->stream, &g);
    if (e == hipSuccess) e = e2;
    if (e == hipSuccess) e = hipGraphInstantiate(&c->round_exec, g, nullptr, nullptr, 0);
    if (g) (void)hipGraphDestroy(g);
    return e;
}

}  // namespace

extern "C" {

felog_status felog_run_graph(felog_ctx *c, uint32_t max_rounds) {
    felog_status s = need_built(c);
    if (s) return s;
    const bool budget = c->budget_on;
    const bool visits = c->d_visits != nullptr;
    // buffers for the largest corpus a round can write
    if (c->opt.trainer != FELOG_TRAIN_HOGBATCH) TRY(c, w2v_reserve(c, c->walkers, c->tok_cap));
    else TRY(c, sgns_reserve(c, c->walkers, c->tok_cap));
    if (c->round_exec && (c->round_exec_budget != budget || c->round_exec_visits != visits)) drop_round_graph(c);
    if (!c->round_exec) {
        TRY(c, capture_round(c, visits));
        c->round_exec_budget = budget;
        c->round_exec_visits = visits;
    }
    bool plain_stopped = budget && !c->budget_plain_left;
    const uint32_t last = budget ? last_budget_round(c) : 0;
    RoundState &h = *c->h_rs;
    std::memset(&h, 0, sizeof(h));
    h.round = c->rounds_done;
    h.stop = plain_stopped ? 1 : 0;
    h.last_eva = c->last_eva_num;
    h.eval_from = c->opt.min_R;
    h.budget = budget ? 1 : 0;
    TRY(c, hipMemcpyAsync(c->d_rs, c->h_rs, sizeof(RoundState), hipMemcpyHostToDevice, c->stream));
    const uint32_t first = c->rounds_done + 1;
    for (uint32_t r = first;; ++r) {
        if (c->opt.max_rounds && r > c->opt.max_rounds) break;
        if (max_rounds && r - first >= max_rounds) break;
        if (budget && r > last && plain_stopped) break;
        TRY(c, hipEventRecord(c->ev0, c->stream));
        TRY(c, hipGraphLaunch(c->round_exec, c->stream));
        TRY(c, hipEventRecord(c->ev1, c->stream));
        TRY(c, hipMemcpyAsync(c->h_rs, c->d_rs, sizeof(RoundState), hipMemcpyDeviceToHost, c->stream));
        TRY(c, hipMemcpyAsync(c->h_counters, c->d_counters, kNumCounters * 8, hipMemcpyDeviceToHost, c->stream));
        TRY(c, hipStreamSynchronize(c->stream));
        float ms = 0.f;
        TRY(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
        const uint64_t *k = c->h_counters;
        c->st.graph_ms_total += ms;
        c->corpus = Corpus();
        c->corpus.round = r;
        c->corpus.n_walks = h.n_active;
        c->corpus.n_tokens = h.n_tokens;
        c->corpus.n_steps = k[C_STEPS];
        c->corpus.max_len = (uint32_t)k[C_MAXLEN];
        c->corpus.valid = true;
        c->corpus.from_walk = true;
        c->st.walk_steps = k[C_STEPS];
        c->st.walk_tokens = h.n_tokens;
        c->st.walk_tokens_total += h.n_tokens;
        c->st.overflow_used = k[C_OVF_NEXT];
        c->st.train_words = k[C_WORDS];
        c->st.train_kept = k[C_KEPT];
        c->st.train_pairs = k[C_PAIRS];
        c->st.train_groups = k[C_GROUPS];
        c->st.train_rows_written = k[C_ROWS_WB];
        c->st.train_rows_lost = k[C_ROWS_LOST];
        if (visits) c->visits_total += h.n_tokens;
        if (k[C_OVF_FAIL])
            return fail(c, FELOG_ERR_CAPACITY,
                        std::to_string(k[C_OVF_FAIL]) + " walks outgrew the overflow pool; raise overflow_capacity");
        if (h.n_active + k[C_TILES] > c->tiles_cap)
            return fail(c, FELOG_ERR_INVARIANT, "SGNS tile descriptors exceeded their bound");
        // the budget loop stops before training a round without walkers
        if (budget && h.n_active == 0) break;
        c->rounds_done = r;
        if (h.evaluated) {
            c->last_eva_num = h.last_eva;
            c->n_unconverged = h.last_eva;
            c->st.n_unconverged = h.last_eva;
            c->stop = h.stop != 0;
        }
        if (budget) {
            plain_stopped = h.stop != 0;
        } else {
            if (h.stop) break;
            if (h.n_active == 0) break;
        }
    }
    return FELOG_OK;
}

felog_status felog_visits_reset(felog_ctx *c) {
    felog_status s = need_built(c);
    if (s) return s;
    if (!c->d_visits) TRY(c, hipMalloc(&c->d_visits, (uint64_t)c->V * 4));
    if (!c->d_vstats) TRY(c, hipMalloc(&c->d_vstats, 4 * sizeof(double)));
    TRY(c, hipMemsetAsync(c->d_visits, 0, (uint64_t)c->V * 4, c->stream));
    TRY(c, hipStreamSynchronize(c->stream));
    c->visits_total = 0;
    return FELOG_OK;
}

felog_status felog_track_lost_updates(felog_ctx *c, int on) {
    felog_status s = need_built(c);
    if (s) return s;
    if (!on) {
        dfree(c->d_row_ver);
        c->d_row_ver = nullptr;
    } else if (!c->d_row_ver) {
        TRY(c, hipMalloc(&c->d_row_ver, 2ull * c->V * 4));
        TRY(c, hipMemsetAsync(c->d_row_ver, 0, 2ull * c->V * 4, c->stream));
        TRY(c, hipStreamSynchronize(c->stream));
    }
    drop_round_graph(c);  // captured rounds hold the tracking pointer
    return FELOG_OK;
}

felog_status felog_visit_stats(felog_ctx *c, double out[4]) {
    felog_status s = need_built(c);
    if (s) return s;
    if (!out) return fail(c, FELOG_ERR_ARG, "null output");
    if (!c->d_visits) return fail(c, FELOG_ERR_STATE, "visits are not tracked: call felog_visits_reset first");
    TRY(c, visit_stats(c, c->d_vstats));
    double h[3] = {0, 0, 0};
    TRY(c, hipMemcpyAsync(h, c->d_vstats, 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    TRY(c, hipStreamSynchronize(c->stream));
    out[0] = h[0];
    out[1] = h[1];
    out[2] = h[2];
    out[3] = (double)c->visits_total;
    return FELOG_OK;
}

// ---- SGNS tile shape and autotuner -----------------------------------------

}  // extern "C"

namespace {

uint32_t group_centres(const felog_ctx *c) {
    return sgns_group_centres(c->opt.batch_size, c->opt.window, c->opt.negative);
}

felog_status check_shape(felog_ctx *c, const felog_tile_shape &s) {
    if (s.rows != 0 && s.rows != 32 && s.rows != 48 && s.rows != 64)
        return fail(c, FELOG_ERR_ARG, "tile rows must be 32, 48 or 64");
    if (s.rows == 48 && group_centres(c) + 2 * c->opt.window > 48)
        return fail(c, FELOG_ERR_ARG, "48-row tiles need batch centres + 2 window <= 48");
    const uint32_t lo = tile_wpe_low(c);
    if (s.waves_per_simd != 0 && s.waves_per_simd != lo && s.waves_per_simd != lo + 1)
        return fail(c, FELOG_ERR_ARG,
                    "waves_per_simd must be " + std::to_string(lo) + " or " + std::to_string(lo + 1));
    if (s.groups_per_cu > 64) return fail(c, FELOG_ERR_ARG, "groups_per_cu must be <= 64");
    if (s.windows_per_wave != 0 && s.windows_per_wave != 1 && s.windows_per_wave != 2 && s.windows_per_wave != 4 &&
        s.windows_per_wave != 8)
        return fail(c, FELOG_ERR_ARG, "windows_per_wave must be 1, 2, 4 or 8");
    if (s.dim_split != 0 && s.dim_split != c->dp && 2 * s.dim_split != c->dp && 4 * s.dim_split != c->dp)
        return fail(c, FELOG_ERR_ARG, "dim_split must be " + std::to_string(c->dp) + ", " +
                                          std::to_string(c->dp / 2) + " or " + std::to_string(c->dp / 4) +
                                          " (d rounded up to 32, divided by 1, 2 or 4)");
    return FELOG_OK;
}

// column parts of a dim split (0: the default)
uint32_t split_parts(const felog_ctx *c, uint32_t dim_split) {
    return dim_split ? c->dp / dim_split : 0u;
}

void apply_shape(felog_ctx *c, uint32_t rows, uint32_t wpe, uint32_t gpc, uint32_t wpw, uint32_t parts) {
    if (rows != c->tile_rows || wpe != c->tile_wpe || gpc != c->tile_gpc || wpw != c->tile_wpw ||
        parts != c->tile_parts)
        drop_round_graph(c);
    c->tile_rows = rows;
    c->tile_wpe = wpe;
    c->tile_gpc = gpc;
    c->tile_wpw = wpw;
    c->tile_parts = parts;
}

}  // namespace

extern "C" {

felog_status felog_tile_shape_set(felog_ctx *c, const felog_tile_shape *shape) {
    if (!c) return FELOG_ERR_ARG;
    const felog_tile_shape zero{};
    const felog_tile_shape &s = shape ? *shape : zero;
    felog_status st = check_shape(c, s);
    if (st) return st;
    apply_shape(c, s.rows, s.waves_per_simd, s.groups_per_cu, s.windows_per_wave, split_parts(c, s.dim_split));
    return FELOG_OK;
}

felog_status felog_tile_shape_get(felog_ctx *c, felog_tile_shape *out) {
    if (!c || !out) return FELOG_ERR_ARG;
    TRY(c, hipSetDevice(c->device));
    out->rows = c->tile_rows ? c->tile_rows : 32u;  // default: the 32-row split
    out->waves_per_simd = c->tile_wpe ? c->tile_wpe : tile_wpe_low(c) + 1;
    int per_cu = 0;
    TRY(c, tile_occupancy(c, out->rows, out->waves_per_simd, &per_cu));
    out->groups_per_cu = c->tile_gpc && (int)c->tile_gpc < per_cu ? c->tile_gpc : (uint32_t)per_cu;
    out->windows_per_wave = c->tile_wpw ? c->tile_wpw : 8u;
    out->dim_split = c->dp / (c->tile_parts ? c->tile_parts : 1u);
    out->tile_ms = 0.f;
    return FELOG_OK;
}

felog_status felog_tune_sgns(felog_ctx *c, uint64_t sample_sentences, uint32_t reps, felog_tile_shape *results,
                             uint32_t max_results, uint32_t *n_results) {
    felog_status s = need_built(c);
    if (s) return s;
    if (!c->corpus.valid || !c->corpus.from_walk)
        return fail(c, FELOG_ERR_STATE, "no walk corpus: call felog_walk_round first");
    if (c->opt.trainer != FELOG_TRAIN_HOGBATCH) return fail(c, FELOG_ERR_STATE, "the tile trainer is not selected");
    if (reps == 0) reps = 1;
    uint64_t n = c->corpus.n_walks;
    if (sample_sentences && sample_sentences < n) n = sample_sentences;
    if (n == 0) return fail(c, FELOG_ERR_STATE, "the walk corpus is empty");
    uint64_t nt = 0;
    TRY(c, hipMemcpy(&nt, c->d_off + n, 8, hipMemcpyDeviceToHost));
    // every shape: rows x register budget x workgroups per CU up to what fits
    std::vector<felog_tile_shape> cand;
    const uint32_t lo = tile_wpe_low(c);
    for (uint32_t rows : {32u, 48u, 64u}) {
        if (rows == 48 && group_centres(c) + 2 * c->opt.window > 48) continue;
        for (uint32_t wpe : {lo, lo + 1}) {
            int per_cu = 0;
            TRY(c, tile_occupancy(c, rows, wpe, &per_cu));
            for (int g = 1; g <= per_cu; ++g)
                cand.push_back(felog_tile_shape{rows, wpe, (uint32_t)g, 0u, 0u, 0.f});
        }
    }
    // both tables (one allocation) are restored before every pass
    const size_t tbytes = 2 * (size_t)c->V * c->dp * (c->opt.precision == FELOG_BF16 ? 2 : 4);
    void *snap = nullptr;
    TRY(c, hipMalloc(&snap, tbytes));
    const uint32_t keep_rows = c->tile_rows, keep_wpe = c->tile_wpe, keep_gpc = c->tile_gpc,
                   keep_wpw = c->tile_wpw, keep_parts = c->tile_parts;
    hipError_t e = hipMemcpyAsync(snap, c->d_syn0, tbytes, hipMemcpyDeviceToDevice, c->stream);
    auto time_shapes = [&](size_t from) {
        for (size_t i = from; i < cand.size() && e == hipSuccess; ++i) {
            apply_shape(c, cand[i].rows, cand[i].waves_per_simd, cand[i].groups_per_cu, cand[i].windows_per_wave,
                        split_parts(c, cand[i].dim_split));
            double sum = 0.0;
            for (uint32_t r = 0; r <= reps && e == hipSuccess; ++r) {
                e = hipMemcpyAsync(c->d_syn0, snap, tbytes, hipMemcpyDeviceToDevice, c->stream);
                if (e == hipSuccess) e = round_set(c, c->corpus.round, -1.f, (uint32_t)n);
                float ms[3] = {0.f, 0.f, 0.f};
                if (e == hipSuccess) e = run_sgns(c, c->d_off, c->d_tok, n, nt, nullptr, ms);
                if (r > 0) sum += ms[2];
            }
            cand[i].tile_ms = (float)(sum / reps);
        }
    };
    auto fastest = [&]() {
        size_t best = 0;
        for (size_t i = 1; i < cand.size(); ++i)
            if (cand[i].tile_ms < cand[best].tile_ms) best = i;
        return best;
    };
    // rows x register budget x workgroups per CU, then windows per
    // wavefront x dim split at the fastest of those
    time_shapes(0);
    const size_t first = cand.size();
    if (e == hipSuccess) {
        const felog_tile_shape b = cand[fastest()];
        for (uint32_t wpw : {1u, 2u, 4u, 8u})
            for (uint32_t parts : {4u, 2u, 1u})
                if (!(wpw == 8u && parts == 1u))  // the default, timed above
                    cand.push_back(felog_tile_shape{b.rows, b.waves_per_simd, b.groups_per_cu, wpw, c->dp / parts,
                                                    0.f});
        time_shapes(first);
    }
    const hipError_t e2 = hipMemcpyAsync(c->d_syn0, snap, tbytes, hipMemcpyDeviceToDevice, c->stream);
    const hipError_t e3 = hipStreamSynchronize(c->stream);
    dfree(snap);
    if (e == hipSuccess) e = e2;
    if (e == hipSuccess) e = e3;
    if (e != hipSuccess) {
        apply_shape(c, keep_rows, keep_wpe, keep_gpc, keep_wpw, keep_parts);
        return hip_fail(c, e, "felog_tune_sgns");
    }
    // shapes of the first sweep report the defaults they ran with
    for (size_t i = 0; i < first; ++i) {
        cand[i].windows_per_wave = 8u;
        cand[i].dim_split = c->dp;
    }
    const size_t best = fastest();
    apply_shape(c, cand[best].rows, cand[best].waves_per_simd, cand[best].groups_per_cu,
                cand[best].windows_per_wave, split_parts(c, cand[best].dim_split));
    const uint32_t m = (uint32_t)(cand.size() < max_results ? cand.size() : max_results);
    if (results)
        for (uint32_t i = 0; i < m; ++i) results[i] = cand[i];
    if (n_results) *n_results = (uint32_t)cand.size();
    return FELOG_OK;
}

felog_status felog_corpus_get(felog_ctx *c, felog_corpus *out) {
    felog_status s = need_built(c);
    if (s) return s;
    if (!out) return fail(c, FELOG_ERR_ARG, "null output");
    if (!c->corpus.valid) return fail(c, FELOG_ERR_STATE, "no corpus");
    out->round = c->corpus.round;
    out->max_len = c->corpus.max_len;
    out->n_walks = c->corpus.n_walks;
    out->n_tokens = c->corpus.n_tokens;
    out->n_steps = c->corpus.n_steps;
    return FELOG_OK;
}

felog_status felog_corpus_copy_out(felog_ctx *c, uint64_t *offsets, uint32_t *tokens, uint32_t *walker_ids) {
    felog_status s = need_built(c);
    if (s) return s;
    if (!c->corpus.valid) return fail(c, FELOG_ERR_STATE, "no corpus");
    const uint64_t n = c->corpus.n_walks;
    if (offsets) TRY(c, hipMemcpy(offsets, c->d_off, (n + 1) * 8, hipMemcpyDeviceToHost));
    if (tokens && c->corpus.n_tokens) TRY(c, hipMemcpy(tokens, c->d_tok, c->corpus.n_tokens * 4, hipMemcpyDeviceToHost));
    if (walker_ids && n) TRY(c, hipMemcpy(walker_ids, c->d_sent_id, n * 4, hipMemcpyDeviceToHost));
    return FELOG_OK;
}

felog_status felog_csr_copy_out(felog_ctx *c, uint64_t *row_ptr, uint32_t *col, float *w) {
    felog_status s = need_built(c);
    if (s) return s;
    if (row_ptr) TRY(c, hipMemcpy(row_ptr, c->d_row_ptr, ((uint64_t)c->V + 1) * 8, hipMemcpyDeviceToHost));
    if (col && c->E) TRY(c, hipMemcpy(col, c->d_col, c->E * 4, hipMemcpyDeviceToHost));
    if (w && c->E) TRY(c, hipMemcpy(w, c->d_w, c->E * 4, hipMemcpyDeviceToHost));
    return FELOG_OK;
}

felog_status felog_alias_copy_out(felog_ctx *c, uint32_t *thresh, uint32_t *alias_idx) {
    felog_status s = need_built(c);
    if (s) return s;
    if (!c->E) return FELOG_OK;
    std::vector<EdgeRec> rec(c->E);
    TRY(c, hipMemcpy(rec.data(), c->d_erec, c->E * sizeof(EdgeRec), hipMemcpyDeviceToHost));
    for (uint64_t i = 0; i < c->E; ++i) {
        if (thresh) thresh[i] = rec[i].thresh;
        if (alias_idx) alias_idx[i] = rec[i].alias_idx;
    }
    return FELOG_OK;
}

felog_status felog_vocab_copy_out(felog_ctx *c, uint32_t *order, uint32_t *degree) {
    felog_status s = need_built(c);
    if (s) return s;
    if (order) TRY(c, hipMemcpy(order, c->d_vocab_order, (uint64_t)c->V * 4, hipMemcpyDeviceToHost));
    if (degree) TRY(c, hipMemcpy(degree, c->d_deg, (uint64_t)c->V * 4, hipMemcpyDeviceToHost));
    return FELOG_OK;
}

felog_status felog_unigram_copy_out(felog_ctx *c, uint32_t *thresh, uint32_t *alias_idx) {
    felog_status s = need_built(c);
    if (s) return s;
    if (thresh || alias_idx) {
        std::vector<uint2> uni(c->V);
        TRY(c, hipMemcpy(uni.data(), c->d_uni, (uint64_t)c->V * 8, hipMemcpyDeviceToHost));
        for (uint32_t v = 0; v < c->V; ++v) {
            if (thresh) thresh[v] = uni[v].x;
            if (alias_idx) alias_idx[v] = uni[v].y;
        }
    }
    return FELOG_OK;
}

felog_status felog_flags_copy_out(felog_ctx *c, uint8_t *flags, float *scores) {
    felog_status s = need_built(c);
    if (s) return s;
    if (flags) TRY(c, hipMemcpy(flags, c->d_flags, c->V, hipMemcpyDeviceToHost));
    if (scores) TRY(c, hipMemcpy(scores, c->d_scores, (uint64_t)c->V * 4, hipMemcpyDeviceToHost));
    return FELOG_OK;
}

felog_status felog_w2v_table_copy_out(felog_ctx *c, int32_t *table) {
    felog_status s = need_built(c);
    if (s) return s;
    if (!table) return fail(c, FELOG_ERR_ARG, "null table");
    TRY(c, w2v_prepare(c));
    TRY(c, hipMemcpy(table, c->d_w2v_table, (uint64_t)c->w2v_table_size * 4, hipMemcpyDeviceToHost));
    return FELOG_OK;
}

static felog_status copy_table_out(felog_ctx *c, const void *dtab, float *out) {
    const uint64_t V = c->V, d = c->d, dp = c->dp;
    if (c->opt.precision == FELOG_FP32) {
        TRY(c, hipMemcpy2D(out, d * 4, dtab, dp * 4, d * 4, V, hipMemcpyDeviceToHost));
    } else {
        std::vector<uint16_t> tmp(V * d);
        TRY(c, hipMemcpy2D(tmp.data(), d * 2, dtab, dp * 2, d * 2, V, hipMemcpyDeviceToHost));
        for (uint64_t i = 0; i < V * d; ++i) {
            const uint32_t b = (uint32_t)tmp[i] << 16;
            std::memcpy(&out[i], &b, 4);
        }
    }
    return FELOG_OK;
}

static felog_status copy_table_in(felog_ctx *c, void *dtab, const float *in) {
    const uint64_t V = c->V, d = c->d, dp = c->dp;
    if (c->opt.precision == FELOG_FP32) {
        TRY(c, hipMemset(dtab, 0, V * dp * 4));
        TRY(c, hipMemcpy2D(dtab, dp * 4, in, d * 4, d * 4, V, hipMemcpyHostToDevice));
    } else {
        std::vector<uint16_t> tmp(V * dp, 0);
        for (uint64_t v = 0; v < V; ++v)
            for (uint64_t j = 0; j < d; ++j) {
                uint32_t b;
                std::memcpy(&b, &in[v * d + j], 4);
                b += 0x7FFFu + ((b >> 16) & 1u);  // round to nearest even
                tmp[v * dp + j] = (uint16_t)(b >> 16);
            }
        TRY(c, hipMemcpy(dtab, tmp.data(), V * dp * 2, hipMemcpyHostToDevice));
    }
    return FELOG_OK;
}

felog_status felog_embeddings_copy_out(felog_ctx *c, float *syn0, float *syn1neg) {
    felog_status s = need_built(c);
    if (s) return s;
    if (syn0 && (s = copy_table_out(c, c->d_syn0, syn0))) return s;
    if (syn1neg && (s = copy_table_out(c, c->d_syn1, syn1neg))) return s;
    return FELOG_OK;
}

felog_status felog_embeddings_copy_in(felog_ctx *c, const float *syn0, const float *syn1neg) {
    felog_status s = need_built(c);
    if (s) return s;
    if (syn0 && (s = copy_table_in(c, c->d_syn0, syn0))) return s;
    if (syn1neg && (s = copy_table_in(c, c->d_syn1, syn1neg))) return s;
    return FELOG_OK;
}

// Walk file of include/path.hpp:87-97 (the reference's writer is disabled):
// <prefix>-<rank>-<round>.txt, one line per non-empty walk, each vertex as
// " %u", in walker order.
felog_status felog_write_walks(felog_ctx *c, const char *prefix) {
    felog_status s = need_built(c);
    if (s) return s;
    if (!prefix) return fail(c, FELOG_ERR_ARG, "null prefix");
    if (!c->corpus.valid || !c->corpus.from_walk) return fail(c, FELOG_ERR_STATE, "no walk corpus");
    const uint64_t n = c->corpus.n_walks;
    std::vector<uint64_t> off(n + 1);
    std::vector<uint32_t> tok(c->corpus.n_tokens);
    if ((s = felog_corpus_copy_out(c, off.data(), tok.data(), nullptr))) return s;
    const std::string path = std::string(prefix) + "-0-" + std::to_string(c->corpus.round) + ".txt";
    FILE *f = std::fopen(path.c_str(), "w");
    if (!f) return fail(c, FELOG_ERR_IO, "cannot open " + path);
    for (uint64_t i = 0; i < n; ++i) {
        if (off[i + 1] == off[i]) continue;
        for (uint64_t k = off[i]; k < off[i + 1]; ++k) std::fprintf(f, " %u", tok[k]);
        std::fputc('\n', f);
    }
    if (std::fclose(f) != 0) return fail(c, FELOG_ERR_IO, "write failed: " + path);
    return FELOG_OK;
}

// Embedding file of src/word2vec.cu:2117-2127: "V d\n", then one row per
// vocabulary entry in vocabulary order: "<vertex id> " followed by d values
// as "%lf " (or d raw f32 with binary), then "\n".
felog_status felog_write_embeddings(felog_ctx *c, const char *path, int binary) {
    felog_status s = need_built(c);
    if (s) return s;
    if (!path) return fail(c, FELOG_ERR_ARG, "null path");
    const uint64_t V = c->V, d = c->d;
    std::vector<float> syn0(V * d);
    if ((s = felog_embeddings_copy_out(c, syn0.data(), nullptr))) return s;
    std::vector<uint32_t> order(V), deg(V);
    if ((s = felog_vocab_copy_out(c, order.data(), deg.data()))) return s;
    FILE *f = std::fopen(path, "wb");
    if (!f) return fail(c, FELOG_ERR_IO, std::string("cannot open ") + path);
    std::fprintf(f, "%lld %lld\n", (long long)c->vocab_size, (long long)d);
    for (uint64_t i = 0; i < V; ++i) {
        const uint32_t v = order[i];
        if (!(deg[v] >= c->opt.min_count || i == 0)) continue;
        std::fprintf(f, "%u ", v);
        if (binary) std::fwrite(&syn0[(uint64_t)v * d], 4, d, f);
        else
            for (uint64_t j = 0; j < d; ++j) std::fprintf(f, "%lf ", (double)syn0[(uint64_t)v * d + j]);
        std::fputc('\n', f);
    }
    if (std::fclose(f) != 0) return fail(c, FELOG_ERR_IO, std::string("write failed: ") + path);
    return FELOG_OK;
}

// ---- the tools either side of the path --------------------------------------

}  // extern "C"

namespace {
// Whitespace-separated edge records "src dst" or "src dst data" read as
// fscanf("%u %u") / fscanf("%u %u %f") would (include/storage.hpp:156-164):
// records are taken until the first one that does not parse.  With src
// null only the records are counted.
felog_status parse_edges_txt(felog_ctx *c, const char *path, int columns, uint32_t *src, uint32_t *dst, float *w,
                             uint64_t cap, uint64_t *m) {
    if (!path || !m || (columns != 2 && columns != 3)) return fail(c, FELOG_ERR_ARG, "bad arguments");
    FILE *f = std::fopen(path, "rb");
    if (!f) return fail(c, FELOG_ERR_IO, std::string("cannot open ") + path);
    std::string buf;
    char chunk[1 << 16];
    size_t got;
    while ((got = std::fread(chunk, 1, sizeof chunk, f)) > 0) buf.append(chunk, got);
    std::fclose(f);
    const char *p = buf.c_str();
    uint64_t n = 0;
    while (true) {
        char *end = nullptr;
        uint32_t rec[2];
        bool ok = true;
        for (int k = 0; k < 2 && ok; ++k) {
            while (*p && std::isspace((unsigned char)*p)) ++p;
            if (!*p || *p == '-' || !std::isdigit((unsigned char)*p)) { ok = false; break; }
            const unsigned long x = std::strtoul(p, &end, 10);
            if (end == p || x > 0xFFFFFFFFul) { ok = false; break; }
            rec[k] = (uint32_t)x;
            p = end;
        }
        float data = 1.f;
        if (ok && columns == 3) {
            const float x = std::strtof(p, &end);
            if (end == p) ok = false;
            else {
                data = x;
                p = end;
            }
        }
        if (!ok) break;
        if (src) {
            if (n >= cap) return fail(c, FELOG_ERR_CAPACITY, "more edges than the buffers hold");
            src[n] = rec[0];
            dst[n] = rec[1];
            if (w) w[n] = data;
        }
        ++n;
    }
    *m = n;
    return FELOG_OK;
}
}  // namespace

extern "C" {

felog_status felog_read_edges_txt(const char *path, int columns, uint32_t *src, uint32_t *dst, float *w,
                                  uint64_t capacity, uint64_t *m) {
    felog_ctx scratch_err;  // parse errors have no context to report to
    return parse_edges_txt(&scratch_err, path, columns, src, dst, w, capacity, m);
}

felog_status felog_load_edges_txt(felog_ctx *c, const char *path, uint32_t v_num, int make_undirected, int weighted) {
    if (!c) return FELOG_ERR_ARG;
    const int cols = weighted ? 3 : 2;
    uint64_t m = 0;
    felog_status s = parse_edges_txt(c, path, cols, nullptr, nullptr, nullptr, 0, &m);
    if (s) return s;
    std::vector<uint32_t> src(m), dst(m);
    std::vector<float> w(m, 1.f);
    if ((s = parse_edges_txt(c, path, cols, src.data(), dst.data(), weighted ? w.data() : nullptr, m, &m))) return s;
    return felog_set_edges(c, src.data(), dst.data(), w.data(), m, v_num, make_undirected);
}

// include/storage.hpp:207-240: "%u %u\n" (no data) or "%u %u %f\n" per edge.
felog_status felog_write_edges_txt(const char *path, const uint32_t *src, const uint32_t *dst, const float *w,
                                   uint64_t m) {
    if (!path || (m && (!src || !dst))) return FELOG_ERR_ARG;
    FILE *f = std::fopen(path, "w");
    if (!f) return FELOG_ERR_IO;
    for (uint64_t i = 0; i < m; ++i) {
        if (w) std::fprintf(f, "%u %u %f\n", src[i], dst[i], (double)w[i]);
        else std::fprintf(f, "%u %u\n", src[i], dst[i]);
    }
    return std::fclose(f) == 0 ? FELOG_OK : FELOG_ERR_IO;
}

// include/storage.hpp:107-115 (write_graph): packed little-endian records,
// {u32 src, u32 dst, f32 data} (12 B) or without data {u32, u32} (8 B).
felog_status felog_write_edges_bin(const char *path, const uint32_t *src, const uint32_t *dst, const float *w,
                                   uint64_t m) {
    if (!path || (m && (!src || !dst))) return FELOG_ERR_ARG;
    FILE *f = std::fopen(path, "wb");
    if (!f) return FELOG_ERR_IO;
    std::vector<unsigned char> rec;
    const size_t rs = w ? 12 : 8;
    const uint64_t step = 1 << 20;
    for (uint64_t i0 = 0; i0 < m; i0 += step) {
        const uint64_t k = m - i0 < step ? m - i0 : step;
        rec.resize(k * rs);
        for (uint64_t i = 0; i < k; ++i) {
            std::memcpy(&rec[i * rs], &src[i0 + i], 4);
            std::memcpy(&rec[i * rs + 4], &dst[i0 + i], 4);
            if (w) std::memcpy(&rec[i * rs + 8], &w[i0 + i], 4);
        }
        if (std::fwrite(rec.data(), 1, rec.size(), f) != rec.size()) {
            std::fclose(f);
            return FELOG_ERR_IO;
        }
    }
    return std::fclose(f) == 0 ? FELOG_OK : FELOG_ERR_IO;
}

felog_status felog_common_neighbours(felog_ctx *c, const uint32_t *src, const uint32_t *dst, uint64_t m,
                                     int32_t *cn) {
    felog_status s = need_built(c);
    if (s) return s;
    if (m == 0) return FELOG_OK;
    if (!src || !dst || !cn) return fail(c, FELOG_ERR_ARG, "null pair or output array");
    for (uint64_t i = 0; i < m; ++i)
        if (src[i] >= c->V || dst[i] >= c->V) return fail(c, FELOG_ERR_ARG, "vertex id >= v_num");
    uint32_t *du = nullptr, *dv = nullptr;
    int32_t *dout = nullptr;
    TRY(c, hipMalloc(&du, m * 4));
    TRY(c, hipMalloc(&dv, m * 4));
    TRY(c, hipMalloc(&dout, m * 4));
    hipError_t e = hipMemcpy(du, src, m * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dv, dst, m * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = common_neighbours_pairs(c, du, dv, m, dout);
    if (e == hipSuccess) e = hipMemcpy(cn, dout, m * 4, hipMemcpyDeviceToHost);
    dfree(du);
    dfree(dv);
    dfree(dout);
    TRY(c, e);
    return FELOG_OK;
}

felog_status felog_use_common_neighbours(felog_ctx *c, int on) {
    felog_status s = need_built(c);
    if (s) return s;
    drop_round_graph(c);  // the captured walk holds the edge-data pointer
    if (!on) {
        dfree(c->d_cn);
        c->d_cn = nullptr;
        return FELOG_OK;
    }
    TRY(c, common_neighbours_csr(c));
    return FELOG_OK;
}

felog_status felog_cn_copy_out(felog_ctx *c, float *cn) {
    felog_status s = need_built(c);
    if (s) return s;
    if (!cn) return fail(c, FELOG_ERR_ARG, "null output");
    if (!c->d_cn) return fail(c, FELOG_ERR_STATE, "no common-neighbour counts: call felog_use_common_neighbours");
    if (c->E) TRY(c, hipMemcpy(cn, c->d_cn, c->E * 4, hipMemcpyDeviceToHost));
    return FELOG_OK;
}

felog_status felog_link_auc(felog_ctx *c, const uint32_t *pairs, const uint8_t *labels, uint64_t n, float *scores,
                            double *auc) {
    felog_status s = need_built(c);
    if (s) return s;
    if (!pairs || !labels || !auc || n == 0) return fail(c, FELOG_ERR_ARG, "pairs, labels and auc are required");
    for (uint64_t i = 0; i < 2 * n; ++i)
        if (pairs[i] >= c->V) return fail(c, FELOG_ERR_ARG, "vertex id >= v_num");
    uint32_t *dp = nullptr;
    uint8_t *dl = nullptr;
    float *ds = nullptr;
    TRY(c, hipMalloc(&dp, n * 8));
    TRY(c, hipMalloc(&dl, n));
    TRY(c, hipMalloc(&ds, n * 4));
    std::vector<uint8_t> lab(labels, labels + n);
    for (auto &x : lab) x = x ? 1 : 0;
    hipError_t e = hipMemcpy(dp, pairs, n * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dl, lab.data(), n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = link_auc(c, dp, dl, n, ds, auc);
    if (e == hipSuccess && scores) e = hipMemcpy(scores, ds, n * 4, hipMemcpyDeviceToHost);
    dfree(dp);
    dfree(dl);
    dfree(ds);
    TRY(c, e);
    return FELOG_OK;
}

felog_status felog_kmeans_classes(felog_ctx *c, uint32_t classes, uint32_t iters, int32_t *class_of) {
    felog_status s = need_built(c);
    if (s) return s;
    if (!class_of || classes == 0) return fail(c, FELOG_ERR_ARG, "classes must be > 0 and class_of non-null");
    if (c->opt.precision != FELOG_FP32) return fail(c, FELOG_ERR_STATE, "k-means runs on FP32 tables");
    int32_t *d_cl = nullptr;
    TRY(c, hipMalloc(&d_cl, (c->vocab_size ? c->vocab_size : 1) * 4));
    hipError_t e = kmeans_classes(c, classes, iters, d_cl);
    if (e == hipSuccess) e = hipMemcpy(class_of, d_cl, c->vocab_size * 4, hipMemcpyDeviceToHost);
    dfree(d_cl);
    TRY(c, e);
    return FELOG_OK;
}

// The classes file of src/word2vec.cu:2172: "<vertex id> <class>\n" per
// vocabulary entry, in vocabulary order, after 10 k-means rounds.
felog_status felog_write_classes(felog_ctx *c, const char *path, uint32_t classes) {
    felog_status s = need_built(c);
    if (s) return s;
    if (!path) return fail(c, FELOG_ERR_ARG, "null path");
    std::vector<int32_t> cl(c->vocab_size);
    if ((s = felog_kmeans_classes(c, classes, 10, cl.data()))) return s;
    std::vector<uint32_t> order(c->V);
    TRY(c, hipMemcpy(order.data(), c->d_vocab_order, (uint64_t)c->V * 4, hipMemcpyDeviceToHost));
    FILE *f = std::fopen(path, "wb");
    if (!f) return fail(c, FELOG_ERR_IO, std::string("cannot open ") + path);
    for (uint64_t a = 0; a < c->vocab_size; ++a) std::fprintf(f, "%u %d\n", order[a], cl[a]);
    if (std::fclose(f) != 0) return fail(c, FELOG_ERR_IO, std::string("write failed: ") + path);
    return FELOG_OK;
}

felog_status felog_stats_get(felog_ctx *c, felog_stats *out) {
    if (!c || !out) return FELOG_ERR_ARG;
    *out = c->st;
    out->rounds_done = c->rounds_done;
    return FELOG_OK;
}

}  // extern "C"
