/*
 * felog.h - C ABI of libfelog_hip.so, the MI355X walk -> SGNS hot path of
 * FeLoG (information-centric graph embedding).
 *
 * One context owns one GPU's copy of the graph (CSR + alias tables), the two
 * embedding tables, the per-vertex convergence flags and the last round's walk
 * corpus.  All device memory is owned by the context; every pointer argument
 * below is a HOST pointer that is borrowed for the duration of the call only.
 *
 * Conventions
 *   - every entry point returns felog_status (0 = OK, < 0 = error class);
 *     felog_last_error() returns a message for the last failure on a context.
 *     Nothing in the library calls exit() or assert().
 *   - calls on one context must be serialised by the caller (one owning host
 *     thread); distinct contexts are independent.
 *   - vertex ids are u32, edge counts u64.  The directed edge count (after
 *     undirecting) must stay below 2^32 in this build.
 *
 * Reference interfaces this boundary replaces (paths relative to the
 * reference repository):
 *   - felog argv surface: include/option_helper.hpp:94-188 and
 *     src/word2vec.cu:2269-2286 (mirrored field by field in felog_options)
 *   - GraphEngine::load_graph: include/graph.hpp:712-919
 *   - WalkEngine::random_walk: include/walk.hpp:529-542, 630-840, 848-1122
 *   - train_corpus_cuda: src/word2vec.cu:2193 (declared at
 *     src/examples/felog.cpp:26), TrainModel src/word2vec.cu:1922-2180
 */
#ifndef FELOG_H
#define FELOG_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t felog_status;

enum {
    FELOG_OK = 0,
    FELOG_ERR_ARG = -1,        /* bad argument or option value            */
    FELOG_ERR_IO = -2,         /* file missing, short or malformed         */
    FELOG_ERR_HIP = -3,        /* a HIP runtime call failed                */
    FELOG_ERR_OOM = -4,        /* device allocation failed                 */
    FELOG_ERR_INVARIANT = -5,  /* an internal consistency check failed     */
    FELOG_ERR_STATE = -6,      /* call made in the wrong order             */
    FELOG_ERR_CAPACITY = -7    /* a bounded device buffer overflowed       */
};

/* Precision of the embedding tables. */
enum { FELOG_FP32 = 0, FELOG_BF16 = 1 };

/*
 * Parsed felog options.  felog_default_options() fills the reference
 * defaults; zero-valued fields that the reference treats as "use default"
 * keep that meaning.
 */
typedef struct felog_options {
    /* graph / walk flags (include/option_helper.hpp:94-188) */
    uint32_t v_num;            /* -v                                         */
    uint32_t walkers;          /* -w  (0 -> v_num; one walker per vertex)    */
    uint32_t min_L;            /* --min_L (0 -> 20, include/walk.hpp:511)    */
    uint32_t min_R;            /* --min_R (0 -> 5,  include/walk.hpp:506)    */
    uint32_t eval_batch;       /* --batch_size (16384): evaluation chunk     */
    int32_t make_undirected;   /* --make-undirected                          */
    /* word2vec flags (src/word2vec.cu:2269-2286) */
    uint32_t size;             /* -size   (100)                              */
    uint32_t window;           /* -window (5)                                */
    uint32_t negative;         /* -negative (5)                              */
    uint32_t batch_size;       /* -batch-size: centres sharing one negative
                                  set (0 -> 2*window+1, include/dsgl.hpp:832) */
    float sample;              /* -sample (1e-3)                             */
    float alpha;               /* -alpha: base of the per-round LR (0.0025;
                                  word2vec's default is 0.025)               */
    uint32_t min_count;        /* -min-count (5)                             */
    int32_t binary;            /* -binary (0)                                */
    /* build switches */
    int32_t precision;         /* FELOG_FP32 or FELOG_BF16                   */
    uint32_t max_walk_len;     /* hard walk cap in vertices (1000)           */
    uint64_t overflow_capacity;/* vertices of walk storage beyond the fixed
                                  per-walker slot (0 -> 1<<24)               */
    uint64_t seed;             /* Philox key                                 */
    int32_t compat_lr;         /* 1: ExponentialDecayLR(0.025,0.1) ignores
                                  -alpha (src/word2vec.cu:1958,2001)         */
    int32_t compat_vocab_break;/* 1: vocab index 0 and min-count-dropped
                                  vertices end a sentence
                                  (src/word2vec.cu:1202-1205)                */
    uint32_t max_rounds;       /* felog_run() round limit (0 -> no limit)    */
    int32_t device;            /* HIP device ordinal                         */
    int32_t update_mode;       /* SGNS row updates: FELOG_UPDATE_*           */
    int32_t compat_walk;       /* 1: the reference's transition - uniform
                                  neighbour, accepted with tanh of the degree
                                  ratio over (deg - common neighbours), else a
                                  step there and back; common neighbours =
                                  (int) edge data (include/walk.hpp:976-1019).
                                  0: weight-proportional alias step          */
    int32_t trainer;           /* FELOG_TRAIN_* (below)                      */
    uint32_t w2v_table_size;   /* unigram table entries of the per-pair
                                  trainers (0 -> 1e8, src/word2vec.cu:171)   */
} felog_options;

/* The trainer of felog_train_round / felog_train_corpus / the captured round.
 * HOGBATCH (the default): the MFMA tile kernel (shared negatives per group of
 *   centres, tile-summed updates; include/dsgl.hpp:832-1407 is its model),
 *   FP32 or BF16 tables; replaces skip_gram_kernel (src/word2vec.cu:442-536).
 * SKIPGRAM, SKIPGRAM_REUSE, CBOW: the reference's per-pair kernels with its
 *   update order and random streams - skip_gram_kernel, __sgNegReuse
 *   (src/word2vec.cu:368-440) and cbow_kernel (:538-646) - on FP32 tables,
 *   with the 1e8-entry unigram table and the sigmoid table. */
enum { FELOG_TRAIN_HOGBATCH = 0, FELOG_TRAIN_SKIPGRAM = 1, FELOG_TRAIN_SKIPGRAM_REUSE = 2, FELOG_TRAIN_CBOW = 3 };

/* How SGNS tiles write their row deltas (Hogwild either way).
 * WRITEBACK (the default): each tile stores staged-row + delta (bf16:
 *   stochastically rounded), like word2vec's unsynchronised CPU threads; a
 *   concurrent tile's delta to the same row can be overwritten, which bounds
 *   the step a hub row takes.
 * ATOMIC: f32 tables get hardware float atomic adds, so no delta is lost
 *   (as src/word2vec.cu:530,533); bf16 tables a compare-and-swap per pair of
 *   elements with stochastic rounding.  Every concurrent tile's delta, each
 *   computed from the same stale row, lands: on degree-skewed graphs hub rows
 *   then take hundreds of steps at once and training diverges. */
enum { FELOG_UPDATE_DEFAULT = 0, FELOG_UPDATE_ATOMIC = 1, FELOG_UPDATE_WRITEBACK = 2 };

typedef struct felog_ctx felog_ctx;

/* Summary of the corpus produced by the last felog_walk_round(); the token
 * data stays on the device inside the context. */
typedef struct felog_corpus {
    uint32_t round;            /* 1-based round number                       */
    uint32_t max_len;          /* longest walk in vertices                   */
    uint64_t n_walks;          /* non-empty walks (walkers started)          */
    uint64_t n_tokens;         /* sum of walk lengths                        */
    uint64_t n_steps;          /* sum of (length - 1): vertices appended     */
} felog_corpus;

typedef struct felog_stats {
    uint64_t v_num, e_num;           /* vertices, directed CSR entries      */
    uint64_t vocab_size;             /* vertices with degree >= min_count   */
    uint32_t rounds_done;
    uint32_t n_unconverged;          /* after the last evaluation           */
    uint64_t walk_steps;             /* last round                          */
    uint64_t walk_tokens;            /* last round                          */
    uint64_t train_words;            /* last round: tokens read pre-subsample */
    uint64_t train_kept;             /* last round: tokens kept              */
    uint64_t train_pairs;            /* last round: (centre, context) pairs */
    uint64_t train_groups;           /* last round: SGNS groups             */
    uint64_t overflow_used;          /* last round: overflow vertices used  */
    double build_ms;                 /* CSR + alias + vocab + init          */
    double walk_kernel_ms;           /* last round, HIP events, walk kernel */
    double compact_kernel_ms;        /* last round, corpus compaction       */
    double train_kernel_ms;          /* last round, HIP events, SGNS kernel */
    double eval_kernel_ms;           /* last evaluation kernel              */
    double walk_ms_total, train_ms_total, eval_ms_total;  /* all rounds     */
    double sgns_prep_ms;             /* last round: SGNS tile preparation   */
    double sgns_tile_ms;             /* last round: SGNS tile kernel        */
    double graph_ms_total;           /* captured rounds (felog_run_graph),
                                        HIP events around each replay       */
    uint64_t walk_tokens_total;      /* corpus tokens of all walk rounds     */
    uint64_t train_rows_written;     /* last round, lost-update tracking on:
                                        table rows the tiles wrote back     */
    uint64_t train_rows_lost;        /* of those, writes that replaced another
                                        tile's update to the same row (the
                                        row changed after the tile read it) */
} felog_stats;

/* ---- lifetime ---------------------------------------------------------- */
void felog_default_options(felog_options *opt);
felog_status felog_create(felog_ctx **out, const felog_options *opt);
felog_status felog_destroy(felog_ctx *ctx);
const char *felog_last_error(const felog_ctx *ctx);

/* ---- graph input (include/storage.hpp:40-61, include/graph.hpp:741-772) */
/* Binary edge file: packed little-endian {u32 src, u32 dst, f32 data}
 * records, no header.  make_undirected appends each edge's reverse right
 * after it. */
felog_status felog_load_edges(felog_ctx *ctx, const char *bin_path,
                              uint32_t v_num, int make_undirected);
felog_status felog_set_edges(felog_ctx *ctx, const uint32_t *src,
                             const uint32_t *dst, const float *w, uint64_t m,
                             uint32_t v_num, int make_undirected);

/* CSR (neighbour order = input order, include/graph.hpp:531-575), alias
 * tables, vocabulary, table initialisation (src/word2vec.cu:864-902),
 * unigram sampler and subsampling thresholds. */
felog_status felog_build(felog_ctx *ctx);

/* ---- the hot path ------------------------------------------------------ */
/* One round of walks: one walker per unconverged vertex
 * (include/walk.hpp:547-610, 848-1122). */
felog_status felog_walk_round(felog_ctx *ctx, uint32_t round,
                              felog_corpus *out);
/* Skip-gram negative-sampling training over the last walk corpus
 * (replaces src/word2vec.cu:1133-1273 and skip_gram_kernel :442-536). */
felog_status felog_train_round(felog_ctx *ctx, const felog_corpus *corpus);
/* Train over a caller-supplied corpus (the SyncQueue hand-off of
 * src/word2vec.cu:1990-1998): sentence s is tokens[offsets[s]:offsets[s+1]]. */
felog_status felog_train_corpus(felog_ctx *ctx, const uint64_t *offsets,
                                const uint32_t *tokens, uint64_t n_sentences,
                                uint32_t round);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): felog_train_corpus with the staged corpus and the SGNS
 * tile descriptors in a static 64 MiB device array of the library; the call
 * makes no device allocation.  A corpus that needs more (8 B per sentence +
 * 4 B per token + 640 B per tile, each part 256-byte aligned) returns
 * FELOG_ERR_CAPACITY.  Every context on a GPU shares that GPU's array:
 * calls that overlap on one GPU race and must be serialised. */
felog_status felog_train_corpus_noalloc(felog_ctx *ctx, const uint64_t *offsets,
                                        const uint32_t *tokens,
                                        uint64_t n_sentences, uint32_t round);
/* Convergence evaluation (src/word2vec.cu:2016-2080): flags vertices whose
 * mean cosine to <= 30 neighbours exceeds 0.65; *stop = 1 when the
 * unconverged count stopped shrinking (ratio > 0.7). */
felog_status felog_evaluate(felog_ctx *ctx, uint32_t *n_unconverged,
                            int *stop);
/* Rounds until the stop rule fires or max_rounds is reached. */
felog_status felog_run(felog_ctx *ctx);
/* felog_run (or, with a walk budget installed, felog_run_budget) with each
 * round - walker select, walks, corpus, SGNS pass, evaluation and the stop
 * decision - captured once as a hipGraph and replayed; the round number,
 * learning rate, walker count and stop state live in device memory, and the
 * host waits once per round.  Same walks, corpus and stop decisions as the
 * eager loop (src/word2vec.cu:1955-2082 drives the same sequence).  Runs at
 * most max_rounds more rounds (0: until the stop rule or the options'
 * max_rounds); a later call continues from the rounds done. */
felog_status felog_run_graph(felog_ctx *ctx, uint32_t max_rounds);

/* ---- SGNS tile shape and its autotuner ---------------------------------- */
/* Shape of the HogBatch tile kernel (no reference counterpart; the
 * reference's HogBatch sizes its MKL GEMMs per sentence pair,
 * include/dsgl.hpp:1099-1159).  Zero fields take the defaults. */
typedef struct felog_tile_shape {
    uint32_t rows;            /* input rows of the tile kernel: 48 or 64 (48
                                 needs batch centres + 2 window <= 48), or 32:
                                 tiles of at most 32 distinct input rows run on
                                 a 32-row kernel (smaller LDS footprint, more
                                 workgroups per CU), the rest on the 48/64-row
                                 one (the default)                            */
    uint32_t waves_per_simd;  /* register budget: FP32 2 or 3, BF16 3 or 4  */
    uint32_t groups_per_cu;   /* persistent workgroups per CU (at most what
                                 LDS and registers allow; with rows 32, of
                                 the 32-row kernel)                         */
    uint32_t windows_per_wave; /* centre windows whose pairs one wavefront
                                 counts at a time: 1, 2, 4 or 8 (64, 32, 16
                                 or 8 lanes per window; default 8)          */
    uint32_t dim_split;       /* columns of a 16-row block that one wavefront
                                 updates in the dVin / dUout products: dp,
                                 dp/2 or dp/4 (dp = d rounded up to 32; d =
                                 128: 128, 64 or 32; default dp)            */
    float tile_ms;            /* felog_tune_sgns: tile-kernel time          */
} felog_tile_shape;
felog_status felog_tile_shape_set(felog_ctx *ctx, const felog_tile_shape *shape);
felog_status felog_tile_shape_get(felog_ctx *ctx, felog_tile_shape *shape);
/* Times tile shapes on the first sample_sentences sentences of the current
 * walk corpus (reps timed passes each, after one untimed), with the tables
 * restored after every pass, and installs the fastest: first every rows x
 * waves_per_simd x groups_per_cu (default windows and split), then every
 * windows_per_wave {1, 2, 4, 8} x dim_split {dp/4, dp/2, dp} at the best
 * of those.  results (up to max_results) receive every shape tried with
 * its mean time. */
felog_status felog_tune_sgns(felog_ctx *ctx, uint64_t sample_sentences,
                             uint32_t reps, felog_tile_shape *results,
                             uint32_t max_results, uint32_t *n_results);

/* ---- read-back (caller-allocated host buffers) ------------------------- */
/* Descriptor of the context's current corpus: the last walk round's, also
 * after felog_run and felog_run_graph. */
felog_status felog_corpus_get(felog_ctx *ctx, felog_corpus *out);
felog_status felog_corpus_copy_out(felog_ctx *ctx, uint64_t *offsets,
                                   uint32_t *tokens, uint32_t *walker_ids);
felog_status felog_csr_copy_out(felog_ctx *ctx, uint64_t *row_ptr,
                                uint32_t *col, float *w);
felog_status felog_alias_copy_out(felog_ctx *ctx, uint32_t *thresh,
                                  uint32_t *alias_idx);
felog_status felog_vocab_copy_out(felog_ctx *ctx, uint32_t *vocab_order,
                                  uint32_t *degree);
felog_status felog_unigram_copy_out(felog_ctx *ctx, uint32_t *thresh,
                                    uint32_t *alias_idx);
felog_status felog_flags_copy_out(felog_ctx *ctx, uint8_t *flags,
                                  float *scores);
/* The per-pair trainers' unigram table (w2v_table_size vocabulary indices,
 * src/word2vec.cu:674-694); built on first use. */
felog_status felog_w2v_table_copy_out(felog_ctx *ctx, int32_t *table);
/* Rows indexed by vertex id, widened to f32; either pointer may be NULL. */
felog_status felog_embeddings_copy_out(felog_ctx *ctx, float *syn0,
                                       float *syn1neg);
felog_status felog_embeddings_copy_in(felog_ctx *ctx, const float *syn0,
                                      const float *syn1neg);

/* ---- writers (include/path.hpp:87-97, src/word2vec.cu:2117-2127) ------- */
felog_status felog_write_walks(felog_ctx *ctx, const char *prefix);
felog_status felog_write_embeddings(felog_ctx *ctx, const char *path,
                                    int binary);
/* ---- the steps either side of the path (src/tools/, include/storage.hpp) */
/* Text edge lists (include/storage.hpp:156-164, src/tools/gconverter.cpp:
 * 57-76): whitespace-separated "src dst" (columns 2) or "src dst data"
 * (columns 3) records, read until the first record that does not parse.
 * felog_read_edges_txt needs no context: with src NULL it only counts (*m);
 * otherwise it fills up to `capacity` records (w may be NULL).
 * felog_load_edges_txt = felog_load_edges for such a file (weighted: 3
 * columns; else every weight is 1). */
felog_status felog_read_edges_txt(const char *path, int columns, uint32_t *src,
                                  uint32_t *dst, float *w, uint64_t capacity,
                                  uint64_t *m);
felog_status felog_load_edges_txt(felog_ctx *ctx, const char *path, uint32_t v_num,
                                  int make_undirected, int weighted);
/* Writers of include/storage.hpp:207-240 ("%u %u %f\n", or "%u %u\n" with w
 * NULL) and :107-115 (binary records of 12 B, or 8 B with w NULL). */
felog_status felog_write_edges_txt(const char *path, const uint32_t *src,
                                   const uint32_t *dst, const float *w, uint64_t m);
felog_status felog_write_edges_bin(const char *path, const uint32_t *src,
                                   const uint32_t *dst, const float *w, uint64_t m);
/* Common neighbours (src/tools/gcn.cpp:304-357) of m vertex pairs over the
 * built graph: the distinct vertices adjacent to both, the pair's own two
 * vertices excluded (the reference's count on graphs without self loops). */
felog_status felog_common_neighbours(felog_ctx *ctx, const uint32_t *src,
                                     const uint32_t *dst, uint64_t m, int32_t *cn);
/* on != 0: count the common neighbours of every CSR entry's two endpoints on
 * the device and let the reference's transition (compat_walk) read them in
 * place of the edge data, as the reference reads gcn's output; 0: back to
 * the edge data.  felog_cn_copy_out returns the counts in CSR order. */
felog_status felog_use_common_neighbours(felog_ctx *ctx, int on);
felog_status felog_cn_copy_out(felog_ctx *ctx, float *cn);
/* Held-out link prediction (the test files of src/tools/gtrain.cpp:369-384):
 * pairs holds n vertex pairs (2n ids), labels 1 / 0; *auc = the
 * Mann-Whitney AUC (ties at half weight) of the cosine of the pairs' syn0
 * rows; scores (may be NULL) receives the cosines. */
felog_status felog_link_auc(felog_ctx *ctx, const uint32_t *pairs,
                            const uint8_t *labels, uint64_t n, float *scores,
                            double *auc);

/* -classes (src/word2vec.cu:2128-2175): k-means of the vocabulary's syn0
 * rows into `classes` classes over `iters` rounds (the reference runs 10),
 * FP32 tables.  felog_kmeans_classes writes one class per vocabulary entry
 * (vocabulary order, vocab_size entries); felog_write_classes writes the
 * reference's "<vertex id> <class>\n" lines instead of the embeddings. */
felog_status felog_kmeans_classes(felog_ctx *ctx, uint32_t classes, uint32_t iters,
                                  int32_t *class_of);
felog_status felog_write_classes(felog_ctx *ctx, const char *path,
                                 uint32_t classes);

felog_status felog_stats_get(felog_ctx *ctx, felog_stats *out);

/* ---- learned walk budget (no counterpart in the reference) ------------- */
/* An 8-64-64-2 ReLU MLP maps per-vertex features to (ln walk count, ln walk
 * length), both standardised.  It is trained on the budgets the plain
 * information-centric rule produced on other graphs; a vertex whose
 * standardised features leave the training box [feat_lo, feat_hi] keeps the
 * plain rule. */
#define FELOG_BUDGET_FEATURES 8
#define FELOG_BUDGET_PARAMS 4866  /* W1 64x8, b1, W2 64x64, b2, W3 2x64, b3 */
typedef struct felog_budget_model {
    float params[FELOG_BUDGET_PARAMS];
    float feat_mean[FELOG_BUDGET_FEATURES], feat_std[FELOG_BUDGET_FEATURES];
    float feat_lo[FELOG_BUDGET_FEATURES], feat_hi[FELOG_BUDGET_FEATURES];
    float targ_mean[2], targ_std[2];  /* of ln(count), ln(length)        */
} felog_budget_model;

/* Raw features of every vertex (V x 8, after felog_build): ln(1+deg),
 * ln(1+sum w), max w, mean w, ln(1+sum of neighbour degrees), ln(1+max
 * neighbour degree), transition entropy / ln deg, in-vocabulary flag. */
felog_status felog_budget_features(felog_ctx *ctx, float *features);
/* Mean-squared-error gradient over n samples of standardised features X
 * (n x 8) and targets T (n x 2) at params; *loss = the mean squared error. */
felog_status felog_budget_grad(felog_ctx *ctx, const float *X, const float *T,
                               uint64_t n, const float *params, float *grad,
                               double *loss);
/* Adam training from params (updated in place): epochs x ceil(n/batch)
 * steps of batch samples drawn with replacement (batch a multiple of 64);
 * loss_per_epoch (may be NULL) receives each epoch's minibatch mean loss. */
felog_status felog_budget_train(felog_ctx *ctx, const float *X, const float *T,
                                uint64_t n, uint32_t epochs, uint32_t batch,
                                float lr, uint64_t seed, float *params,
                                double *loss_per_epoch);
/* Install a model: predicts every vertex's budget now; later walk rounds
 * start budgeted vertices in rounds 1..count with walks of exactly length
 * vertices (no entropy state) and leave the rest to the plain rule.
 * model NULL removes the budget. */
felog_status felog_budget_set(felog_ctx *ctx, const felog_budget_model *model);
/* Current budgets (0 = plain rule); any pointer may be NULL.  The counts
 * are the predicted ones as raised by the min_R recheck, if it ran;
 * *rounds receives the largest current count. */
felog_status felog_budget_copy_out(felog_ctx *ctx, uint16_t *count,
                                   uint16_t *length, uint32_t *rounds);
/* Rounds of the budgeted run: walks and training until every budgeted
 * vertex has used its count and (when plain-rule vertices remain) the plain
 * stop rule fires; evaluation runs only for plain-rule vertices (and, with
 * the recheck on, once for the budgeted vertices at min_R). */
felog_status felog_run_budget(felog_ctx *ctx);
/* No counterpart in the reference.  on != 0: at round min_R the budgeted
 * vertices whose count ends there are scored once with the plain rule's
 * convergence test (src/word2vec.cu:2016-2080, mean cosine to at most 30
 * neighbours > 0.65); each that fails walks one more round (its count
 * becomes min_R + 1).  Their flags and the plain rule's stop count are left
 * alone.  Off by default; applies to felog_run_budget and felog_run_graph. */
felog_status felog_budget_recheck(felog_ctx *ctx, int on);

/* ---- visit diagnostics (include/walk.hpp:754-788) ---------------------- */
/* Start (or restart) counting corpus visits; each later walk round adds its
 * tokens to a per-vertex histogram. */
felog_status felog_visits_reset(felog_ctx *ctx);
/* From the histogram q (visit shares) and p (degree shares):
 * out[0] = KL(p || q) = sum p (ln p - ln q) over visited vertices,
 * out[1] = Shannon entropy -sum q ln q (nats), out[2] = vertices with p > 0
 * and no visit, out[3] = visits counted. */
felog_status felog_visit_stats(felog_ctx *ctx, double out[4]);
/* Lost-update tracking of the HogBatch tile trainer's write-back (no
 * reference counterpart: word2vec's unsynchronised CPU threads, the same
 * Hogwild scheme, lose such updates uncounted).  With on = 1 every table
 * row gets a version counter (8 bytes per vertex); a tile reads a row's
 * version when it gathers the row and increments it when it writes the row
 * back, and a write-back that finds the version moved replaces another
 * tile's update.  felog_stats' train_rows_written / train_rows_lost report
 * the last round's counts.  on = 0 frees the counters. */
felog_status felog_track_lost_updates(felog_ctx *ctx, int on);

/* ---- measurement --------------------------------------------------------- */
/* No counterpart in the reference: a ceiling for the walk's memory access,
 * measured on the device.  Reads `reads` rows of read_bytes (16, 32, 64 or
 * 128) at independent pseudo-random rows of a zeroed table of table_bytes
 * on GPU `device` (four reads in flight per thread, no dependency between
 * them), five timed launches after a warm-up, and returns the rate in
 * *reads_per_s.  A walk step is one such 32-byte read whose address the
 * previous step decides, so the walk's steps per second divided by this
 * rate (same table size) is the walk's fraction of what the memory system
 * returns for its access pattern.  Allocates and frees the table itself. */
felog_status felog_probe_random_reads(int device, uint64_t table_bytes, uint64_t reads,
                                      uint32_t read_bytes, double *reads_per_s);

/* ---- test hook --------------------------------------------------------- */
/* Runs the SGNS kernels on a caller-supplied corpus with updates diverted:
 * instead of touching the tables they write, for every tile (a run of
 * centre groups trained together), the f32 logits and the table deltas into
 * caller buffers.  Used to compare the kernels' arithmetic with an FP64
 * restatement; every tile reads the tables as they were before the call.
 * Slot s < n_sentences holds the first tile of sentence s (all zero when the
 * sentence keeps fewer than two tokens); later slots hold further tiles.
 *   groups_out: per slot {sentence, first centre, n_centres, n_in, n_out}
 *   rows_out:   per slot 64 input + 32 output vertex ids (0xFFFFFFFF = pad)
 *   logits_out: per slot 64 x 32 f32
 *   dv_out:     per slot 64 x size f32 (input-row deltas)
 *   du_out:     per slot 32 x size f32 (output-row deltas)
 * max_groups bounds all buffers; *n_groups receives the number of slots
 * written; FELOG_ERR_CAPACITY when more were needed. */
felog_status felog_debug_sgns(felog_ctx *ctx, const uint64_t *offsets,
                              const uint32_t *tokens, uint64_t n_sentences,
                              uint32_t round, float alpha,
                              uint32_t max_groups, uint32_t *n_groups,
                              uint32_t *groups_out, uint32_t *rows_out,
                              float *logits_out, float *dv_out, float *du_out);

#ifdef __cplusplus
}
#endif
#endif /* FELOG_H */
