/*
 * felog_oracle.h - CPU restatement of the FeLoG walk -> SGNS path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the tests, the
 * smoke check and the CPU-baseline leg of bench.py compare the HIP path
 * against; the product library never links or calls it.
 *
 * Every function names the reference code (paths relative to the reference
 * repository) or the published method it restates.  Pins (each in tests/):
 *   orc_philox4x32_10      Random123 known-answer vectors
 *   orc_build_csr          brute-force adjacency on small graphs; order rule
 *   orc_build_alias        exact probability identity vs w_j / sum(w);
 *                          every bucket holds one unit; chi-square draws
 *   orc_corr_reference /
 *   orc_corr_exact         Shannon entropy and two-pass Pearson recomputed
 *                          from the visit histogram (closed forms)
 *   orc_walks              invariants (edges exist, stop rule) + the above;
 *                          compat transition: chi-square of the first step
 *                          and trace frequencies against closed form
 *   orc_tanh_restated      libm tanh within a few ulp
 *   orc_init_syn0          the reference's sequential LCG loop, element 0..n
 *   orc_sgns_groups        finite differences of the SGNS loss on d <= 8
 *   orc_w2v_train          single-pair update == hand-computed SGNS step;
 *                          with negatives and the sigmoid table: an FP64
 *                          restatement driven by the same LCG draws
 *   orc_w2v_unigram_table  cumulative slot counts vs table_size * the
 *                          cn^0.75 law (within one slot)
 *   orc_w2v_lut_index /
 *   orc_w2v_exp_table      hand-computed indices; entries vs the logistic
 *   orc_w2v_negative       a hand-traced LCG step and table lookup
 *   orc_eval               brute-force cosine on small tables
 *   orc_eval_subset        distinct in-range columns; inclusion frequency
 *                          30/deg (chi-square over many draws)
 *   orc_auc                scikit-learn roc_auc_score (fixtures)
 */
#ifndef FELOG_ORACLE_H
#define FELOG_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* CSR of the (optionally undirected) edge list, neighbours in input order. */
int orc_build_csr(uint32_t V, uint64_t m, const uint32_t *src, const uint32_t *dst, const float *w,
                  int undirected, uint64_t *row_ptr, uint32_t *col, float *wout);

/* Alias tables per vertex: canonical FP64 sum, integer shares of n * 2^32,
 * integer sweep pairing (u32 thresholds, column indices). */
void orc_build_alias(uint32_t V, const uint64_t *row_ptr, const float *w, uint32_t *thresh,
                     uint32_t *alias_idx);
/* Alias over one weight vector. */
void orc_alias_1d(uint32_t n, const double *wts, uint32_t *thresh, uint32_t *alias_idx);
/* Negative sampler over vertex ids: deg^0.75 (as sqrt(deg) * sqrt(sqrt(deg)))
 * for vocabulary vertices, 0 otherwise. */
void orc_unigram(uint32_t V, const uint32_t *deg, const uint8_t *in_vocab, uint32_t *thresh, uint32_t *alias_idx);

/* n*log2(n) for n = 0..L+1 and its forward differences for f = 0..L. */
void orc_xlog2x(uint32_t L, double *T, double *dT);

/* Walker statistics over a sequence of prior-visit counts f_1..f_n of the
 * vertices entered at steps 1..n; H[i], R[i] after step i+1.
 * _reference: Correlation::iterate, include/walk.hpp:296-366 (pow/log2);
 * _exact: the table form the GPU walk uses. */
void orc_corr_reference(uint32_t n, const uint32_t *f, double *H, double *R);
void orc_corr_exact(uint32_t n, const uint32_t *f, const double *T, const double *dT, double *H, double *R);

/* Host replay of one round of walks: alias steps, or with cn_data (edge
 * data per CSR entry) the reference's common-neighbour transition.  Walker
 * w starts at w % V and is skipped when flags[w % V] != 0.  Writes the flat
 * corpus in walker order: offsets[n_walks+1], tokens[...],
 * walker_ids[n_walks].  Returns n_walks; *n_tokens and *n_steps receive the
 * totals.  tokens may be NULL to count. */
uint64_t orc_walks(uint32_t V, const uint64_t *row_ptr, const uint32_t *col, const uint32_t *thresh,
                   const uint32_t *alias_idx, const float *cn_data, const uint8_t *flags, uint32_t walkers,
                   uint32_t min_L, uint32_t max_len, uint32_t round, uint64_t seed, uint64_t *offsets,
                   uint32_t *tokens, uint32_t *walker_ids, uint64_t *n_tokens, uint64_t *n_steps, int threads);
/* tanh restated from + - * / and exact scaling (the compat transition's
 * acceptance, include/walk.hpp:1005-1006). */
double orc_tanh_restated(double x);

/* Vocabulary (src/word2vec.cu:697-755,834-862): order by degree desc, id
 * asc; in_vocab = degree >= min_count or rank 0. */
void orc_vocab(uint32_t V, const uint32_t *deg, uint32_t min_count, uint32_t *order, uint32_t *rank,
               uint8_t *in_vocab, uint64_t *vocab_size, uint64_t *train_words);
/* Subsampling keep thresholds (src/word2vec.cu:180-212), u32 scale. */
void orc_keep_thresholds(uint32_t V, const uint32_t *deg, float sample, uint64_t train_words,
                         uint32_t *keep);
/* syn0 initialisation (src/word2vec.cu:864-902), rows by vertex id. */
void orc_init_syn0(uint32_t V, uint32_t d, const uint32_t *rank, float *syn0);

/* SGNS group oracle in FP64.  Parameters mirror the GPU kernel.  mode 0:
 * every group reads the initial tables and its logits / deltas are written
 * to the per-group buffers (layout of felog_debug_sgns); mode 1: groups are
 * applied in sentence order to syn0/syn1 (FP64 copies), no buffers.
 * Returns the number of groups. */
typedef struct orc_sgns_params {
    uint32_t V, d, window, negative, bgroup, rmax;
    int compat_break;
    float alpha;
    uint32_t round;
    uint64_t seed;
    const uint8_t *in_vocab;
    const uint32_t *vocab_rank, *keep_thr, *uni_thr, *uni_alias;
} orc_sgns_params;
uint32_t orc_sgns_groups(const orc_sgns_params *p, const uint64_t *offsets, const uint32_t *tokens,
                         uint64_t n_sent, double *syn0, double *syn1, int mode, uint32_t max_groups,
                         uint32_t *groups_out, uint32_t *rows_out, double *logits_out, double *dv_out,
                         double *du_out, uint64_t *pairs_out);

/* word2vec skip-gram restated (src/word2vec.cu:442-536 per pair, batching
 * :1169-1232), f32 Hogwild over `threads` OpenMP threads.  Tables are rows
 * of d floats indexed by vertex id; rank/in_vocab from orc_vocab.
 * use_lut: sigmoid from the 1000-entry expTable (src/word2vec.cu:2292-2297).
 * compat_break: vocabulary index 0 and dropped vertices end a sentence.
 * Returns the number of words read (pre-subsampling). */
uint64_t orc_w2v_train(uint32_t V, uint32_t d, float *syn0, float *syn1, const uint64_t *offsets,
                       const uint32_t *tokens, uint64_t n_sent, const uint32_t *deg, const uint32_t *order,
                       const uint32_t *rank, const uint8_t *in_vocab, uint64_t vocab_size,
                       uint64_t train_words, uint32_t window, uint32_t negative, float sample, float alpha,
                       int use_lut, int compat_break, uint32_t table_size, int threads, uint64_t *pairs);

/* The same with the trainer chosen by mode: 0 skip_gram_kernel, 1
 * __sgNegReuse (src/word2vec.cu:368-440: a full window and one set of
 * negatives per sentence, drawn as Philox(seed; j, W2V, sentence, round)
 * word 0 through the unigram table), 2 cbow_kernel (src/word2vec.cu:538-646).
 * pairs counts (centre, context) pairs, for CBOW the context rows averaged. */
uint64_t orc_w2v_train_mode(uint32_t V, uint32_t d, float *syn0, float *syn1, const uint64_t *offsets,
                            const uint32_t *tokens, uint64_t n_sent, const uint32_t *deg, const uint32_t *order,
                            const uint32_t *rank, const uint8_t *in_vocab, uint64_t vocab_size, uint64_t train_words,
                            uint32_t window, uint32_t negative, float sample, float alpha, int use_lut,
                            int compat_break, uint32_t table_size, int threads, uint64_t *pairs, int mode,
                            uint64_t seed, uint32_t round);

/* word2vec's unigram table over vocabulary indices (src/word2vec.cu:674-694);
 * cn[] = counts in vocabulary order. */
void orc_w2v_unigram_table(uint64_t vocab_size, const uint32_t *cn, uint32_t table_size, int32_t *table);
/* expTable[0..1000] (src/word2vec.cu:2292-2297) and the index a logit
 * f in (-6, 6] reads (src/word2vec.cu:525). */
void orc_w2v_exp_table(float *expTable);
int orc_w2v_lut_index(float f);
/* One negative draw of skip_gram_kernel (src/word2vec.cu:500-505); advances
 * *next_random and returns a vocabulary index. */
uint32_t orc_w2v_negative(uint64_t *next_random, const int32_t *table, uint32_t table_size, uint64_t vocab_size);

/* Neighbour columns the evaluation scores for vertex v of degree deg:
 * 0..deg-1 when deg <= 30, else 30 distinct columns by Floyd's sampling.
 * Returns the count written to chosen[30]. */
uint32_t orc_eval_subset(uint32_t deg, uint32_t v, uint32_t round, uint64_t seed, uint32_t *chosen);

/* Convergence scores (src/word2vec.cu:1748-1889,2016-2080) for every vertex
 * with in_vocab && !flags; others get score 0.  Returns #not converged and
 * sets flags of converged vertices. */
uint32_t orc_eval(uint32_t V, uint32_t d, const float *syn0, const uint64_t *row_ptr, const uint32_t *col,
                  const uint8_t *in_vocab, uint8_t *flags, float *scores, uint32_t round, uint64_t seed);

/* Common neighbours (src/tools/gcn.cpp:304-357) of m pairs (u[i], v[i]):
 * distinct vertices adjacent to both, the pair's own vertices excluded. */
void orc_common_neighbours(uint32_t V, const uint64_t *row_ptr, const uint32_t *col, uint64_t m, const uint32_t *u,
                           const uint32_t *v, int32_t *cn);

/* -classes k-means (src/word2vec.cu:2128-2175) over the vocabulary's rows;
 * cl[vocab_size] receives the classes in vocabulary order. */
void orc_kmeans_classes(uint32_t d, const float *syn0, const uint32_t *order, uint64_t vocab_size, uint32_t classes,
                        uint32_t iter, int32_t *cl);

/* Mann-Whitney AUC with ties counted 1/2. */
double orc_auc(uint64_t n_pos, const float *pos, uint64_t n_neg, const float *neg);

#ifdef __cplusplus
}
#endif
#endif
