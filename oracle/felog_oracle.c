/*
 * felog_oracle.c - CPU restatement of the FeLoG walk -> SGNS path.
 * TEST INFRASTRUCTURE ONLY (see felog_oracle.h): the checker for the HIP
 * path and the CPU baseline of bench.py; never linked by the product.
 *
 * Built with -ffp-contract=off so that every floating-point operation of the
 * walk statistics and of the alias construction is a single IEEE operation.
 */
#include "felog_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---- Philox4x32-10 (Salmon et al., SC'11; Random123 philox.h) ---------- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t x0 = ctr[0], x1 = ctr[1], x2 = ctr[2], x3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        const uint64_t a = (uint64_t)0xD2511F53u * x0;
        const uint64_t b = (uint64_t)0xCD9E8D57u * x2;
        const uint32_t y0 = (uint32_t)(b >> 32) ^ x1 ^ k0;
        const uint32_t y1 = (uint32_t)b;
        const uint32_t y2 = (uint32_t)(a >> 32) ^ x3 ^ k1;
        const uint32_t y3 = (uint32_t)a;
        x0 = y0; x1 = y1; x2 = y2; x3 = y3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = x0; out[1] = x1; out[2] = x2; out[3] = x3;
}

/* counter domains (y word) */
#define D_WALK 0x00000000u
/* One SGNS draw per (index x, sentence z): word 0 subsamples raw position
 * x, word 1 is the window shrink of kept centre x, words 2-3 draw negative
 * j of group g for x = g * negative + j (bucket, coin). */
#define D_SGNS 0x10000000u
#define D_EVAL 0x40000000u

static void rnd(uint64_t seed, uint32_t x, uint32_t y, uint32_t z, uint32_t w, uint32_t out[4]) {
    const uint32_t c[4] = {x, y, z, w};
    const uint32_t k[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    orc_philox4x32_10(c, k, out);
}

static uint32_t mulhi32(uint32_t a, uint32_t b) { return (uint32_t)(((uint64_t)a * b) >> 32); }

/* ---- CSR: include/graph.hpp:741-772 (undirect), :775-805 (degrees),
 *      :531-575 (adjacency in input order) ------------------------------ */
int orc_build_csr(uint32_t V, uint64_t m, const uint32_t *src, const uint32_t *dst, const float *w,
                  int undirected, uint64_t *row_ptr, uint32_t *col, float *wout) {
    const uint64_t E = undirected ? 2 * m : m;
    for (uint64_t i = 0; i < m; ++i)
        if (src[i] >= V || dst[i] >= V) return -1;
    memset(row_ptr, 0, ((uint64_t)V + 1) * sizeof(uint64_t));
    for (uint64_t i = 0; i < m; ++i) {
        row_ptr[src[i] + 1]++;
        if (undirected) row_ptr[dst[i] + 1]++;
    }
    for (uint32_t v = 0; v < V; ++v) row_ptr[v + 1] += row_ptr[v];
    uint64_t *end = (uint64_t *)malloc(((uint64_t)V + 1) * sizeof(uint64_t));
    memcpy(end, row_ptr, ((uint64_t)V + 1) * sizeof(uint64_t));
    for (uint64_t e = 0; e < E; ++e) {
        const uint64_t i = undirected ? e / 2 : e;
        const int rev = undirected && (e & 1);
        const uint32_t s = rev ? dst[i] : src[i], t = rev ? src[i] : dst[i];
        const uint64_t p = end[s]++;
        col[p] = t;
        wout[p] = w[i];
    }
    free(end);
    return 0;
}

/* ---- alias tables: exact integer construction --------------------------
 * The reference declares alias buckets (include/walk.hpp:87-103) but never
 * builds them; this construction is the build's own.  For a vertex with
 * n columns and weights w_j (negative weights count as 0):
 *  1. S = the canonical FP64 sum of the weights: the columns are cut into
 *     chunks of 4096; a chunk's sum is the xor-butterfly combination of 64
 *     lane partials, lane l summing the chunk's entries j = l (mod 64) in
 *     order; the chunk sums are combined by the same rule.
 *  2. q_j = trunc(((w_j * n) / S) * 2^32) in FP64, an integer share of the
 *     total n * 2^32; the remainder n * 2^32 - sum q goes to the largest
 *     q_j (lowest column on ties).  S = 0 gives q_j = 2^32 for every j.
 *  3. The sweep of Hubschle-Schneider and Sanders ("Parallel weighted
 *     random sampling", ESA 2019) over integers: light columns (q < 2^32)
 *     in column order take the current heavy column (q >= 2^32) as their
 *     alias and its residual drops by 2^32 - q_light; a heavy column whose
 *     residual falls below 2^32 becomes a bucket of that residual aliased
 *     to the next heavy column, which absorbs the difference; heavy
 *     columns never reduced keep their whole bucket.
 * Every bucket then holds exactly 2^32 units, so P(j) = q_j / (n 2^32):
 * the only error against w_j / S is the truncation.  Column j keeps itself
 * with probability thresh_j / 2^32 (0xFFFFFFFF = a whole bucket). */
#define ALIAS_ONE 4294967296ull

static double lanesum(const double *x, uint64_t n) {
    double p[64], q[64];
    for (int l = 0; l < 64; ++l) p[l] = 0.0;
    for (uint64_t j = 0; j < n; ++j) p[j & 63] = p[j & 63] + x[j];
    for (int o = 32; o > 0; o >>= 1) {
        for (int l = 0; l < 64; ++l) q[l] = p[l] + p[l ^ o];
        memcpy(p, q, sizeof p);
    }
    return p[0];
}

static double canonical_sum(const double *x, uint64_t n) {
    const uint64_t nc = (n + 4095) / 4096;
    double *cs = (double *)malloc((nc ? nc : 1) * sizeof(double));
    for (uint64_t k = 0; k < nc; ++k) cs[k] = lanesum(x + 4096 * k, n - 4096 * k < 4096 ? n - 4096 * k : 4096);
    const double s = lanesum(cs, nc);
    free(cs);
    return s;
}

/* x: n clamped weights; q, lt, hv: scratch of n entries. */
static void alias_exact(uint64_t n, const double *x, uint64_t *q, uint32_t *lt, uint32_t *hv, uint32_t *thresh,
                        uint32_t *alias_idx) {
    if (n == 0) return;
    const double S = canonical_sum(x, n);
    if (S > 0.0) {
        const double dn = (double)n;
        uint64_t sq = 0, jmax = 0;
        for (uint64_t j = 0; j < n; ++j) {
            q[j] = (uint64_t)(((x[j] * dn) / S) * 4294967296.0);
            sq += q[j];
            if (q[j] > q[jmax]) jmax = j;
        }
        q[jmax] += n * ALIAS_ONE - sq;  /* exact: the true result is in range */
    } else {
        for (uint64_t j = 0; j < n; ++j) q[j] = ALIAS_ONE;
    }
    uint64_t nl = 0, nh = 0;
    for (uint64_t j = 0; j < n; ++j) {
        if (q[j] < ALIAS_ONE) lt[nl++] = (uint32_t)j;
        else hv[nh++] = (uint32_t)j;
    }
    for (uint64_t j = 0; j < n; ++j) { thresh[j] = 0xFFFFFFFFu; alias_idx[j] = (uint32_t)j; }
    uint64_t h = 0, r = nh ? q[hv[0]] : 0;
    for (uint64_t i = 0; i < nl; ++i) {
        const uint32_t l = lt[i];
        thresh[l] = (uint32_t)q[l];
        alias_idx[l] = hv[h];
        r -= ALIAS_ONE - q[l];
        while (r < ALIAS_ONE && h + 1 < nh) {
            thresh[hv[h]] = (uint32_t)r;
            alias_idx[hv[h]] = hv[h + 1];
            r = q[hv[h + 1]] - (ALIAS_ONE - r);
            ++h;
        }
    }
}

void orc_build_alias(uint32_t V, const uint64_t *row_ptr, const float *w, uint32_t *thresh, uint32_t *alias_idx) {
    uint64_t maxd = 0;
    for (uint32_t v = 0; v < V; ++v)
        if (row_ptr[v + 1] - row_ptr[v] > maxd) maxd = row_ptr[v + 1] - row_ptr[v];
    double *wd = (double *)malloc((maxd + 1) * sizeof(double));
    uint64_t *q = (uint64_t *)malloc((maxd + 1) * sizeof(uint64_t));
    uint32_t *lt = (uint32_t *)malloc((maxd + 1) * sizeof(uint32_t));
    uint32_t *hv = (uint32_t *)malloc((maxd + 1) * sizeof(uint32_t));
    for (uint32_t v = 0; v < V; ++v) {
        const uint64_t b = row_ptr[v];
        const uint64_t n = row_ptr[v + 1] - b;
        for (uint64_t j = 0; j < n; ++j) wd[j] = w[b + j] > 0.f ? (double)w[b + j] : 0.0;
        alias_exact(n, wd, q, lt, hv, thresh + b, alias_idx + b);
    }
    free(wd); free(q); free(lt); free(hv);
}

void orc_alias_1d(uint32_t n, const double *wts, uint32_t *thresh, uint32_t *alias_idx) {
    double *x = (double *)malloc(((uint64_t)n + 1) * sizeof(double));
    uint64_t *q = (uint64_t *)malloc(((uint64_t)n + 1) * sizeof(uint64_t));
    uint32_t *lt = (uint32_t *)malloc(((uint64_t)n + 1) * sizeof(uint32_t));
    uint32_t *hv = (uint32_t *)malloc(((uint64_t)n + 1) * sizeof(uint32_t));
    for (uint32_t j = 0; j < n; ++j) x[j] = wts[j] > 0.0 ? wts[j] : 0.0;
    alias_exact(n, x, q, lt, hv, thresh, alias_idx);
    free(x); free(q); free(lt); free(hv);
}

/* Negative sampler: weight deg^0.75 for vocabulary vertices, 0 otherwise
 * (src/word2vec.cu:674-694 draws from the same law through a table), as an
 * alias table over vertex ids.  deg^0.75 = sqrt(deg) * sqrt(sqrt(deg)):
 * correctly rounded square roots give the same bits on every IEEE machine,
 * where pow() may differ by an ulp between libraries. */
void orc_unigram(uint32_t V, const uint32_t *deg, const uint8_t *in_vocab, uint32_t *thresh, uint32_t *alias_idx) {
    double *wts = (double *)malloc(((uint64_t)V + 1) * sizeof(double));
    for (uint32_t v = 0; v < V; ++v) {
        const double x = (double)deg[v];
        wts[v] = in_vocab[v] ? sqrt(x) * sqrt(sqrt(x)) : 0.0;
    }
    orc_alias_1d(V, wts, thresh, alias_idx);
    free(wts);
}

/* ---- walker statistics ----------------------------------------------- */
void orc_xlog2x(uint32_t L, double *T, double *dT) {
    for (uint32_t n = 0; n < L + 2; ++n) T[n] = n <= 1 ? 0.0 : (double)n * log2((double)n);
    for (uint32_t f = 0; f < L + 1; ++f) dT[f] = T[f + 1] - T[f];
}

/* Correlation::iterate of include/walk.hpp:296-366 with step_k = 1, started
 * from the state set at a walker's first vertex (:937-939). */
void orc_corr_reference(uint32_t n, const uint32_t *f, double *H, double *R) {
    double Hn = 0.0, Sn = 1.0, EH = 0.0, ES = 1.0, EHS = 0.0, EHH = 0.0, ESS = 1.0, N = 1.0;
    for (uint32_t i = 0; i < n; ++i) {
        const double fi = (double)f[i];
        const double Snp1 = Sn + 1.0;
        const double ss = pow(Sn / Snp1, Sn) * pow(1.0 / Snp1, 1.0);
        const double t = fi == 0.0 ? log2(ss) : log2(ss * (pow((fi + 1.0) / fi, fi) * pow(fi + 1.0, 1.0)));
        const double Hnp1 = (Sn * Hn - t) / Snp1;
        const double Np1 = N + 1.0;
        EH = EH + (Hnp1 - EH) / Np1;
        ES = ES + (Snp1 - ES) / Np1;
        EHS = EHS + (Hnp1 * Snp1 - EHS) / Np1;
        EHH = EHH + (Hnp1 * Hnp1 - EHH) / Np1;
        ESS = ESS + (Snp1 * Snp1 - ESS) / Np1;
        const double DH = EHH - EH * EH, DS = ESS - ES * ES, e = EHS - EH * ES;
        R[i] = e == 0.0 ? 0.0 : e / sqrt(DH * DS);
        H[i] = Hnp1;
        Hn = Hnp1; Sn = Snp1; N = Np1;
    }
}

typedef struct { double A, EH, ES, EHS, EHH, ESS; } stats_t;

static void stats_init(stats_t *s) { s->A = 0.0; s->EH = 0.0; s->ES = 1.0; s->EHS = 0.0; s->EHH = 0.0; s->ESS = 1.0; }

/* returns H, writes R; S = vertices on the path after the step */
static double stats_step(stats_t *s, const double *T, const double *dT, uint32_t f, uint32_t S_int, double *Rout) {
    s->A = s->A + dT[f];
    const double S = (double)S_int;
    const double H = (T[S_int] - s->A) / S;
    const double N = S;
    s->EH = s->EH + (H - s->EH) / N;
    s->ES = s->ES + (S - s->ES) / N;
    s->EHS = s->EHS + (H * S - s->EHS) / N;
    s->EHH = s->EHH + (H * H - s->EHH) / N;
    s->ESS = s->ESS + (S * S - s->ESS) / N;
    const double DH = s->EHH - s->EH * s->EH;
    const double DS = s->ESS - s->ES * s->ES;
    const double e = s->EHS - s->EH * s->ES;
    const double den = DH * DS;
    double R = 0.0;
    if (e != 0.0 && den > 0.0) R = e / sqrt(den);
    *Rout = R;
    return H;
}

void orc_corr_exact(uint32_t n, const uint32_t *f, const double *T, const double *dT, double *H, double *R) {
    stats_t s;
    stats_init(&s);
    for (uint32_t i = 0; i < n; ++i) H[i] = stats_step(&s, T, dT, f[i], i + 2, &R[i]);
}

/* One alias walk (the GPU path's semantics):
 *   include/walk.hpp:964-969  degree 0 ends the walk
 *   include/walk.hpp:977-979  degree 2 with itself as first neighbour ends it
 *   include/walk.hpp:981-983  a self loop at step 0 is redrawn (<= 64 times)
 *   include/walk.hpp:1056-1062 stop when step > min_L and (R^2 < 0.999 or R < 0)
 * Draw for step s, attempt a: Philox(seed; s, a, walker, round); column
 * hi32(r0*deg); keep it if r1 < thresh else take its alias. */
static uint32_t walk_one(uint32_t w, uint32_t V, const uint64_t *row_ptr, const uint32_t *col,
                         const uint32_t *thresh, const uint32_t *alias_idx, uint32_t min_L, uint32_t max_len,
                         uint32_t round, uint64_t seed, const double *T, const double *dT, uint32_t *path) {
    uint32_t v = w % V, s = 0, attempt = 0;
    stats_t st;
    stats_init(&st);
    path[0] = v;
    for (;;) {
        if (s + 1 >= max_len) break;
        const uint64_t b = row_ptr[v];
        const uint32_t deg = (uint32_t)(row_ptr[v + 1] - b);
        if (deg == 0) break;
        if (deg == 2 && col[b] == v) break;
        uint32_t r[4];
        rnd(seed, s, D_WALK | attempt, w, round, r);
        const uint32_t j = mulhi32(r[0], deg);
        const uint32_t x = r[1] < thresh[b + j] ? col[b + j] : col[b + alias_idx[b + j]];
        if (s == 0 && x == v) {
            if (++attempt >= 64) break;
            continue;
        }
        attempt = 0;
        uint32_t f = 0;
        for (uint32_t k = 0; k <= s; ++k) f += path[k] == x;
        ++s;
        path[s] = x;
        double R;
        stats_step(&st, T, dT, f, s + 1, &R);
        v = x;
        if (s > min_L && (R * R < 0.999 || R < 0.0)) break;
    }
    return s + 1;
}

/* The reference's transition (include/walk.hpp:962-1101): a uniform
 * neighbour x of the current vertex v (Philox word 0, as the alias column),
 * u = a uniform double from words 1-2, p = (1 / (deg v - cn)) * max(deg v,
 * deg x) / min(deg v, deg x) with cn = (int) edge data; when tanh(p) <= u
 * the walker appends x and then v again (a "trace", :1015-1019,1063-1094)
 * and stays at v, else it moves to x.  Each appended vertex runs the stop
 * rule.  tanh is restated from + - * / and exact scaling (Taylor series of
 * e^y - 1 on |y| <= 0.5 to the y^18 term, 2^k scaling with fdlibm's
 * two-part ln2 beyond), evaluated in the same order as the device so the
 * decisions agree bit for bit; its accuracy is pinned against libm. */
static double expm1_series(double y) {
    double p = 1.0;
    for (int i = 18; i >= 2; --i) p = 1.0 + (y / (double)i) * p;
    return y * p;
}

double orc_tanh_restated(double x) {
    const int neg = x < 0.0;
    const double a = neg ? -x : x;
    if (!(a < 20.0)) return neg ? -1.0 : 1.0;
    const double y = a + a;
    double t;
    if (y <= 0.5) {
        t = expm1_series(y);
    } else {
        const double k = floor(y * 1.4426950408889634 + 0.5);
        const double r = (y - k * 6.93147180369123816490e-01) - k * 1.90821492927058770002e-10;
        t = ldexp(1.0 + expm1_series(r), (int)k) - 1.0;
    }
    const double th = t / (t + 2.0);
    return neg ? -th : th;
}

static uint32_t walk_one_cn(uint32_t w, uint32_t V, const uint64_t *row_ptr, const uint32_t *col,
                            const float *cn_data, uint32_t min_L, uint32_t max_len, uint32_t round, uint64_t seed,
                            const double *T, const double *dT, uint32_t *path) {
    uint32_t v = w % V, s = 0, attempt = 0;
    stats_t st;
    stats_init(&st);
    path[0] = v;
    for (;;) {
        if (s + 1 >= max_len) break;
        const uint64_t b = row_ptr[v];
        const uint32_t deg = (uint32_t)(row_ptr[v + 1] - b);
        if (deg == 0) break;
        if (deg == 2 && col[b] == v) break;
        uint32_t r[4];
        rnd(seed, s, D_WALK | attempt, w, round, r);
        const uint32_t j = mulhi32(r[0], deg);
        const uint32_t x = col[b + j];
        if (s == 0 && x == v) {
            if (++attempt >= 64) break;
            continue;
        }
        attempt = 0;
        const uint32_t dx = (uint32_t)(row_ptr[x + 1] - row_ptr[x]);
        const int64_t cn = (int64_t)(int)cn_data[b + j];
        const double hi = (double)(deg > dx ? deg : dx), lo = (double)(deg < dx ? deg : dx);
        const double p = ((1.0 / (double)((int64_t)deg - cn)) * hi) / lo;
        const double u = (double)((((uint64_t)r[1]) << 21) | (r[2] >> 11)) * 0x1.0p-53;
        const int trace = orc_tanh_restated(p) <= u;
        int stop = 0;
        for (int leg = 0; leg < (trace ? 2 : 1) && !stop; ++leg) {
            const uint32_t y = leg ? v : x;
            if (leg && s + 1 >= max_len) { stop = 1; break; }
            uint32_t f = 0;
            for (uint32_t k = 0; k <= s; ++k) f += path[k] == y;
            ++s;
            path[s] = y;
            double R;
            stats_step(&st, T, dT, f, s + 1, &R);
            stop = s > min_L && (R * R < 0.999 || R < 0.0);
        }
        if (stop) break;
        if (!trace) v = x;
    }
    return s + 1;
}

/* Walks of one round; cn_data (edge data per CSR entry) selects the
 * reference's transition, NULL the alias step. */
uint64_t orc_walks(uint32_t V, const uint64_t *row_ptr, const uint32_t *col, const uint32_t *thresh,
                   const uint32_t *alias_idx, const float *cn_data, const uint8_t *flags, uint32_t walkers,
                   uint32_t min_L, uint32_t max_len, uint32_t round, uint64_t seed, uint64_t *offsets,
                   uint32_t *tokens, uint32_t *walker_ids, uint64_t *n_tokens, uint64_t *n_steps, int threads) {
    double *T = (double *)malloc(((uint64_t)max_len + 2) * sizeof(double));
    double *dT = (double *)malloc(((uint64_t)max_len + 1) * sizeof(double));
    orc_xlog2x(max_len, T, dT);
    uint32_t *lens = (uint32_t *)calloc(walkers ? walkers : 1, sizeof(uint32_t));
    uint32_t *act = (uint32_t *)malloc((walkers ? walkers : 1) * sizeof(uint32_t));
    uint64_t n = 0;
    for (uint32_t w = 0; w < walkers; ++w)
        if (!flags || !flags[w % V]) act[n++] = w;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#endif
    /* pass 1: lengths */
#pragma omp parallel
    {
        uint32_t *path = (uint32_t *)malloc((uint64_t)max_len * sizeof(uint32_t));
#pragma omp for schedule(dynamic, 256)
        for (int64_t i = 0; i < (int64_t)n; ++i)
            lens[i] = cn_data ? walk_one_cn(act[i], V, row_ptr, col, cn_data, min_L, max_len, round, seed, T, dT, path)
                              : walk_one(act[i], V, row_ptr, col, thresh, alias_idx, min_L, max_len, round, seed, T,
                                         dT, path);
        free(path);
    }
    uint64_t tot = 0, steps = 0;
    if (offsets) offsets[0] = 0;
    for (uint64_t i = 0; i < n; ++i) {
        tot += lens[i];
        steps += lens[i] - 1;
        if (offsets) offsets[i + 1] = tot;
        if (walker_ids) walker_ids[i] = act[i];
    }
    /* pass 2: paths */
    if (tokens && offsets) {
#pragma omp parallel
        {
            uint32_t *path = (uint32_t *)malloc((uint64_t)max_len * sizeof(uint32_t));
#pragma omp for schedule(dynamic, 256)
            for (int64_t i = 0; i < (int64_t)n; ++i) {
                const uint32_t len =
                    cn_data ? walk_one_cn(act[i], V, row_ptr, col, cn_data, min_L, max_len, round, seed, T, dT, path)
                            : walk_one(act[i], V, row_ptr, col, thresh, alias_idx, min_L, max_len, round, seed, T,
                                       dT, path);
                memcpy(tokens + offsets[i], path, (uint64_t)len * sizeof(uint32_t));
            }
            free(path);
        }
    }
    if (n_tokens) *n_tokens = tot;
    if (n_steps) *n_steps = steps;
    free(T); free(dT); free(lens); free(act);
    return n;
}

/* ---- vocabulary, subsampling, initialisation ------------------------- */
static const uint32_t *g_sort_deg;
static int by_degree(const void *a, const void *b) {
    const uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
    if (g_sort_deg[x] != g_sort_deg[y]) return g_sort_deg[x] > g_sort_deg[y] ? -1 : 1;
    return x < y ? -1 : (x > y);
}

void orc_vocab(uint32_t V, const uint32_t *deg, uint32_t min_count, uint32_t *order, uint32_t *rank,
               uint8_t *in_vocab, uint64_t *vocab_size, uint64_t *train_words) {
    for (uint32_t v = 0; v < V; ++v) order[v] = v;
    g_sort_deg = deg;
    qsort(order, V, sizeof(uint32_t), by_degree);
    for (uint32_t i = 0; i < V; ++i) rank[order[i]] = i;
    uint64_t vs = 0, tw = 0;
    for (uint32_t v = 0; v < V; ++v) {
        in_vocab[v] = (deg[v] >= min_count || rank[v] == 0) ? 1 : 0;
        if (in_vocab[v]) { ++vs; tw += deg[v]; }
    }
    *vocab_size = vs;
    *train_words = tw;
}

void orc_keep_thresholds(uint32_t V, const uint32_t *deg, float sample, uint64_t train_words, uint32_t *keep) {
    const double sT = (double)sample * (double)train_words;
    for (uint32_t v = 0; v < V; ++v) {
        keep[v] = 0xFFFFFFFFu;
        if (!(sample > 0.f) || train_words == 0 || deg[v] == 0) continue;
        const double c = (double)deg[v];
        const double prob = (sqrt(c / sT) + 1.0) * (sT / c);
        if (prob < 1.0) keep[v] = prob <= 0.0 ? 0u : (uint32_t)(prob * 4294967296.0);
    }
}

void orc_init_syn0(uint32_t V, uint32_t d, const uint32_t *rank, float *syn0) {
    uint32_t *order = (uint32_t *)malloc((uint64_t)V * sizeof(uint32_t));
    for (uint32_t v = 0; v < V; ++v) order[rank[v]] = v;
    unsigned long long next_random = 1;
    for (uint32_t a = 0; a < V; ++a)
        for (uint32_t b = 0; b < d; ++b) {
            next_random = next_random * 25214903917ull + 11;
            syn0[(uint64_t)order[a] * d + b] = (float)(((double)(float)((next_random & 0xFFFF) / (float)65536) - 0.5) / (double)d);
        }
    free(order);
}

/* ---- SGNS group oracle (FP64) ----------------------------------------- */
#define RMAX 64
#define OMAX 32
#define MAXKEEP 1000

static double sigmoid_clamped(double x) {
    if (x > 6.0) return 1.0;
    if (x < -6.0) return 0.0;
    return 1.0 / (1.0 + exp(-x));
}

/* Sentence filter shared by the SGNS oracle: vocabulary, subsampling from
 * word 0 of Philox(seed; position, SGNS, sentence, round), 1000-token cut. */
static uint32_t filter_sentence(const orc_sgns_params *p, const uint32_t *t, uint32_t n, uint64_t sid, uint32_t *K) {
    uint32_t m = 0;
    for (uint32_t i = 0; i < n && m < MAXKEEP; ++i) {
        const uint32_t x = t[i];
        const int iv = p->in_vocab[x] != 0;
        if (p->compat_break && (!iv || p->vocab_rank[x] == 0)) break;
        if (!iv) continue;
        const uint32_t thr = p->keep_thr[x];
        if (thr != 0xFFFFFFFFu) {
            uint32_t r[4];
            rnd(p->seed, i, D_SGNS, (uint32_t)sid, p->round, r);
            if (!(r[0] < thr)) continue;
        }
        K[m++] = x;
    }
    return m;
}

uint32_t orc_sgns_groups(const orc_sgns_params *p, const uint64_t *offsets, const uint32_t *tokens, uint64_t n_sent,
                         double *syn0, double *syn1, int mode, uint32_t max_groups, uint32_t *groups_out,
                         uint32_t *rows_out, double *logits_out, double *dv_out, double *du_out, uint64_t *pairs_out) {
    const uint32_t d = p->d, w = p->window, k = p->negative, B = p->bgroup;
    uint32_t *K = (uint32_t *)malloc(MAXKEEP * sizeof(uint32_t));
    double *X = (double *)malloc(RMAX * OMAX * sizeof(double));
    double *G = (double *)malloc(RMAX * OMAX * sizeof(double));
    double *dV = (double *)malloc((uint64_t)RMAX * d * sizeof(double));
    double *dU = (double *)malloc((uint64_t)OMAX * d * sizeof(double));
    uint32_t P[RMAX][OMAX], Q[RMAX][OMAX];
    uint32_t ntile = 0;
    uint64_t pairs = 0;
    for (uint64_t sid = 0; sid < n_sent; ++sid) {
        const uint64_t b0 = offsets[sid];
        const uint32_t m = filter_sentence(p, tokens + b0, (uint32_t)(offsets[sid + 1] - b0), sid, K);
        if (m < 2) continue;
        for (uint32_t g0 = 0; g0 * B < m;) {
            /* Tile: consecutive groups g0.. while centres + k per group <= 32
             * output slots and the windows span <= rmax positions. */
            const uint32_t c0 = g0 * B;
            const uint32_t lo = c0 >= w ? c0 - w : 0;
            uint32_t ng = 1;
            for (;;) {
                const uint32_t gn = g0 + ng;
                if (gn * B >= m) break;
                const uint32_t ce = (gn + 1) * B < m ? (gn + 1) * B : m;
                const uint32_t he = ce + w < m ? ce + w : m;
                if ((ce - c0) + k * (ng + 1) > OMAX || he - lo > p->rmax) break;
                ++ng;
            }
            const uint32_t c_end = (g0 + ng) * B < m ? (g0 + ng) * B : m;
            const uint32_t nc = c_end - c0, nneg = k * ng;
            const uint32_t hi = c_end + w < m ? c_end + w : m;
            uint32_t bshr[OMAX], neg[OMAX], rin_tok[RMAX], rout_tok[OMAX], pos_rin[RMAX], cen_rout[OMAX], neg_rout[OMAX];
            for (uint32_t i = 0; i < nc; ++i) {
                uint32_t r[4];
                rnd(p->seed, c0 + i, D_SGNS, (uint32_t)sid, p->round, r);
                bshr[i] = r[1] % w;
            }
            for (uint32_t q = 0; q < nneg; ++q) {
                uint32_t r[4];
                /* negative q % k of group g0 + q / k: index g0 * k + q */
                rnd(p->seed, g0 * k + q, D_SGNS, (uint32_t)sid, p->round, r);
                const uint32_t bk = mulhi32(r[2], p->V);
                neg[q] = r[3] < p->uni_thr[bk] ? bk : p->uni_alias[bk];
            }
            uint32_t n_in = 0, n_out = 0;
            for (uint32_t q = lo; q < hi; ++q) {
                uint32_t f = n_in;
                for (uint32_t r = 0; r < n_in; ++r)
                    if (rin_tok[r] == K[q]) { f = r; break; }
                if (f == n_in) rin_tok[n_in++] = K[q];
                pos_rin[q - lo] = f;
            }
            for (uint32_t q = 0; q < nc + nneg; ++q) {
                const uint32_t t = q < nc ? K[c0 + q] : neg[q - nc];
                uint32_t f = n_out;
                for (uint32_t r = 0; r < n_out; ++r)
                    if (rout_tok[r] == t) { f = r; break; }
                if (f == n_out) rout_tok[n_out++] = t;
                if (q < nc) cen_rout[q] = f;
                else neg_rout[q - nc] = f;
            }
            memset(P, 0, sizeof(P));
            memset(Q, 0, sizeof(Q));
            for (uint32_t i = 0; i < nc; ++i)
                for (uint32_t a = bshr[i]; a < 2 * w + 1 - bshr[i]; ++a) {
                    if (a == w) continue;
                    const int64_t pp = (int64_t)(c0 + i) - (int64_t)w + (int64_t)a;
                    if (pp < 0 || pp >= (int64_t)m) continue;
                    const uint32_t r = pos_rin[pp - lo];
                    P[r][cen_rout[i]]++;
                    const uint32_t nb = ((c0 + i) / B - g0) * k;  /* this centre's group */
                    for (uint32_t j = 0; j < k; ++j)
                        if (neg[nb + j] != K[c0 + i]) Q[r][neg_rout[nb + j]]++;
                    ++pairs;
                }
            for (uint32_t r = 0; r < n_in; ++r)
                for (uint32_t o = 0; o < n_out; ++o) {
                    double x = 0.0;
                    const double *vr = syn0 + (uint64_t)rin_tok[r] * d;
                    const double *uo = syn1 + (uint64_t)rout_tok[o] * d;
                    for (uint32_t e = 0; e < d; ++e) x += vr[e] * uo[e];
                    X[r * OMAX + o] = x;
                    G[r * OMAX + o] = (double)p->alpha * ((double)P[r][o] - (double)(P[r][o] + Q[r][o]) * sigmoid_clamped(x));
                }
            for (uint32_t r = 0; r < n_in; ++r)
                for (uint32_t e = 0; e < d; ++e) {
                    double acc = 0.0;
                    for (uint32_t o = 0; o < n_out; ++o) acc += G[r * OMAX + o] * syn1[(uint64_t)rout_tok[o] * d + e];
                    dV[(uint64_t)r * d + e] = acc;
                }
            for (uint32_t o = 0; o < n_out; ++o)
                for (uint32_t e = 0; e < d; ++e) {
                    double acc = 0.0;
                    for (uint32_t r = 0; r < n_in; ++r) acc += G[r * OMAX + o] * syn0[(uint64_t)rin_tok[r] * d + e];
                    dU[(uint64_t)o * d + e] = acc;
                }
            if (mode == 0) {
                if (ntile < max_groups) {
                    uint32_t *gq = groups_out + (uint64_t)ntile * 5;
                    gq[0] = (uint32_t)sid; gq[1] = c0; gq[2] = nc; gq[3] = n_in; gq[4] = n_out;
                    uint32_t *rr = rows_out + (uint64_t)ntile * (RMAX + OMAX);
                    for (uint32_t r = 0; r < RMAX; ++r) rr[r] = r < n_in ? rin_tok[r] : 0xFFFFFFFFu;
                    for (uint32_t o = 0; o < OMAX; ++o) rr[RMAX + o] = o < n_out ? rout_tok[o] : 0xFFFFFFFFu;
                    double *lg = logits_out + (uint64_t)ntile * RMAX * OMAX;
                    for (uint32_t r = 0; r < RMAX; ++r)
                        for (uint32_t o = 0; o < OMAX; ++o)
                            lg[r * OMAX + o] = (r < n_in && o < n_out) ? X[r * OMAX + o] : 0.0;
                    double *dv = dv_out + (uint64_t)ntile * RMAX * d;
                    double *du = du_out + (uint64_t)ntile * OMAX * d;
                    memset(dv, 0, (uint64_t)RMAX * d * sizeof(double));
                    memset(du, 0, (uint64_t)OMAX * d * sizeof(double));
                    memcpy(dv, dV, (uint64_t)n_in * d * sizeof(double));
                    memcpy(du, dU, (uint64_t)n_out * d * sizeof(double));
                }
            } else {
                for (uint32_t r = 0; r < n_in; ++r)
                    for (uint32_t e = 0; e < d; ++e) syn0[(uint64_t)rin_tok[r] * d + e] += dV[(uint64_t)r * d + e];
                for (uint32_t o = 0; o < n_out; ++o)
                    for (uint32_t e = 0; e < d; ++e) syn1[(uint64_t)rout_tok[o] * d + e] += dU[(uint64_t)o * d + e];
            }
            ++ntile;
            g0 += ng;
        }
    }
    if (pairs_out) *pairs_out = pairs;
    free(K); free(X); free(G); free(dV); free(dU);
    return ntile;
}

/* ---- word2vec skip-gram restatement --------------------------------- */
#define EXP_TABLE_SIZE 1000
#define MAX_EXP 6
#define MAX_SENTENCE 15000

/* InitUnigramTable, src/word2vec.cu:674-694: slot a holds vocabulary index
 * i while a / table_size has not passed the cumulative share of indices
 * 0..i under cn^0.75; cn[] is in vocabulary (rank) order. */
void orc_w2v_unigram_table(uint64_t vocab_size, const uint32_t *cn, uint32_t table_size, int32_t *table) {
    double pw = 0.0;
    for (uint64_t a = 0; a < vocab_size; ++a) pw += pow((double)cn[a], 0.75);
    uint64_t i = 0;
    double d1 = pow((double)cn[0], 0.75) / pw;
    for (uint64_t a = 0; a < table_size; ++a) {
        table[a] = (int32_t)i;
        if (a / (double)table_size > d1) {
            ++i;
            if (i < vocab_size) d1 += pow((double)cn[i], 0.75) / pw;
        }
        if (i >= vocab_size) i = vocab_size - 1;
    }
}

/* expTable, src/word2vec.cu:2292-2297: entry i = sigma((i / 1000 * 2 - 1) * 6),
 * plus a copy of the last entry at index 1000. */
void orc_w2v_exp_table(float *expTable) {
    for (int i = 0; i < EXP_TABLE_SIZE; ++i) {
        expTable[i] = expf((i / (float)EXP_TABLE_SIZE * 2 - 1) * MAX_EXP);
        expTable[i] = expTable[i] / (expTable[i] + 1);
    }
    expTable[EXP_TABLE_SIZE] = expTable[EXP_TABLE_SIZE - 1];
}

/* Index of the expTable entry for a logit f in (-6, 6]: the integer
 * division EXP_TABLE_SIZE / MAX_EXP / 2 is 83, not 83.33
 * (src/word2vec.cu:525). */
int orc_w2v_lut_index(float f) { return (int)((f + MAX_EXP) * (EXP_TABLE_SIZE / MAX_EXP / 2)); }

/* One negative draw of skip_gram_kernel (src/word2vec.cu:500-505): advance
 * the LCG, take table[(x >> 16) % table_size]; index 0 is replaced by
 * x % (vocab_size - 1) + 1.  Returns the vocabulary index. */
uint32_t orc_w2v_negative(uint64_t *next_random, const int32_t *table, uint32_t table_size, uint64_t vocab_size) {
    const unsigned long long x = (unsigned long long)*next_random * 25214903917ull + 11ull;
    *next_random = x;
    uint32_t tr = (uint32_t)table[(x >> 16) % table_size];
    if (tr == 0) tr = (uint32_t)(x % (vocab_size - 1) + 1);
    return tr;
}

#define D_W2V 0x70000000u

static float w2v_grad(float f, float label, float alpha, int use_lut, const float *expTable) {
    if (f > MAX_EXP) return (label - 1) * alpha;
    if (f < -MAX_EXP) return (label - 0) * alpha;
    if (use_lut) return (label - expTable[orc_w2v_lut_index(f)]) * alpha;
    return (label - 1.f / (1.f + expf(-f))) * alpha;
}

uint64_t orc_w2v_train(uint32_t V, uint32_t d, float *syn0, float *syn1, const uint64_t *offsets, const uint32_t *tokens,
                       uint64_t n_sent, const uint32_t *deg, const uint32_t *order, const uint32_t *rank,
                       const uint8_t *in_vocab, uint64_t vocab_size, uint64_t train_words, uint32_t window,
                       uint32_t negative, float sample, float alpha, int use_lut, int compat_break,
                       uint32_t table_size, int threads, uint64_t *pairs_out) {
    return orc_w2v_train_mode(V, d, syn0, syn1, offsets, tokens, n_sent, deg, order, rank, in_vocab, vocab_size,
                              train_words, window, negative, sample, alpha, use_lut, compat_break, table_size, threads,
                              pairs_out, 0, 0, 0);
}

uint64_t orc_w2v_train_mode(uint32_t V, uint32_t d, float *syn0, float *syn1, const uint64_t *offsets,
                            const uint32_t *tokens, uint64_t n_sent, const uint32_t *deg, const uint32_t *order,
                            const uint32_t *rank, const uint8_t *in_vocab, uint64_t vocab_size, uint64_t train_words,
                            uint32_t window, uint32_t negative, float sample, float alpha, int use_lut,
                            int compat_break, uint32_t table_size, int threads, uint64_t *pairs_out, int mode,
                            uint64_t seed, uint32_t round) {
    /* expTable, src/word2vec.cu:2292-2297 */
    float expTable[EXP_TABLE_SIZE + 1];
    orc_w2v_exp_table(expTable);
    /* unigram table over vocabulary ranks, src/word2vec.cu:674-694 */
    int32_t *table = (int32_t *)malloc((uint64_t)table_size * sizeof(int32_t));
    {
        uint32_t *cn = (uint32_t *)malloc((vocab_size ? vocab_size : 1) * sizeof(uint32_t));
        for (uint64_t a = 0; a < vocab_size; ++a) cn[a] = deg[order[a]];
        orc_w2v_unigram_table(vocab_size, cn, table_size, table);
        free(cn);
    }
    /* subsampling thresholds (u16) and the batching RNG, src/word2vec.cu:134-152,180-212 */
    uint16_t *thr16 = (uint16_t *)malloc((uint64_t)V * sizeof(uint16_t));
    const double sT = (double)sample * (double)train_words;
    for (uint32_t v = 0; v < V; ++v) {
        thr16[v] = 0xFFFF;
        if (sample > 0.f && deg[v] > 0) {
            double prob = (sqrt(deg[v] / sT) + 1.0) * (sT / (double)deg[v]);
            if (prob < 1.0) thr16[v] = (uint16_t)(prob * 65535.0 + 0.5);
        }
    }
    unsigned long long lcg = 0x9E3779B97F4A7C15ull ^ (1ull << 1) ^ 0x2ull;
    /* serial batching pass (src/word2vec.cu:1194-1248): filtered sentences,
     * capped at 1000 tokens; a batch closes at 15000 sentences or 299,980
     * tokens, and then draws its per-sentence negatives from the same LCG
     * (:1251-1259) whether or not they are used.  blk[s] = the sentence's
     * index in its batch, the kernel's blockIdx.x. */
    uint64_t nt = offsets[n_sent];
    uint32_t *sen = (uint32_t *)malloc((nt ? nt : 1) * sizeof(uint32_t));
    uint64_t *soff = (uint64_t *)malloc((n_sent + 1) * sizeof(uint64_t));
    uint32_t *blk = (uint32_t *)malloc((n_sent ? n_sent : 1) * sizeof(uint32_t));
    uint64_t words = 0, total = 0, batch_start = 0;
    uint32_t in_batch = 0;
    soff[0] = 0;
    for (uint64_t s = 0; s < n_sent; ++s) {
        blk[s] = in_batch;
        uint32_t len = 0;
        for (uint64_t i = offsets[s]; i < offsets[s + 1]; ++i) {
            const uint32_t x = tokens[i];
            ++words;
            if (!in_vocab[x]) {
                if (compat_break) break;
                continue;
            }
            if (compat_break && rank[x] == 0) break;
            if (thr16[x] < 0xFFFF) {
                lcg = lcg * 25214903917ull + 11ull;
                const uint16_t r16 = (uint16_t)((uint32_t)(lcg >> 16) >> 16);
                if (r16 > thr16[x]) continue;
            }
            sen[total++] = x;
            if (++len >= 1000) break;
        }
        soff[s + 1] = total;
        ++in_batch;
        if (in_batch == MAX_SENTENCE || total - batch_start >= (uint64_t)(MAX_SENTENCE - 1) * 20 || s + 1 == n_sent) {
            for (uint64_t q = 0; q < (uint64_t)in_batch * negative; ++q) lcg = lcg * 25214903917ull + 11ull;
            in_batch = 0;
            batch_start = total;
        }
    }
    uint64_t pairs = 0;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#endif
#pragma omp parallel reduction(+ : pairs)
    {
        float *neu1e = (float *)malloc((uint64_t)d * sizeof(float));
        float *neu1 = (float *)malloc((uint64_t)d * sizeof(float));
        uint32_t negs[64];
#pragma omp for schedule(dynamic, 64)
        for (int64_t s = 0; s < (int64_t)n_sent; ++s) {
            const uint64_t ss = soff[s], se = soff[s + 1];
            unsigned long long next_random = (unsigned long long)blk[s];
            if (mode == 1) {
                /* __sgNegReuse (src/word2vec.cu:368-440): one set of negatives
                 * per sentence, drawn from the unigram table; here the draw is
                 * Philox(seed; j, W2V, sentence, round) word 0 in place of the
                 * host LCG of :1251-1259 */
                for (uint32_t j = 0; j < negative && j < 64; ++j) {
                    uint32_t r[4];
                    rnd(seed, j, D_W2V, (uint32_t)s, round, r);
                    uint32_t tr = (uint32_t)table[r[0] % table_size];
                    if (tr == 0) tr = (uint32_t)(r[0] % (vocab_size - 1) + 1);
                    negs[j] = order[tr];
                }
            }
            for (uint64_t pos = ss; pos < se; ++pos) {
                const uint32_t word = sen[pos];
                uint32_t b = 0;
                if (mode != 1) {
                    next_random = next_random * 2514903917ull + 11;  /* src/word2vec.cu:458,555 (sic) */
                    b = (uint32_t)(next_random % window);
                }
                if (mode == 2) {
                    /* cbow_kernel (src/word2vec.cu:538-646): the mean of the
                     * context rows predicts the centre */
                    uint32_t cw = 0;
                    memset(neu1, 0, (uint64_t)d * sizeof(float));
                    for (uint32_t a = b; a < window * 2 + 1 - b; ++a) {
                        if (a == window) continue;
                        const int64_t c = (int64_t)pos - window + a;
                        if (c < (int64_t)ss || c >= (int64_t)se) continue;
                        const float *l1 = syn0 + (uint64_t)sen[c] * d;
                        for (uint32_t e = 0; e < d; ++e) neu1[e] += l1[e];
                        ++cw;
                    }
                    if (!cw) continue;
                    for (uint32_t e = 0; e < d; ++e) neu1[e] /= (float)cw;
                    memset(neu1e, 0, (uint64_t)d * sizeof(float));
                    pairs += cw;
                    for (uint32_t dd = 0; dd < negative + 1; ++dd) {
                        uint32_t target;
                        float label;
                        if (dd == 0) {
                            target = word;
                            label = 1.f;
                        } else {
                            uint64_t nr = next_random;
                            const uint32_t tr = orc_w2v_negative(&nr, table, table_size, vocab_size);
                            next_random = nr;
                            target = order[tr];
                            if (target == word) continue;
                            label = 0.f;
                        }
                        float *l2 = syn1 + (uint64_t)target * d;
                        float f = 0.f;
                        for (uint32_t e = 0; e < d; ++e) f += neu1[e] * l2[e];
                        const float g = w2v_grad(f, label, alpha, use_lut, expTable);
                        for (uint32_t e = 0; e < d; ++e) neu1e[e] += g * l2[e];
                        for (uint32_t e = 0; e < d; ++e) l2[e] += g * neu1[e];
                    }
                    for (uint32_t a = b; a < window * 2 + 1 - b; ++a) {
                        if (a == window) continue;
                        const int64_t c = (int64_t)pos - window + a;
                        if (c < (int64_t)ss || c >= (int64_t)se) continue;
                        float *l1 = syn0 + (uint64_t)sen[c] * d;
                        for (uint32_t e = 0; e < d; ++e) l1[e] += neu1e[e];
                    }
                    continue;
                }
                for (uint32_t a = b; a < window * 2 + 1 - b; ++a) {
                    if (a == window) continue;
                    const int64_t c = (int64_t)pos - window + a;
                    if (c < (int64_t)ss || c >= (int64_t)se) continue;
                    const uint32_t last = sen[c];
                    float *l1 = syn0 + (uint64_t)last * d;
                    memset(neu1e, 0, (uint64_t)d * sizeof(float));
                    ++pairs;
                    for (uint32_t dd = 0; dd < negative + 1; ++dd) {
                        uint32_t target;
                        float label;
                        if (dd == 0) {
                            target = word;
                            label = 1.f;
                        } else if (mode == 1) {
                            target = negs[dd - 1];
                            if (target == word) continue;
                            label = 0.f;
                        } else {
                            uint64_t nr = next_random;
                            const uint32_t tr = orc_w2v_negative(&nr, table, table_size, vocab_size);
                            next_random = nr;
                            target = order[tr];
                            if (target == word) continue;
                            label = 0.f;
                        }
                        float *l2 = syn1 + (uint64_t)target * d;
                        float f = 0.f;
                        for (uint32_t e = 0; e < d; ++e) f += l1[e] * l2[e];
                        const float g = w2v_grad(f, label, alpha, use_lut, expTable);
                        for (uint32_t e = 0; e < d; ++e) neu1e[e] += g * l2[e];
                        for (uint32_t e = 0; e < d; ++e) l2[e] += g * l1[e];
                    }
                    for (uint32_t e = 0; e < d; ++e) l1[e] += neu1e[e];
                }
            }
        }
        free(neu1e);
        free(neu1);
    }
    if (pairs_out) *pairs_out = pairs;
    free(table); free(thr16); free(sen); free(soff); free(blk);
    return words;
}

/* ---- convergence evaluation ------------------------------------------ */
/* Neighbour columns scored for vertex v (src/word2vec.cu:1785-1800 takes
 * at most 30 random neighbours): all of them when deg <= 30, else 30
 * distinct columns by Floyd's sampling (Bentley & Floyd, CACM 30(9),
 * 1987): for j = deg-30 .. deg-1 draw t uniform in [0, j] (Philox word 0
 * mod j+1) and take t, or j when t was already taken. */
uint32_t orc_eval_subset(uint32_t deg, uint32_t v, uint32_t round, uint64_t seed, uint32_t *chosen) {
    if (deg <= 30) {
        for (uint32_t i = 0; i < deg; ++i) chosen[i] = i;
        return deg;
    }
    for (uint32_t i = 0; i < 30; ++i) {
        const uint32_t j = deg - 30 + i;
        uint32_t r[4];
        rnd(seed, v, D_EVAL | i, 0xFFFFFFFFu, round, r);
        const uint32_t t = r[0] % (j + 1);
        int dup = 0;
        for (uint32_t q = 0; q < i; ++q) dup |= chosen[q] == t;
        chosen[i] = dup ? j : t;
    }
    return 30;
}

uint32_t orc_eval(uint32_t V, uint32_t d, const float *syn0, const uint64_t *row_ptr, const uint32_t *col,
                  const uint8_t *in_vocab, uint8_t *flags, float *scores, uint32_t round, uint64_t seed) {
    uint32_t eva = 0;
    for (uint32_t v = 0; v < V; ++v) {
        if (!in_vocab[v] || flags[v]) continue;
        const uint64_t b = row_ptr[v];
        const uint32_t deg = (uint32_t)(row_ptr[v + 1] - b);
        float score = 0.f;
        if (deg > 0) {
            uint32_t chosen[30];
            const uint32_t cnt = orc_eval_subset(deg, v, round, seed, chosen);
            const float *x = syn0 + (uint64_t)v * d;
            float nn = 0.f;
            for (uint32_t e = 0; e < d; ++e) nn += x[e] * x[e];
            const float nv = sqrtf(nn);
            float sum = 0.f;
            for (uint32_t i = 0; i < cnt; ++i) {
                const float *y = syn0 + (uint64_t)col[b + chosen[i]] * d;
                float dot = 0.f, mm = 0.f;
                for (uint32_t e = 0; e < d; ++e) { dot += x[e] * y[e]; mm += y[e] * y[e]; }
                const float nu = sqrtf(mm);
                sum += (nv == 0.f || nu == 0.f) ? 0.f : dot / (nv * nu);
            }
            score = sum / (float)cnt;
        }
        scores[v] = score;
        if (score > 0.65f) flags[v] = 1;
        else ++eva;
    }
    return eva;
}

/* ---- common neighbours, src/tools/gcn.cpp:304-357 ----------------------
 * Each row's neighbours sorted (the reference's sort_csr), the two rows
 * merged; the count skips repeated entries and the pair's own vertices. */
static int cmp_u32(const void *a, const void *b) {
    const uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
    return (x > y) - (x < y);
}

void orc_common_neighbours(uint32_t V, const uint64_t *row_ptr, const uint32_t *col, uint64_t m, const uint32_t *u,
                           const uint32_t *v, int32_t *cn) {
    const uint64_t E = row_ptr[V];
    uint32_t *s = (uint32_t *)malloc((E ? E : 1) * sizeof(uint32_t));
    memcpy(s, col, E * sizeof(uint32_t));
    for (uint32_t x = 0; x < V; ++x) qsort(s + row_ptr[x], row_ptr[x + 1] - row_ptr[x], sizeof(uint32_t), cmp_u32);
    for (uint64_t i = 0; i < m; ++i) {
        const uint32_t a = u[i], b = v[i];
        uint64_t p = row_ptr[a], pe = row_ptr[a + 1], q = row_ptr[b], qe = row_ptr[b + 1];
        int32_t c = 0;
        int64_t last = -1;
        while (p < pe && q < qe) {
            if (s[p] < s[q]) ++p;
            else if (s[p] > s[q]) ++q;
            else {
                const uint32_t x = s[p];
                if (x != a && x != b && (int64_t)x != last) ++c;
                last = x;
                ++p;
                ++q;
            }
        }
        cn[i] = c;
    }
    free(s);
}

/* ---- -classes: k-means, src/word2vec.cu:2128-2175 ----------------------
 * Rows in vocabulary order (row a = syn0 row of vertex order[a]); float
 * arithmetic in the reference's order. */
void orc_kmeans_classes(uint32_t d, const float *syn0, const uint32_t *order, uint64_t vocab_size, uint32_t classes,
                        uint32_t iter, int32_t *cl) {
    int *centcn = (int *)malloc(classes * sizeof(int));
    float *cent = (float *)calloc((uint64_t)classes * d, sizeof(float));
    for (uint64_t a = 0; a < vocab_size; ++a) cl[a] = (int32_t)(a % classes);
    for (uint32_t it = 0; it < iter; ++it) {
        for (uint64_t b = 0; b < (uint64_t)classes * d; ++b) cent[b] = 0;
        for (uint32_t b = 0; b < classes; ++b) centcn[b] = 1;
        for (uint64_t c = 0; c < vocab_size; ++c) {
            const float *x = syn0 + (uint64_t)order[c] * d;
            for (uint32_t e = 0; e < d; ++e) cent[(uint64_t)d * cl[c] + e] += x[e];
            centcn[cl[c]]++;
        }
        for (uint32_t b = 0; b < classes; ++b) {
            float closev = 0;
            for (uint32_t c = 0; c < d; ++c) {
                cent[(uint64_t)d * b + c] /= centcn[b];
                closev += cent[(uint64_t)d * b + c] * cent[(uint64_t)d * b + c];
            }
            closev = sqrtf(closev);
            for (uint32_t c = 0; c < d; ++c) cent[(uint64_t)d * b + c] /= closev;
        }
        for (uint64_t c = 0; c < vocab_size; ++c) {
            const float *x = syn0 + (uint64_t)order[c] * d;
            float closev = -10;
            int32_t closeid = 0;
            for (uint32_t k = 0; k < classes; ++k) {
                float v = 0;
                for (uint32_t e = 0; e < d; ++e) v += cent[(uint64_t)d * k + e] * x[e];
                if (v > closev) {
                    closev = v;
                    closeid = (int32_t)k;
                }
            }
            cl[c] = closeid;
        }
    }
    free(centcn);
    free(cent);
}

/* ---- AUC -------------------------------------------------------------- */
typedef struct { float s; int pos; } scored_t;
static int by_score(const void *a, const void *b) {
    const float x = ((const scored_t *)a)->s, y = ((const scored_t *)b)->s;
    return (x > y) - (x < y);
}

double orc_auc(uint64_t n_pos, const float *pos, uint64_t n_neg, const float *neg) {
    const uint64_t n = n_pos + n_neg;
    if (n_pos == 0 || n_neg == 0) return 0.5;
    scored_t *a = (scored_t *)malloc(n * sizeof(scored_t));
    for (uint64_t i = 0; i < n_pos; ++i) { a[i].s = pos[i]; a[i].pos = 1; }
    for (uint64_t i = 0; i < n_neg; ++i) { a[n_pos + i].s = neg[i]; a[n_pos + i].pos = 0; }
    qsort(a, n, sizeof(scored_t), by_score);
    double rank_sum = 0.0;
    uint64_t i = 0;
    while (i < n) {
        uint64_t j = i;
        while (j + 1 < n && a[j + 1].s == a[i].s) ++j;
        const double r = 0.5 * (double)(i + j) + 1.0;  /* average 1-based rank */
        for (uint64_t q = i; q <= j; ++q)
            if (a[q].pos) rank_sum += r;
        i = j + 1;
    }
    free(a);
    return (rank_sum - (double)n_pos * (double)(n_pos + 1) / 2.0) / ((double)n_pos * (double)n_neg);
}
