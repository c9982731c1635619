// budget_kernels.hip - the learned walk budget and the visit diagnostics.
//
// The plain information-centric rule decides, per vertex, in how many rounds
// a walker starts there (every round until the vertex converges,
// src/word2vec.cu:2016-2080) and how long each walk runs (the entropy /
// Pearson stop rule, include/walk.hpp:945-951,1056-1062).  The learned
// budget predicts both from eight per-vertex features with an 8-64-64-2 MLP
// (ReLU), trained on the budgets the plain rule produced on other graphs:
//   features   one wavefront per vertex over its CSR row; rows longer than
//              kHubChunk entries in chunks of kHubChunk (one wavefront each,
//              partials combined per hub in chunk order: deterministic)
//   training   mean squared error on standardised log targets; a 64-sample
//              workgroup runs forward and backward with its activations in
//              LDS, reduces the weight gradients over its samples and adds
//              them to the global gradient once; then an Adam step
//   inference  one thread per vertex, weights in LDS; a vertex whose
//              standardised features leave the training set's box keeps the
//              plain rule (budget 0)
//   walks      round r starts the walker of v when r <= count(v) (plain-rule
//              vertices: when v is not flagged) and it runs exactly length(v)
//              vertices, without the FP64 entropy state
// Quality measures: the corpus visit histogram, its Shannon entropy and the
// visit-frequency KL diagnostic of include/walk.hpp:754-788.
#include <hip/hip_runtime.h>
#include "felog_dev.h"
#include "felog_internal.h"

namespace felog {

namespace {

constexpr int kF = FELOG_BUDGET_FEATURES;  // 8
constexpr int kH = 64;
constexpr int kO = 2;
// parameter layout (row-major weights, out x in)
constexpr int oW1 = 0;
constexpr int oB1 = oW1 + kH * kF;
constexpr int oW2 = oB1 + kH;
constexpr int oB2 = oW2 + kH * kH;
constexpr int oW3 = oB2 + kH;
constexpr int oB3 = oW3 + kO * kH;
static_assert(oB3 + kO == FELOG_BUDGET_PARAMS, "parameter count");

__device__ __forceinline__ float wave_sum(float x) {
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    return x;
}
__device__ __forceinline__ float wave_max(float x) {
    for (int o = 32; o > 0; o >>= 1) x = fmaxf(x, __shfl_xor(x, o));
    return x;
}

// Features of vertex v (raw, before standardisation):
//   0 ln(1 + deg)        1 ln(1 + sum w)      2 max w      3 mean w
//   4 ln(1 + sum of the neighbours' degrees)  (two-hop degree estimate)
//   5 ln(1 + max neighbour degree)
//   6 entropy of the transition distribution w / sum w over ln deg (0 if
//     deg <= 1)
//   7 1 if v is in the vocabulary (its tokens are trained)
// Four vertices per wavefront, 16 lanes each (most rows are short); the
// entropy in one pass as ln S - (sum w ln w) / S.  Rows longer than
// kHubChunk are left to k_feat_chunk / k_feat_hub.
__device__ __forceinline__ float group_sum16(float x) {
    for (int o = 8; o > 0; o >>= 1) x += __shfl_xor(x, o);
    return x;
}
__device__ __forceinline__ float group_max16(float x) {
    for (int o = 8; o > 0; o >>= 1) x = fmaxf(x, __shfl_xor(x, o));
    return x;
}

__global__ __launch_bounds__(256) void k_features(const VertexRec *__restrict__ vrec,
                                                  const uint32_t *__restrict__ col, const float *__restrict__ w,
                                                  const uint8_t *__restrict__ in_vocab, uint32_t V,
                                                  float *__restrict__ out) {
    const uint32_t lane = threadIdx.x & 63, grp = lane >> 4, l = lane & 15;
    const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    for (uint64_t base = (((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6) * 4; base < V; base += nw * 4) {
        const uint64_t v = base + grp;
        VertexRec vr{0, 0};
        if (v < V) vr = vrec[v];
        const bool mine = v < V && vr.deg <= kHubChunk;
        float ws = 0.f, wm = 0.f, nd = 0.f, ndm = 0.f, wl = 0.f;
        if (mine) {
            for (uint32_t j = l; j < vr.deg; j += 16) {
                const float x = w[vr.begin + j];
                const float du = (float)vrec[col[vr.begin + j]].deg;
                ws += x;
                wm = fmaxf(wm, x);
                nd += du;
                ndm = fmaxf(ndm, du);
                if (x > 0.f) wl += x * __logf(x);
            }
        }
        ws = group_sum16(ws);
        wm = group_max16(wm);
        nd = group_sum16(nd);
        ndm = group_max16(ndm);
        wl = group_sum16(wl);
        if (mine && l == 0) {
            const float h = (vr.deg > 1 && ws > 0.f) ? (__logf(ws) - wl / ws) / __logf((float)vr.deg) : 0.f;
            float *f = out + v * kF;
            f[0] = log1pf((float)vr.deg);
            f[1] = log1pf(ws);
            f[2] = wm;
            f[3] = vr.deg ? ws / (float)vr.deg : 0.f;
            f[4] = log1pf(nd);
            f[5] = log1pf(ndm);
            f[6] = h;
            f[7] = in_vocab[v] ? 1.f : 0.f;
        }
    }
}

// Hub rows: chunk t sums (w, max w, neighbour degree, max neighbour degree,
// w ln w) over its kHubChunk entries into part[8t ..].  The entropy of a hub
// row is ln S - (sum w ln w) / S, so one pass over the row suffices.
constexpr int kPart = 8;
__global__ __launch_bounds__(256) void k_feat_chunk(const VertexRec *__restrict__ vrec,
                                                    const uint32_t *__restrict__ col, const float *__restrict__ w,
                                                    const uint32_t *__restrict__ hub,
                                                    const uint64_t *__restrict__ cstart, uint32_t nh,
                                                    float *__restrict__ part) {
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t nchunks = cstart[nh];
    const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    for (uint64_t t = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; t < nchunks; t += nw) {
        const uint32_t k = hub_of_chunk(cstart, nh, t);
        const VertexRec vr = vrec[hub[k]];
        const uint64_t c0 = (t - cstart[k]) * kHubChunk;
        const uint64_t cn = vr.deg - c0 < kHubChunk ? vr.deg - c0 : kHubChunk;
        const uint64_t b = vr.begin + c0;
        float ws = 0.f, wm = 0.f, nd = 0.f, ndm = 0.f, wl = 0.f;
        for (uint64_t j = lane; j < cn; j += 64) {
            const float x = w[b + j];
            const float du = (float)vrec[col[b + j]].deg;
            ws += x;
            wm = fmaxf(wm, x);
            nd += du;
            ndm = fmaxf(ndm, du);
            if (x > 0.f) wl += x * __logf(x);
        }
        ws = wave_sum(ws);
        wm = wave_max(wm);
        nd = wave_sum(nd);
        ndm = wave_max(ndm);
        wl = wave_sum(wl);
        if (lane == 0) {
            float *o = part + t * kPart;
            o[0] = ws;
            o[1] = wm;
            o[2] = nd;
            o[3] = ndm;
            o[4] = wl;
        }
    }
}

// One wavefront per hub: its chunk partials (lane l takes chunks l, l + 64,
// ... in order) and the features.
__global__ __launch_bounds__(256) void k_feat_hub(const VertexRec *__restrict__ vrec,
                                                  const uint32_t *__restrict__ hub,
                                                  const uint64_t *__restrict__ cstart, uint32_t nh,
                                                  const float *__restrict__ part,
                                                  const uint8_t *__restrict__ in_vocab, float *__restrict__ out) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t nw = (gridDim.x * blockDim.x) >> 6;
    for (uint32_t k = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; k < nh; k += nw) {
        float ws = 0.f, wm = 0.f, nd = 0.f, ndm = 0.f, wl = 0.f;
        for (uint64_t t = cstart[k] + lane; t < cstart[k + 1]; t += 64) {
            const float *p = part + t * kPart;
            ws += p[0];
            wm = fmaxf(wm, p[1]);
            nd += p[2];
            ndm = fmaxf(ndm, p[3]);
            wl += p[4];
        }
        ws = wave_sum(ws);
        wm = wave_max(wm);
        nd = wave_sum(nd);
        ndm = wave_max(ndm);
        wl = wave_sum(wl);
        if (lane == 0) {
            const uint32_t v = hub[k];
            const uint32_t deg = vrec[v].deg;
            float *f = out + (uint64_t)v * kF;
            f[0] = log1pf((float)deg);
            f[1] = log1pf(ws);
            f[2] = wm;
            f[3] = ws / (float)deg;
            f[4] = log1pf(nd);
            f[5] = log1pf(ndm);
            f[6] = ws > 0.f ? (logf(ws) - wl / ws) / logf((float)deg) : 0.f;
            f[7] = in_vocab[v] ? 1.f : 0.f;
        }
    }
}

// Gradient of the mean squared error over n samples (X standardised
// features, T standardised targets), accumulated into grad; 64 samples per
// workgroup.  sample_seed != 0 draws the block's samples with replacement
// from Philox (stochastic minibatch), else the block takes samples
// 64*block .. 64*block+63.
__global__ __launch_bounds__(64) void k_mlp_grad(const float *__restrict__ X, const float *__restrict__ T,
                                                 uint64_t n, uint64_t sample_seed, uint32_t step,
                                                 const float *__restrict__ P, float *__restrict__ grad,
                                                 double *__restrict__ loss, float inv_n) {
    __shared__ float sp[FELOG_BUDGET_PARAMS];
    __shared__ float h1[64][kH + 1], h2[64][kH + 1], d1[64][kH + 1], d2[64][kH + 1];
    __shared__ float sx[64][kF], sdy[64][kO];
    const uint32_t t = threadIdx.x;
    for (uint32_t i = t; i < (uint32_t)FELOG_BUDGET_PARAMS; i += 64) sp[i] = P[i];

    uint64_t s = (uint64_t)blockIdx.x * 64 + t;
    if (sample_seed) {
        const u32x4 r = draw(sample_seed, blockIdx.x * 64 + t, DOM_MLP, 0u, step);
        s = (((uint64_t)r.y << 32) | r.x) % n;
    }
    const bool valid = s < n;
    float x[kF], tg[kO] = {0.f, 0.f};
#pragma unroll
    for (int c = 0; c < kF; ++c) {
        x[c] = valid ? X[s * kF + c] : 0.f;
        sx[t][c] = x[c];
    }
    if (valid) { tg[0] = T[s * kO]; tg[1] = T[s * kO + 1]; }
    __syncthreads();

    for (int j = 0; j < kH; ++j) {
        float z = sp[oB1 + j];
#pragma unroll
        for (int c = 0; c < kF; ++c) z += sp[oW1 + j * kF + c] * x[c];
        h1[t][j] = fmaxf(z, 0.f);
    }
    float y0 = sp[oB3], y1 = sp[oB3 + 1];
    for (int j = 0; j < kH; ++j) {
        float z = sp[oB2 + j];
        for (int i = 0; i < kH; ++i) z += sp[oW2 + j * kH + i] * h1[t][i];
        const float h = fmaxf(z, 0.f);
        h2[t][j] = h;
        y0 += sp[oW3 + j] * h;
        y1 += sp[oW3 + kH + j] * h;
    }
    const float e0 = valid ? y0 - tg[0] : 0.f, e1 = valid ? y1 - tg[1] : 0.f;
    const float dy0 = 2.f * e0 * inv_n, dy1 = 2.f * e1 * inv_n;
    sdy[t][0] = dy0;
    sdy[t][1] = dy1;
    for (int j = 0; j < kH; ++j)
        d2[t][j] = h2[t][j] > 0.f ? sp[oW3 + j] * dy0 + sp[oW3 + kH + j] * dy1 : 0.f;
    for (int i = 0; i < kH; ++i) {
        float g = 0.f;
        for (int j = 0; j < kH; ++j) g += sp[oW2 + j * kH + i] * d2[t][j];
        d1[t][i] = h1[t][i] > 0.f ? g : 0.f;
    }
    const float l = wave_sum(e0 * e0 + e1 * e1);
    if (t == 0) atomicAdd(loss, (double)l);
    __syncthreads();

    // workgroup reduction over the 64 samples: thread j owns hidden unit j
    const uint32_t j = t;
    float gb1 = 0.f, gb2 = 0.f;
    float gw1[kF];
#pragma unroll
    for (int c = 0; c < kF; ++c) gw1[c] = 0.f;
    float gw3a = 0.f, gw3b = 0.f;
    for (int q = 0; q < 64; ++q) {
        gb1 += d1[q][j];
        gb2 += d2[q][j];
#pragma unroll
        for (int c = 0; c < kF; ++c) gw1[c] += d1[q][j] * sx[q][c];
        gw3a += sdy[q][0] * h2[q][j];
        gw3b += sdy[q][1] * h2[q][j];
    }
    atomicAdd(&grad[oB1 + j], gb1);
    atomicAdd(&grad[oB2 + j], gb2);
#pragma unroll
    for (int c = 0; c < kF; ++c) atomicAdd(&grad[oW1 + j * kF + c], gw1[c]);
    atomicAdd(&grad[oW3 + j], gw3a);
    atomicAdd(&grad[oW3 + kH + j], gw3b);
    for (int i = 0; i < kH; ++i) {
        float g = 0.f;
        for (int q = 0; q < 64; ++q) g += d2[q][j] * h1[q][i];
        atomicAdd(&grad[oW2 + j * kH + i], g);
    }
    if (j < (uint32_t)kO) {
        float g = 0.f;
        for (int q = 0; q < 64; ++q) g += sdy[q][j];
        atomicAdd(&grad[oB3 + j], g);
    }
}

// Adam (Kingma & Ba, 2015) on the flat parameters; clears the gradient.
__global__ void k_adam(float *__restrict__ P, float *__restrict__ grad, float *__restrict__ m, float *__restrict__ v,
                       float lr, float b1, float b2, float eps, float c1, float c2) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint32_t)FELOG_BUDGET_PARAMS) return;
    const float g = grad[i];
    const float mi = b1 * m[i] + (1.f - b1) * g;
    const float vi = b2 * v[i] + (1.f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    P[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    grad[i] = 0.f;
}

// Budgets of every vertex from its raw features and the model (parameters,
// feature standardisation and box, target standardisation).  Outputs 0 for
// vertices outside the box (plain rule).
__global__ __launch_bounds__(256) void k_mlp_predict(const float *__restrict__ F, uint32_t V,
                                                     const felog_budget_model *__restrict__ M, uint32_t max_len,
                                                     uint32_t max_count, uint16_t *__restrict__ cnt,
                                                     uint16_t *__restrict__ len, uint32_t *__restrict__ sum) {
    // The weights are read with wave-uniform indices straight from the
    // model (scalar loads into SGPR operands of the FMAs): staged in LDS,
    // every FMA's broadcast read returned a kilobyte per wavefront and the
    // kernel was bound by that.
    const float *__restrict__ sp = M->params;
    __shared__ float st[6 * kF + 2 * kO];
    if (threadIdx.x < (uint32_t)kF) {
        st[threadIdx.x] = M->feat_mean[threadIdx.x];
        st[kF + threadIdx.x] = M->feat_std[threadIdx.x];
        st[2 * kF + threadIdx.x] = M->feat_lo[threadIdx.x];
        st[3 * kF + threadIdx.x] = M->feat_hi[threadIdx.x];
    }
    if (threadIdx.x < (uint32_t)kO) {
        st[4 * kF + threadIdx.x] = M->targ_mean[threadIdx.x];
        st[4 * kF + kO + threadIdx.x] = M->targ_std[threadIdx.x];
    }
    __syncthreads();
    const uint32_t stride = gridDim.x * blockDim.x;
    uint32_t mx = 0, n_plain = 0, n_plain_vocab = 0;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += stride) {
        float x[kF];
        bool inside = true;
#pragma unroll
        for (int c = 0; c < kF; ++c) {
            const float sd = st[kF + c];
            x[c] = sd > 0.f ? (F[(uint64_t)v * kF + c] - st[c]) / sd : 0.f;
            inside = inside && x[c] >= st[2 * kF + c] && x[c] <= st[3 * kF + c];
        }
        float h1[kH];
        for (int j = 0; j < kH; ++j) {
            float z = sp[oB1 + j];
#pragma unroll
            for (int c = 0; c < kF; ++c) z += sp[oW1 + j * kF + c] * x[c];
            h1[j] = fmaxf(z, 0.f);
        }
        float y0 = sp[oB3], y1 = sp[oB3 + 1];
        for (int j = 0; j < kH; ++j) {
            // four independent partial sums: a 16-deep dependency chain per
            // output instead of 64
            float z[4] = {sp[oB2 + j], 0.f, 0.f, 0.f};
#pragma unroll
            for (int i = 0; i < kH; i += 4)
#pragma unroll
                for (int q = 0; q < 4; ++q) z[q] += sp[oW2 + j * kH + i + q] * h1[i + q];
            const float h = fmaxf((z[0] + z[1]) + (z[2] + z[3]), 0.f);
            y0 += sp[oW3 + j] * h;
            y1 += sp[oW3 + kH + j] * h;
        }
        // targets: ln(count), ln(length)
        const float c_ = __expf(y0 * st[4 * kF + kO] + st[4 * kF]);
        const float l_ = __expf(y1 * st[4 * kF + kO + 1] + st[4 * kF + 1]);
        uint32_t cv = (uint32_t)fminf(fmaxf(rintf(c_), 1.f), (float)max_count);
        uint32_t lv = (uint32_t)fminf(fmaxf(rintf(l_), 2.f), (float)max_len);
        cnt[v] = inside ? (uint16_t)cv : 0;
        len[v] = inside ? (uint16_t)lv : 0;
        if (inside) mx = cv > mx ? cv : mx;
        n_plain += inside ? 0u : 1u;
        n_plain_vocab += (!inside && F[(uint64_t)v * kF + 7] != 0.f) ? 1u : 0u;
    }
    for (int o = 32; o > 0; o >>= 1) {
        mx = max(mx, (uint32_t)__shfl_xor((int)mx, o));
        n_plain += (uint32_t)__shfl_xor((int)n_plain, o);
        n_plain_vocab += (uint32_t)__shfl_xor((int)n_plain_vocab, o);
    }
    __shared__ uint32_t part[3][4];
    auto add = [](uint32_t p, uint32_t q) { return p + q; };
    mx = block_combine<4>(mx, part[0], [](uint32_t p, uint32_t q) { return p > q ? p : q; });
    n_plain = block_combine<4>(n_plain, part[1], add);
    n_plain_vocab = block_combine<4>(n_plain_vocab, part[2], add);
    if (threadIdx.x == 0) {
        atomicMax(&sum[0], mx);
        atomicAdd(&sum[1], n_plain);
        atomicAdd(&sum[2], n_plain_vocab);
    }
}

// Walkers of round r: walker i starts at i % V; budgeted vertices walk in
// rounds 1..count, the others while not flagged.
__global__ void k_walk_ok_budget(const uint8_t *__restrict__ flags, const uint16_t *__restrict__ cnt,
                                 uint32_t walkers, uint32_t V, const RoundState *__restrict__ rs,
                                 uint32_t *__restrict__ ok) {
    const uint32_t round = rs->round;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < walkers; i += stride) {
        const uint32_t v = i % V;
        ok[i] = cnt[v] ? (round <= cnt[v] ? 1 : 0) : (flags[v] ? 0 : 1);
    }
}

// Visit counts, indexed by vocabulary rank (vertices by degree, descending).
// A walk visits a vertex about in proportion to its degree, so the first
// kVisitHubs ranks take most tokens: those are counted in an LDS histogram
// per workgroup and added once at its end (global atomics on the same few
// addresses serialise), the rest go straight to memory.  n_dev (when set)
// holds the token count instead of n.
constexpr uint32_t kVisitHubs = 12288;
constexpr uint32_t kVisitBlock = 1024;
__global__ __launch_bounds__(kVisitBlock) void k_visits(const uint32_t *__restrict__ tok, uint64_t n,
                                                        const uint64_t *__restrict__ n_dev,
                                                        const uint32_t *__restrict__ rank,
                                                        uint32_t *__restrict__ visits) {
    __shared__ uint32_t hub[kVisitHubs];
    for (uint32_t i = threadIdx.x; i < kVisitHubs; i += blockDim.x) hub[i] = 0;
    __syncthreads();
    if (n_dev) n = *n_dev;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint32_t r = rank[tok[i]];
        if (r < kVisitHubs) atomicAdd(&hub[r], 1u);
        else atomicAdd(&visits[r], 1u);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < kVisitHubs; i += blockDim.x)
        if (hub[i]) atomicAdd(&visits[i], hub[i]);
}

// out[0] += sum_i p_i (ln p_i - ln q_i) over vertices with p_i, q_i > 0
// (p = degree share, q = visit share); out[1] += -sum q ln q;
// out[2] += vertices with p_i > 0 and no visit.
__global__ __launch_bounds__(256) void k_visit_stats(const uint32_t *__restrict__ visits,
                                                     const uint32_t *__restrict__ rank,
                                                     const uint32_t *__restrict__ deg, uint32_t V, double inv_vis,
                                                     double inv_deg, double *__restrict__ out) {
    double kl = 0.0, ent = 0.0, unv = 0.0;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < V; i += stride) {
        const double p = (double)deg[i] * inv_deg, q = (double)visits[rank[i]] * inv_vis;
        if (q > 0.0) ent -= q * log(q);
        if (p > 0.0) {
            if (q > 0.0) kl += p * (log(p) - log(q));
            else unv += 1.0;
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        kl += __shfl_xor(kl, o);
        ent += __shfl_xor(ent, o);
        unv += __shfl_xor(unv, o);
    }
    __shared__ double part[3][4];
    auto add = [](double p, double q) { return p + q; };
    kl = block_combine<4>(kl, part[0], add);
    ent = block_combine<4>(ent, part[1], add);
    unv = block_combine<4>(unv, part[2], add);
    if (threadIdx.x == 0) {
        atomicAdd(&out[0], kl);
        atomicAdd(&out[1], ent);
        atomicAdd(&out[2], unv);
    }
}

inline unsigned grid_for(uint64_t n, unsigned block, unsigned cap) {
    uint64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (unsigned)g;
}

}  // namespace

#define HIP_TRY(x)                              \
    do {                                        \
        hipError_t e_ = (x);                    \
        if (e_ != hipSuccess) return e_;        \
    } while (0)

hipError_t budget_features(felog_ctx *c, float *d_out) {
    k_features<<<grid_for((uint64_t)c->V * 64, 256, 1u << 16), 256, 0, c->stream>>>(c->d_vrec, c->d_col, c->d_w,
                                                                                      c->d_in_vocab, c->V, d_out);
    HIP_TRY(hipGetLastError());
    const HubPlan &h = c->hubs;
    if (h.n == 0) return hipSuccess;
    if (!c->d_feat_part) HIP_TRY(hipMalloc(&c->d_feat_part, h.chunks * kPart * 4));
    k_feat_chunk<<<grid_for(h.chunks * 64, 256, 1u << 16), 256, 0, c->stream>>>(c->d_vrec, c->d_col, c->d_w, h.hub,
                                                                                 h.cstart, h.n, c->d_feat_part);
    HIP_TRY(hipGetLastError());
    k_feat_hub<<<grid_for((uint64_t)h.n * 64, 256, 1u << 16), 256, 0, c->stream>>>(c->d_vrec, h.hub, h.cstart, h.n,
                                                                                    c->d_feat_part, c->d_in_vocab,
                                                                                    d_out);
    return hipGetLastError();
}

hipError_t budget_grad(felog_ctx *c, const float *d_x, const float *d_t, uint64_t n, uint64_t sample_seed,
                       uint32_t step, uint32_t blocks, const float *d_p, float *d_grad, double *d_loss,
                       float inv_n) {
    k_mlp_grad<<<blocks, 64, 0, c->stream>>>(d_x, d_t, n, sample_seed, step, d_p, d_grad, d_loss, inv_n);
    return hipGetLastError();
}

hipError_t budget_adam(felog_ctx *c, float *d_p, float *d_grad, float *d_m, float *d_v, float lr, uint32_t t) {
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - powf(b1, (float)t), c2 = 1.f - powf(b2, (float)t);
    k_adam<<<grid_for(FELOG_BUDGET_PARAMS, 256, 64), 256, 0, c->stream>>>(d_p, d_grad, d_m, d_v, lr, b1, b2, 1e-8f,
                                                                         c1, c2);
    return hipGetLastError();
}

hipError_t budget_predict(felog_ctx *c, const float *d_feat, const felog_budget_model *d_model, uint32_t *d_sum) {
    HIP_TRY(hipMemsetAsync(d_sum, 0, 16, c->stream));
    k_mlp_predict<<<grid_for(c->V, 256, c->cu_count * 4), 256, 0, c->stream>>>(d_feat, c->V, d_model, c->opt.max_walk_len,
                                                                     kBudgetMaxCount, c->d_budget_cnt,
                                                                     c->d_budget_len, d_sum);
    return hipGetLastError();
}

hipError_t budget_walk_ok(felog_ctx *c, uint32_t *d_ok) {
    k_walk_ok_budget<<<grid_for(c->walkers, 256, 65535), 256, 0, c->stream>>>(c->d_flags, c->d_budget_cnt,
                                                                             c->walkers, c->V, c->d_rs,
                                                                             d_ok);
    return hipGetLastError();
}

hipError_t add_visits(felog_ctx *c, const uint32_t *d_tok, uint64_t n_tok) {
    if (n_tok == 0) return hipSuccess;
    k_visits<<<grid_for(n_tok, kVisitBlock, c->cu_count * 2), kVisitBlock, 0, c->stream>>>(
        d_tok, n_tok, nullptr, c->d_vocab_rank, c->d_visits);
    HIP_TRY(hipGetLastError());
    c->visits_total += n_tok;
    return hipSuccess;
}

// The caller adds the round's token count to visits_total once it is known.
hipError_t add_round_visits(felog_ctx *c) {
    k_visits<<<c->cu_count * 2, kVisitBlock, 0, c->stream>>>(c->d_tok, 0, &c->d_rs->n_tokens, c->d_vocab_rank,
                                                            c->d_visits);
    return hipGetLastError();
}

hipError_t visit_stats(felog_ctx *c, double *d_out) {
    HIP_TRY(hipMemsetAsync(d_out, 0, 3 * sizeof(double), c->stream));
    const double inv_vis = c->visits_total ? 1.0 / (double)c->visits_total : 0.0;
    const double inv_deg = c->E ? 1.0 / (double)c->E : 0.0;
    k_visit_stats<<<grid_for(c->V, 256, c->cu_count * 4), 256, 0, c->stream>>>(c->d_visits, c->d_vocab_rank, c->d_deg, c->V,
                                                                     inv_vis, inv_deg,
                                                                     d_out);
    return hipGetLastError();
}

}  // namespace felog
