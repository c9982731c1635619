// train_noalloc.hip - felog_train_corpus without device allocation.
//
// felog_train_corpus stages the caller's corpus in freshly allocated device
// buffers and grows the context's SGNS tile-descriptor buffer when a corpus
// needs more descriptors.  Callers that cannot allocate device memory (code
// under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc) use felog_train_corpus_noalloc instead: the staged offsets and
// tokens and the tile descriptors live in one statically sized device array
// of this library, `scratch`, and the call allocates nothing.
//
// Every context on a GPU shares that GPU's copy of the array, which keeps
// the last call's contents; the prep kernel writes every descriptor the tile
// kernel reads and the staged corpus is copied in by each call, so nothing
// in it needs clearing.  Calls that overlap on one GPU (two host threads,
// two contexts) race on the array: serialise them.
#include <hip/hip_runtime.h>
#include <mutex>
#include <string>
#include "felog_internal.h"

namespace felog {
namespace {

// 64 MiB per GPU, held from the library's load until the process ends.
constexpr uint64_t kScratchBytes = 64ull << 20;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

// The array's device address on each GPU (the symbol's host-side name is
// only a stand-in for it), looked up once per device.
constexpr int kMaxDevices = 64;
void *g_scratch[kMaxDevices] = {};
std::mutex g_scratch_mu;

// With device `dev` current: the address of its copy of the array.
hipError_t scratch_ptr(int dev, void **out) {
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch[dev] = p;
    }
    *out = g_scratch[dev];
    return hipSuccess;
}

inline uint64_t round256(uint64_t x) { return (x + 255) / 256 * 256; }

}  // namespace
}  // namespace felog

using namespace felog;

extern "C" felog_status felog_train_corpus_noalloc(felog_ctx *c, const uint64_t *offsets, const uint32_t *tokens,
                                                   uint64_t n, uint32_t round) {
    if (!c) return FELOG_ERR_ARG;
    if (!c->built) {
        c->err = "felog_build has not run";
        return FELOG_ERR_STATE;
    }
    auto fail = [&](felog_status s, const std::string &m) {
        c->err = m;
        return s;
    };
    auto hip_fail = [&](hipError_t e, const char *what) {
        c->err = std::string(what) + ": " + hipGetErrorString(e);
        return FELOG_ERR_HIP;
    };
    if (c->opt.trainer != FELOG_TRAIN_HOGBATCH) return fail(FELOG_ERR_STATE, "the tile trainer is not selected");
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    if (round == 0) return fail(FELOG_ERR_ARG, "rounds are numbered from 1");
    if (n > 0 && !offsets) return fail(FELOG_ERR_ARG, "null offsets");
    if (n == 0) return FELOG_OK;
    if (n >= 0xFFFFFFFFull) return fail(FELOG_ERR_ARG, "a corpus holds fewer than 2^32 - 1 sentences");
    if (offsets[0] != 0) return fail(FELOG_ERR_ARG, "offsets[0] must be 0");
    for (uint64_t i = 0; i < n; ++i)
        if (offsets[i + 1] < offsets[i]) return fail(FELOG_ERR_ARG, "offsets must be non-decreasing");
    const uint64_t nt = offsets[n];
    if (nt > 0 && !tokens) return fail(FELOG_ERR_ARG, "null tokens");
    for (uint64_t i = 0; i < nt; ++i)
        if (tokens[i] >= c->V) return fail(FELOG_ERR_ARG, "token id >= v_num");
    // layout: offsets | tokens | tile descriptors, each 256-byte aligned
    const uint64_t off_bytes = round256((n + 1) * 8), tok_bytes = round256((nt ? nt : 1) * 4);
    const uint64_t tiles = sgns_tiles_needed(c, n, nt);
    const uint64_t need = off_bytes + tok_bytes + tiles * kSgnsDescBytes;
    if (need > kScratchBytes)
        return fail(FELOG_ERR_CAPACITY, "corpus of " + std::to_string(n) + " sentences and " + std::to_string(nt) +
                                            " tokens needs " + std::to_string(need) + " bytes of device scratch; " +
                                            "felog_train_corpus_noalloc has " + std::to_string(kScratchBytes));
    void *base = nullptr;
    if ((e = scratch_ptr(c->device, &base)) != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    unsigned char *b = (unsigned char *)base;
    uint64_t *d_off = (uint64_t *)b;
    uint32_t *d_tok = (uint32_t *)(b + off_bytes);
    void *d_tiles = b + off_bytes + tok_bytes;
    if ((e = hipMemcpyAsync(d_off, offsets, (n + 1) * 8, hipMemcpyHostToDevice, c->stream)) != hipSuccess)
        return hip_fail(e, "hipMemcpyAsync(offsets)");
    if (nt && (e = hipMemcpyAsync(d_tok, tokens, nt * 4, hipMemcpyHostToDevice, c->stream)) != hipSuccess)
        return hip_fail(e, "hipMemcpyAsync(tokens)");
    if ((e = round_set(c, round, -1.f, (uint32_t)n)) != hipSuccess) return hip_fail(e, "round_set");
    float ms[3] = {0.f, 0.f, 0.f};
    if ((e = run_sgns_into(c, d_off, d_tok, nullptr, ms, d_tiles, tiles)) != hipSuccess)
        return hip_fail(e, "run_sgns_into");
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
    if (n + c->h_counters[C_TILES] > tiles)
        return fail(FELOG_ERR_INVARIANT, "SGNS tile descriptors exceeded their bound");
    return FELOG_OK;
}
