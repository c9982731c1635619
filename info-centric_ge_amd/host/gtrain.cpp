// gtrain - the link-prediction split of src/tools/gtrain.cpp:225-384.
//   gtrain -i <edges.txt> -o <train.txt> -e <test_edges.txt> -l <test_labels.txt>
//          -s <fraction> -v <vertices> [-t threads] [--seed N]
// Input: "src dst data" records (unsigned data).  k = fraction * edges
// edges are held out one at a time, each drawn uniformly from the edges not
// yet held out and kept only while both endpoints still have more than one
// neighbour (degrees over both directions, decremented per held-out edge);
// k vertex pairs that are not edges of the input and not drawn before
// (ordered pairs) are the negatives.  Outputs:
//   train        the remaining edges as "%u %u %u\n" (in the order the
//                held-out draws leave them, as the reference's swaps do)
//   test_edges   "%u %u\n": the held-out edges, then the negatives
//   test_labels  "1\n" per held-out edge, then "0\n" per negative
// The reference draws from rand() seeded with the time; --seed (default 1)
// makes the split reproducible.  -t is accepted and unused.
#include <algorithm>
#include <cstdint>
#include <random>
#include <set>
#include <vector>
#include "felog.h"
#include "tool_args.h"

namespace {

bool read_u32_triples(const char *path, std::vector<uint32_t> &s, std::vector<uint32_t> &d, std::vector<uint32_t> &x) {
    FILE *f = std::fopen(path, "r");
    if (!f) return false;
    unsigned a, b, c;
    while (std::fscanf(f, "%u %u %u", &a, &b, &c) == 3) {
        s.push_back(a);
        d.push_back(b);
        x.push_back(c);
    }
    std::fclose(f);
    return true;
}

}  // namespace

int main(int argc, char **argv) {
    const std::string in = tool::str(argc, argv, "-i"), lp = tool::str(argc, argv, "-o");
    const std::string te = tool::str(argc, argv, "-e"), tl = tool::str(argc, argv, "-l");
    const double sample = std::atof(tool::str(argc, argv, "-s").c_str());
    const long long vn = std::atoll(tool::str(argc, argv, "-v").c_str());
    const unsigned long long seed = std::strtoull(tool::str(argc, argv, "--seed", "1").c_str(), nullptr, 0);
    if (sample > 1.0 || sample < 0.0 || vn <= 0) {
        std::fprintf(stderr, "[error] need 0 <= -s <= 1 and -v > 0\n");
        return 1;
    }
    const uint32_t V = (uint32_t)vn;
    std::vector<uint32_t> src, dst, data;
    if (!read_u32_triples(in.c_str(), src, dst, data)) {
        std::fprintf(stderr, "gtrain: cannot read %s\n", in.c_str());
        return 1;
    }
    const uint64_t m = src.size();
    std::printf("input file has %llu edges\n", (unsigned long long)m);
    std::vector<uint32_t> deg(V, 0);
    for (uint64_t e = 0; e < m; ++e) {
        if (src[e] >= V || dst[e] >= V) {
            std::fprintf(stderr, "gtrain: vertex id >= %u\n", V);
            return 1;
        }
        ++deg[src[e]];
        ++deg[dst[e]];
    }
    // sorted adjacency (both directions) to reject negatives that are edges
    std::vector<uint64_t> rp(V + 1, 0);
    for (uint64_t e = 0; e < m; ++e) {
        ++rp[src[e] + 1];
        ++rp[dst[e] + 1];
    }
    for (uint32_t v = 0; v < V; ++v) rp[v + 1] += rp[v];
    std::vector<uint32_t> adj(rp[V]);
    {
        std::vector<uint64_t> end(rp.begin(), rp.end() - 1);
        for (uint64_t e = 0; e < m; ++e) {
            adj[end[src[e]]++] = dst[e];
            adj[end[dst[e]]++] = src[e];
        }
        for (uint32_t v = 0; v < V; ++v) std::sort(adj.begin() + rp[v], adj.begin() + rp[v + 1]);
    }
    auto is_edge = [&](uint32_t a, uint32_t b) {
        return std::binary_search(adj.begin() + rp[a], adj.begin() + rp[a + 1], b);
    };
    const uint64_t npos = (uint64_t)(float)(sample * (double)m), nneg = npos;
    std::mt19937_64 rng(seed);
    auto draw = [&](uint64_t range) { return rng() % range; };
    std::vector<uint32_t> ps(npos), pd(npos);
    int64_t tail = (int64_t)m - 1;
    for (uint64_t i = 0; i < npos; ++i) {
        uint64_t tries = 0, ep;
        do {
            if (tail < 0 || ++tries > 1000 * m + 1000) {
                std::fprintf(stderr, "gtrain: no edge left whose endpoints keep a neighbour\n");
                return 1;
            }
            ep = draw((uint64_t)tail + 1);
        } while (deg[src[ep]] <= 1 || deg[dst[ep]] <= 1);
        ps[i] = src[ep];
        pd[i] = dst[ep];
        --deg[src[ep]];
        --deg[dst[ep]];
        std::swap(src[ep], src[tail]);
        std::swap(dst[ep], dst[tail]);
        std::swap(data[ep], data[tail]);
        --tail;
    }
    std::vector<uint32_t> ns(nneg), nd(nneg);
    std::set<uint64_t> seen;
    for (uint64_t i = 0; i < nneg; ++i) {
        uint32_t a, b;
        uint64_t tries = 0;
        do {
            if (++tries > 1000 * nneg + 1000000) {
                std::fprintf(stderr, "gtrain: cannot find enough vertex pairs that are not edges\n");
                return 1;
            }
            a = (uint32_t)draw(V);
            b = (uint32_t)draw(V);
        } while (is_edge(a, b) || !seen.insert(((uint64_t)a << 32) | b).second);
        ns[i] = a;
        nd[i] = b;
    }
    FILE *fl = std::fopen(lp.c_str(), "w"), *fe = std::fopen(te.c_str(), "w"), *fb = std::fopen(tl.c_str(), "w");
    if (!fl || !fe || !fb) {
        std::fprintf(stderr, "gtrain: cannot open an output file\n");
        return 1;
    }
    for (int64_t e = 0; e <= tail; ++e) std::fprintf(fl, "%u %u %u\n", src[e], dst[e], data[e]);
    for (uint64_t i = 0; i < npos; ++i) {
        std::fprintf(fe, "%u %u\n", ps[i], pd[i]);
        std::fprintf(fb, "1\n");
    }
    for (uint64_t i = 0; i < nneg; ++i) {
        std::fprintf(fe, "%u %u\n", ns[i], nd[i]);
        std::fprintf(fb, "0\n");
    }
    const bool ok = std::fclose(fl) == 0 && std::fclose(fe) == 0 && std::fclose(fb) == 0;
    std::printf("%lld train edges\nnpos edge : %llu\nnneg edge: %llu\n", (long long)(tail + 1),
                (unsigned long long)npos, (unsigned long long)nneg);
    return ok ? 0 : 1;
}
