// gcn - common-neighbour counts of every edge (src/tools/gcn.cpp:249-368),
// counted on the GPU.
//   gcn -i <edges.txt> -o <out.txt> -v <vertices> -s weighted|unweighted
//       [--make-undirected] [-t threads] [--device N]
// Reads "src dst [weight]" records, builds the graph (each edge also
// reversed with --make-undirected) and writes "src dst cn\n" per input
// edge, cn = the distinct common neighbours of src and dst other than
// themselves.  felog's compat transition reads this third column as the
// edge data.  -t is accepted and unused.
#include <vector>
#include "felog.h"
#include "tool_args.h"

int main(int argc, char **argv) {
    const std::string in = tool::str(argc, argv, "-i"), out = tool::str(argc, argv, "-o");
    const std::string kind = tool::str(argc, argv, "-s", "unweighted");
    const long long vn = std::atoll(tool::str(argc, argv, "-v").c_str());
    const bool undirected = tool::flag(argc, argv, "--make-undirected");
    if (vn <= 0) {
        std::fprintf(stderr, "gcn: -v must be > 0\n");
        return 1;
    }
    const int cols = kind == "weighted" ? 3 : 2;
    uint64_t m = 0;
    if (felog_read_edges_txt(in.c_str(), cols, nullptr, nullptr, nullptr, 0, &m) != FELOG_OK) {
        std::fprintf(stderr, "gcn: cannot read %s\n", in.c_str());
        return 1;
    }
    std::vector<uint32_t> src(m), dst(m);
    std::vector<float> w(m, 1.f);
    felog_read_edges_txt(in.c_str(), cols, src.data(), dst.data(), cols == 3 ? w.data() : nullptr, m, &m);
    felog_options o;
    felog_default_options(&o);
    o.v_num = (uint32_t)vn;
    o.size = 32;  // the tables are not used
    o.device = std::atoi(tool::str(argc, argv, "--device", "0").c_str());
    felog_ctx *c = nullptr;
    felog_status s = felog_create(&c, &o);
    std::vector<int32_t> cn(m);
    if (s == FELOG_OK) s = felog_set_edges(c, src.data(), dst.data(), w.data(), m, (uint32_t)vn, undirected);
    if (s == FELOG_OK) s = felog_build(c);
    if (s == FELOG_OK) s = felog_common_neighbours(c, src.data(), dst.data(), m, cn.data());
    if (s != FELOG_OK) {
        std::fprintf(stderr, "gcn: failed (status %d): %s\n", s, c ? felog_last_error(c) : "");
        if (c) felog_destroy(c);
        return 1;
    }
    felog_destroy(c);
    FILE *f = std::fopen(out.c_str(), "w");
    if (!f) {
        std::fprintf(stderr, "gcn: cannot open %s\n", out.c_str());
        return 1;
    }
    for (uint64_t i = 0; i < m; ++i) std::fprintf(f, "%u %u %d\n", src[i], dst[i], cn[i]);
    return std::fclose(f) == 0 ? 0 : 1;
}
