// gconverter - text edge list to the binary graph file felog loads
// (src/tools/gconverter.cpp:57-76): "src dst weight" records to packed
// {u32, u32, f32} (-s weighted) or "src dst" to {u32, u32} (-s unweighted).
//   gconverter -i <edges.txt> -o <graph.bin> -s weighted|unweighted
#include <vector>
#include "felog.h"
#include "tool_args.h"

int main(int argc, char **argv) {
    const std::string in = tool::str(argc, argv, "-i"), out = tool::str(argc, argv, "-o");
    const std::string kind = tool::str(argc, argv, "-s");
    if (kind != "weighted" && kind != "unweighted") {
        std::fprintf(stderr, "gconverter: -s must be weighted or unweighted\n");
        return 1;
    }
    const bool weighted = kind == "weighted";
    const int cols = weighted ? 3 : 2;
    uint64_t m = 0;
    if (felog_read_edges_txt(in.c_str(), cols, nullptr, nullptr, nullptr, 0, &m) != FELOG_OK) {
        std::fprintf(stderr, "gconverter: cannot read %s\n", in.c_str());
        return 1;
    }
    std::vector<uint32_t> src(m), dst(m);
    std::vector<float> w(weighted ? m : 0);
    if (felog_read_edges_txt(in.c_str(), cols, src.data(), dst.data(), weighted ? w.data() : nullptr, m, &m) !=
            FELOG_OK ||
        felog_write_edges_bin(out.c_str(), src.data(), dst.data(), weighted ? w.data() : nullptr, m) != FELOG_OK) {
        std::fprintf(stderr, "gconverter: conversion failed\n");
        return 1;
    }
    std::printf("%llu edges are converted\n", (unsigned long long)m);
    return 0;
}
