// tool_args.h - flag lookup shared by the graph tools (gconverter, gtrain,
// gcn, lpauc): "-x value" and "--flag" forms of the reference tools'
// option helpers (src/tools/*.cpp), found anywhere in argv.
#pragma once
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

namespace tool {

inline int find(int argc, char **argv, const char *name) {
    for (int a = 1; a < argc; ++a)
        if (!std::strcmp(argv[a], name)) return a;
    return -1;
}

// Value of "-x value"; def when absent.  Exits when the value is missing.
inline std::string str(int argc, char **argv, const char *name, const char *def = nullptr) {
    const int a = find(argc, argv, name);
    if (a < 0) {
        if (def) return def;
        std::fprintf(stderr, "missing required flag %s\n", name);
        std::exit(1);
    }
    if (a + 1 >= argc) {
        std::fprintf(stderr, "flag %s needs a value\n", name);
        std::exit(1);
    }
    return argv[a + 1];
}

inline bool flag(int argc, char **argv, const char *name) { return find(argc, argv, name) > 0; }

}  // namespace tool
