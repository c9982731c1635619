// lpauc - held-out link-prediction AUC of an embedding file on gtrain's
// test files (the evaluation the reference leaves to outside scripts).
//   lpauc -emb <embeddings> -e <test_edges.txt> -l <test_labels.txt> [-binary 1]
// The embedding file is felog's output (src/word2vec.cu:2117-2127): "V d\n",
// then per row "<vertex id> " and d values as text ("%lf ") or, with
// -binary 1, d raw floats, then "\n".  A pair's score is the cosine of its
// two rows (0 when a vertex has no row); the AUC is the Mann-Whitney
// statistic with tied scores at half weight.  Prints "AUC <value>".
#include <algorithm>
#include <cmath>
#include <unordered_map>
#include <vector>
#include "tool_args.h"

int main(int argc, char **argv) {
    const std::string ep = tool::str(argc, argv, "-emb"), te = tool::str(argc, argv, "-e");
    const std::string tl = tool::str(argc, argv, "-l");
    const bool binary = std::atoi(tool::str(argc, argv, "-binary", "0").c_str()) != 0;
    FILE *f = std::fopen(ep.c_str(), "rb");
    if (!f) {
        std::fprintf(stderr, "lpauc: cannot open %s\n", ep.c_str());
        return 1;
    }
    long long V = 0, d = 0;
    if (std::fscanf(f, "%lld %lld", &V, &d) != 2 || V < 0 || d <= 0) {
        std::fprintf(stderr, "lpauc: bad header in %s\n", ep.c_str());
        return 1;
    }
    std::fgetc(f);  // the header's newline
    std::unordered_map<uint32_t, size_t> row;
    std::vector<float> emb((size_t)V * d);
    for (long long r = 0; r < V; ++r) {
        unsigned id;
        if (std::fscanf(f, "%u", &id) != 1 || std::fgetc(f) != ' ') {
            std::fprintf(stderr, "lpauc: row %lld is malformed\n", r);
            return 1;
        }
        float *x = &emb[(size_t)r * d];
        if (binary) {
            if (std::fread(x, 4, (size_t)d, f) != (size_t)d) {
                std::fprintf(stderr, "lpauc: row %lld is short\n", r);
                return 1;
            }
        } else {
            for (long long j = 0; j < d; ++j) {
                double v;
                if (std::fscanf(f, "%lf", &v) != 1) {
                    std::fprintf(stderr, "lpauc: row %lld is short\n", r);
                    return 1;
                }
                x[j] = (float)v;
            }
        }
        row[id] = (size_t)r;
    }
    std::fclose(f);
    FILE *fe = std::fopen(te.c_str(), "r"), *fl = std::fopen(tl.c_str(), "r");
    if (!fe || !fl) {
        std::fprintf(stderr, "lpauc: cannot open the test files\n");
        return 1;
    }
    std::vector<std::pair<double, int>> sc;
    unsigned a, b;
    int lab;
    while (std::fscanf(fe, "%u %u", &a, &b) == 2 && std::fscanf(fl, "%d", &lab) == 1) {
        double s = 0.0;
        auto ia = row.find(a), ib = row.find(b);
        if (ia != row.end() && ib != row.end()) {
            const float *x = &emb[ia->second * d], *y = &emb[ib->second * d];
            double xy = 0, xx = 0, yy = 0;
            for (long long j = 0; j < d; ++j) {
                xy += (double)x[j] * y[j];
                xx += (double)x[j] * x[j];
                yy += (double)y[j] * y[j];
            }
            s = (xx > 0 && yy > 0) ? xy / std::sqrt(xx * yy) : 0.0;
        }
        sc.push_back({s, lab ? 1 : 0});
    }
    std::fclose(fe);
    std::fclose(fl);
    std::sort(sc.begin(), sc.end());
    double rank_pos = 0, P = 0;
    for (size_t i = 0; i < sc.size();) {
        size_t j = i;
        while (j < sc.size() && sc[j].first == sc[i].first) ++j;
        const double r = 0.5 * (double)(i + 1 + j);  // average of ranks i+1 .. j
        for (size_t q = i; q < j; ++q)
            if (sc[q].second) {
                rank_pos += r;
                P += 1;
            }
        i = j;
    }
    const double N = (double)sc.size() - P;
    const double auc = (P == 0 || N == 0) ? 0.5 : (rank_pos - P * (P + 1) / 2) / (P * N);
    std::printf("AUC %.6f\npairs %zu positives %.0f\n", auc, sc.size(), P);
    return 0;
}
