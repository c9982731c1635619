// felog_main.cpp - the felog command line on top of libfelog_hip.so.
//
// Accepts the reference's argv (README.md:64-69):
//   felog -g graph.data -v V -w V --min_L 20 --min_R 2 [--make-undirected]
//         [-o prefix] -size 128 -window 10 -negative 5 -sample 1e-3
//         -cbow 0 -reuse-neg 0 -emb_output emb.txt ...
// Parsing follows the reference's two passes:
//   1. graph / walk flags (include/option_helper.hpp:94-188) are read from
//      the front of argv and parsing stops silently at the first token it
//      does not know (include/option_helper.hpp:40-45);
//   2. word2vec flags are found anywhere in argv by exact match
//      (ArgPos, src/word2vec.cu:2182-2192,2269-2286).
// Switches of this build (also found anywhere): --precision fp32|bf16,
// --seed N, --max-rounds N, --max-walk-len N, --device N, --write-walks,
// --felog-compat-lr, --felog-compat-vocab, --felog-compat-walk,
// --felog-compat-sgns.
// Trainers, as the reference's -cbow / -reuse-neg pick its kernels
// (src/word2vec.cu:163, 931-1007): -cbow 1 (the default) trains cbow_kernel's
// rule, -cbow 0 -reuse-neg 1 (the default) __sgNegReuse's, -cbow 0
// -reuse-neg 0 the MFMA HogBatch tile trainer that replaces skip_gram_kernel
// (--felog-compat-sgns: skip_gram_kernel's per-pair rule instead).  The
// per-pair trainers follow the reference's step ExponentialDecayLR(0.025,
// 0.1) and ignore -alpha, as the reference does (src/word2vec.cu:1958,2001);
// the tile trainer honours -alpha (default 0.0025, its stable step).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include "felog.h"

namespace {

int arg_pos(const char *name, int argc, char **argv, bool has_value) {
    for (int a = 1; a < argc; ++a)
        if (!std::strcmp(name, argv[a])) {
            if (has_value && a == argc - 1) {
                std::fprintf(stderr, "Argument missing for %s\n", name);
                std::exit(1);
            }
            return a;
        }
    return -1;
}

struct Front {
    std::string g, p, o;
    long long v = -1, w = -1, min_L = -1, min_R = -1, batch = -1;
    bool undirected = false;
    double rate = 0.0;
};

// First pass: the args-library flags, stopping at the first unknown token.
Front parse_front(int argc, char **argv) {
    Front f;
    for (int a = 1; a < argc; ++a) {
        std::string t = argv[a], val;
        bool has_eq = false;
        const size_t eq = t.find('=');
        if (t.rfind("--", 0) == 0 && eq != std::string::npos) {
            val = t.substr(eq + 1);
            t = t.substr(0, eq);
            has_eq = true;
        }
        auto take = [&](std::string &out) -> bool {
            if (has_eq) { out = val; return true; }
            if (a + 1 >= argc) return false;
            out = argv[++a];
            return true;
        };
        std::string x;
        if (t == "--make-undirected") { f.undirected = true; continue; }
        if (t == "-g") { if (!take(f.g)) break; continue; }
        if (t == "-p") { if (!take(f.p)) break; continue; }
        if (t == "-o") { if (!take(f.o)) break; continue; }
        if (t == "-v") { if (!take(x)) break; f.v = std::atoll(x.c_str()); continue; }
        if (t == "-w") { if (!take(x)) break; f.w = std::atoll(x.c_str()); continue; }
        if (t == "--min_L") { if (!take(x)) break; f.min_L = std::atoll(x.c_str()); continue; }
        if (t == "--min_R") { if (!take(x)) break; f.min_R = std::atoll(x.c_str()); continue; }
        if (t == "--batch_size") { if (!take(x)) break; f.batch = std::atoll(x.c_str()); continue; }
        if (t == "-r") { if (!take(x)) break; f.rate = std::atof(x.c_str()); continue; }
        break;  // unknown token: the args parser's error is swallowed
    }
    return f;
}

int die(felog_ctx *c, const char *what, felog_status s) {
    std::fprintf(stderr, "felog: %s failed (status %d): %s\n", what, s, c ? felog_last_error(c) : "");
    if (c) felog_destroy(c);
    return 1;
}

}  // namespace

int main(int argc, char **argv) {
    const Front f = parse_front(argc, argv);
    if (f.g.empty() || f.v <= 0 || f.w < 0 || f.min_L < 0 || f.min_R < 0) {
        std::fprintf(stderr,
                     "usage: felog -g <graph.bin> -v <vertices> -w <walkers> --min_L <n> --min_R <n> "
                     "[--make-undirected] [-o <prefix>] -emb_output <file> [word2vec flags]\n"
                     "  -cbow 1 (default): CBOW; -cbow 0 -reuse-neg 1 (default): skip-gram with per-sentence\n"
                     "  negatives; -cbow 0 -reuse-neg 0: the HogBatch tile trainer (-alpha honoured, default\n"
                     "  0.0025); the per-pair trainers use the reference's step 0.025 e^{-0.1 r} and ignore -alpha\n");
        return 1;
    }
    felog_options o;
    felog_default_options(&o);
    o.v_num = (uint32_t)f.v;
    o.walkers = f.w > 0 ? (uint32_t)f.w : 0;
    o.min_L = (uint32_t)f.min_L;
    o.min_R = (uint32_t)f.min_R;
    o.make_undirected = f.undirected;
    if (f.batch > 0) o.eval_batch = (uint32_t)f.batch;
    int i;
    std::string emb;
    if ((i = arg_pos("-size", argc, argv, true)) > 0) o.size = (uint32_t)std::atoi(argv[i + 1]);
    if ((i = arg_pos("-binary", argc, argv, true)) > 0) o.binary = std::atoi(argv[i + 1]);
    int cbow = 1;
    if ((i = arg_pos("-cbow", argc, argv, true)) > 0) cbow = std::atoi(argv[i + 1]);
    if ((i = arg_pos("-alpha", argc, argv, true)) > 0) o.alpha = (float)std::atof(argv[i + 1]);
    if ((i = arg_pos("-emb_output", argc, argv, true)) > 0) emb = argv[i + 1];
    if ((i = arg_pos("-window", argc, argv, true)) > 0) o.window = (uint32_t)std::atoi(argv[i + 1]);
    if ((i = arg_pos("-sample", argc, argv, true)) > 0) o.sample = (float)std::atof(argv[i + 1]);
    int hs = 0, classes = 0;
    if ((i = arg_pos("-hs", argc, argv, true)) > 0) hs = std::atoi(argv[i + 1]);
    if ((i = arg_pos("-negative", argc, argv, true)) > 0) o.negative = (uint32_t)std::atoi(argv[i + 1]);
    if ((i = arg_pos("-min-count", argc, argv, true)) > 0) o.min_count = (uint32_t)std::atoi(argv[i + 1]);
    if ((i = arg_pos("-classes", argc, argv, true)) > 0) classes = std::atoi(argv[i + 1]);
    if ((i = arg_pos("-batch-size", argc, argv, true)) > 0) o.batch_size = (uint32_t)std::atoi(argv[i + 1]);
    if ((i = arg_pos("-train", argc, argv, true)) > 0)
        std::printf("[ Warning ] '-train' option is ignored; training data comes from in-memory corpus.\n");
    // build switches
    if ((i = arg_pos("--precision", argc, argv, true)) > 0)
        o.precision = !std::strcmp(argv[i + 1], "bf16") ? FELOG_BF16 : FELOG_FP32;
    if ((i = arg_pos("--seed", argc, argv, true)) > 0) o.seed = std::strtoull(argv[i + 1], nullptr, 0);
    if ((i = arg_pos("--max-rounds", argc, argv, true)) > 0) o.max_rounds = (uint32_t)std::atoi(argv[i + 1]);
    if ((i = arg_pos("--max-walk-len", argc, argv, true)) > 0) o.max_walk_len = (uint32_t)std::atoi(argv[i + 1]);
    if ((i = arg_pos("--device", argc, argv, true)) > 0) o.device = std::atoi(argv[i + 1]);
    o.compat_lr = arg_pos("--felog-compat-lr", argc, argv, false) > 0;
    o.compat_vocab_break = arg_pos("--felog-compat-vocab", argc, argv, false) > 0;
    o.compat_walk = arg_pos("--felog-compat-walk", argc, argv, false) > 0;
    const bool write_walks = arg_pos("--write-walks", argc, argv, false) > 0;

    if (emb.empty()) {
        std::fprintf(stderr, "felog: -emb_output <file> is required\n");
        return 1;
    }
    if (hs) {
        std::fprintf(stderr, "felog: -hs (hierarchical softmax) is not part of this build\n");
        return 1;
    }
    if (classes < 0) {
        std::fprintf(stderr, "felog: -classes must be >= 0\n");
        return 1;
    }
    if (classes && o.precision != FELOG_FP32) {
        std::fprintf(stderr, "felog: -classes runs k-means on FP32 tables\n");
        return 1;
    }
    int reuse_neg = 1;  // src/word2vec.cu:163
    if ((i = arg_pos("-reuse-neg", argc, argv, true)) > 0) reuse_neg = std::atoi(argv[i + 1]);
    const bool compat_sgns = arg_pos("--felog-compat-sgns", argc, argv, false) > 0;
    o.trainer = cbow ? FELOG_TRAIN_CBOW
                     : reuse_neg ? FELOG_TRAIN_SKIPGRAM_REUSE
                                 : compat_sgns ? FELOG_TRAIN_SKIPGRAM : FELOG_TRAIN_HOGBATCH;
    if (o.trainer != FELOG_TRAIN_HOGBATCH) {
        o.compat_lr = 1;
        if (o.precision != FELOG_FP32) {
            std::fprintf(stderr, "felog: the CBOW / negative-reuse / per-pair skip-gram trainers use FP32 tables\n");
            return 1;
        }
    }
    static const char *kTrainer[] = {"HogBatch tile (skip-gram)", "skip-gram, per pair", "skip-gram, per-sentence negatives", "CBOW"};
    std::printf("[ 0 ] trainer: %s\n", kTrainer[o.trainer]);
    if (write_walks && f.o.empty()) {
        std::fprintf(stderr, "felog: --write-walks needs -o <prefix>\n");
        return 1;
    }

    felog_ctx *c = nullptr;
    felog_status s = felog_create(&c, &o);
    if (s) return die(nullptr, "felog_create", s);
    const auto t0 = std::chrono::steady_clock::now();
    if ((s = felog_load_edges(c, f.g.c_str(), o.v_num, o.make_undirected))) return die(c, "load", s);
    if ((s = felog_build(c))) return die(c, "build", s);
    felog_stats st;
    felog_stats_get(c, &st);
    std::printf("[ 0 ] graph: %llu vertices, %llu CSR entries, vocab %llu, build %.1f ms\n",
                (unsigned long long)st.v_num, (unsigned long long)st.e_num, (unsigned long long)st.vocab_size,
                st.build_ms);
    for (uint32_t r = 1;; ++r) {
        if (o.max_rounds && r > o.max_rounds) break;
        felog_corpus corpus;
        if ((s = felog_walk_round(c, r, &corpus))) return die(c, "walk", s);
        if (write_walks && (s = felog_write_walks(c, f.o.c_str()))) return die(c, "write walks", s);
        if ((s = felog_train_round(c, &corpus))) return die(c, "train", s);
        felog_stats_get(c, &st);
        std::printf("[ 0 ] round %u: %llu walks, %llu steps (%.2f ms), %llu words (%.2f ms)\n", r,
                    (unsigned long long)corpus.n_walks, (unsigned long long)corpus.n_steps, st.walk_kernel_ms,
                    (unsigned long long)st.train_words, st.train_kernel_ms);
        if (r >= o.min_R) {
            uint32_t nu = 0;
            int stop = 0;
            if ((s = felog_evaluate(c, &nu, &stop))) return die(c, "evaluate", s);
            std::printf("[ 0 ] round %u: %u vertices not converged\n", r, nu);
            if (stop) break;
        } else {
            std::printf("[ 0 ]Iter %u Skipping evaluation (init_round=%u)\n", r, o.min_R);
        }
        if (corpus.n_walks == 0) break;
    }
    // src/word2vec.cu:2118-2172: the vectors, or with -classes their k-means classes
    if (classes) {
        if ((s = felog_write_classes(c, emb.c_str(), (uint32_t)classes))) return die(c, "write classes", s);
    } else if ((s = felog_write_embeddings(c, emb.c_str(), o.binary))) {
        return die(c, "write embeddings", s);
    }
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    felog_stats_get(c, &st);
    std::printf("walk time %.3f s, train time %.3f s, whole time %.3f s\n", st.walk_ms_total / 1e3,
                st.train_ms_total / 1e3, secs);
    felog_destroy(c);
    return 0;
}
