"""FP64 restatement of the reference's three word2vec trainers, for tests.

Sentences are lists of vertex ids (already filtered); sentence s runs with
block index blk[s] (its index in its training batch), the kernels' LCG seed:
  skip-gram   skip_gram_kernel, src/word2vec.cu:442-536
  reuse       __sgNegReuse, src/word2vec.cu:368-440: full window, one set of
              negatives per sentence (here negs[s], drawn by the caller)
  cbow        cbow_kernel, src/word2vec.cu:538-646
Each pair (or, for CBOW, each centre) is applied before the next, in order.
The sigmoid is read from the 1000-entry expTable at (int)((f + 6) * 83).
"""
import numpy as np

import orc

M64 = (1 << 64) - 1


def _g(f, label, alpha, et):
    if f > 6:
        return (label - 1) * alpha
    if f < -6:
        return label * alpha
    return (label - et[int((f + 6) * 83)]) * alpha


def _neg(nr, table, vocab_size, order):
    nr = (nr * 25214903917 + 11) & M64
    tr = int(table[(nr >> 16) % len(table)])
    if tr == 0:
        tr = nr % (vocab_size - 1) + 1
    return nr, int(order[tr])


def train(mode, syn0, syn1, sentences, blk, order, window, negative, alpha, table, vocab_size, negs=None):
    s0 = syn0.astype(np.float64)
    s1 = syn1.astype(np.float64)
    et = orc.w2v_exp_table().astype(np.float64)
    for si, words in enumerate(sentences):
        nr = int(blk[si])
        n = len(words)
        for pos, word in enumerate(words):
            if mode == "reuse":
                b = 0
            else:
                nr = (nr * 2514903917 + 11) & M64
                b = nr % window
            ctx = [words[pos - window + a] for a in range(b, 2 * window + 1 - b)
                   if a != window and 0 <= pos - window + a < n]
            if mode == "cbow":
                if not ctx:
                    continue
                h = sum(s0[c] for c in ctx) / len(ctx)
                e = np.zeros(s0.shape[1])
                for dd in range(negative + 1):
                    if dd == 0:
                        t, label = word, 1.0
                    else:
                        nr, t = _neg(nr, table, vocab_size, order)
                        if t == word:
                            continue
                        label = 0.0
                    g = _g(h @ s1[t], label, alpha, et)
                    e += g * s1[t]
                    s1[t] = s1[t] + g * h
                for c in ctx:
                    s0[c] = s0[c] + e
                continue
            for last in ctx:
                e = np.zeros(s0.shape[1])
                for dd in range(negative + 1):
                    if dd == 0:
                        t, label = word, 1.0
                    elif mode == "reuse":
                        t, label = int(negs[si][dd - 1]), 0.0
                        if t == word:
                            continue
                    else:
                        nr, t = _neg(nr, table, vocab_size, order)
                        if t == word:
                            continue
                        label = 0.0
                    g = _g(s0[last] @ s1[t], label, alpha, et)
                    e += g * s1[t]
                    s1[t] = s1[t] + g * s0[last]
                s0[last] = s0[last] + e
    return s0, s1


def reuse_negatives(n_sent, negative, table, vocab_size, order, seed, rnd):
    """Per-sentence negatives of the reuse trainer: Philox(seed; j, W2V,
    sentence, round) word 0 through the unigram table."""
    out = []
    for s in range(n_sent):
        row = []
        for j in range(negative):
            r = int(orc.philox([j, 0x70000000, s, rnd], [seed & 0xFFFFFFFF, seed >> 32])[0])
            tr = int(table[r % len(table)])
            if tr == 0:
                tr = r % (vocab_size - 1) + 1
            row.append(int(order[tr]))
        out.append(row)
    return out
