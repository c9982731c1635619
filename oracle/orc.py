"""ctypes binding of liboracle.so - the CPU restatement of the walk -> SGNS path.

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, always as the reference results are compared
against, never as the product.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# FELOG_ORACLE_LIB selects another build of the same sources (the
# address/undefined-behaviour sanitizer build, liboracle_san.so)
LIB_PATH = os.environ.get("FELOG_ORACLE_LIB") or os.path.join(HERE, "liboracle.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} missing: run `make -C oracle`")
        _lib = C.CDLL(LIB_PATH)
        _lib.orc_auc.restype = C.c_double
        _lib.orc_walks.restype = C.c_uint64
        _lib.orc_tanh_restated.restype = C.c_double
        _lib.orc_tanh_restated.argtypes = [C.c_double]
        _lib.orc_w2v_train.restype = C.c_uint64
        _lib.orc_eval.restype = C.c_uint32
        _lib.orc_sgns_groups.restype = C.c_uint32
        _lib.orc_build_csr.restype = C.c_int
    return _lib


def P(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def philox(ctr, key):
    c = np.asarray(ctr, np.uint32)
    k = np.asarray(key, np.uint32)
    out = np.empty(4, np.uint32)
    lib().orc_philox4x32_10(P(c), P(k), P(out))
    return out


def build_csr(V, src, dst, w, undirected=True):
    src = np.ascontiguousarray(src, np.uint32)
    dst = np.ascontiguousarray(dst, np.uint32)
    w = np.ascontiguousarray(w, np.float32)
    m = len(src)
    E = 2 * m if undirected else m
    rp = np.empty(V + 1, np.uint64)
    col = np.empty(max(E, 1), np.uint32)
    wo = np.empty(max(E, 1), np.float32)
    rc = lib().orc_build_csr(C.c_uint32(V), C.c_uint64(m), P(src), P(dst), P(w), C.c_int(int(undirected)),
                             P(rp), P(col), P(wo))
    if rc != 0:
        raise ValueError("edge id out of range")
    return rp, col[:E], wo[:E]


def build_alias(V, row_ptr, w):
    E = int(row_ptr[-1])
    t = np.empty(max(E, 1), np.uint32)
    a = np.empty(max(E, 1), np.uint32)
    lib().orc_build_alias(C.c_uint32(V), P(np.ascontiguousarray(row_ptr, np.uint64)),
                          P(np.ascontiguousarray(w, np.float32)), P(t), P(a))
    return t[:E], a[:E]


def alias_1d(wts):
    wts = np.ascontiguousarray(wts, np.float64)
    n = len(wts)
    t = np.empty(n, np.uint32)
    a = np.empty(n, np.uint32)
    lib().orc_alias_1d(C.c_uint32(n), P(wts), P(t), P(a))
    return t, a


def xlog2x(L):
    T = np.empty(L + 2, np.float64)
    dT = np.empty(L + 1, np.float64)
    lib().orc_xlog2x(C.c_uint32(L), P(T), P(dT))
    return T, dT


def corr_reference(f):
    f = np.ascontiguousarray(f, np.uint32)
    H = np.empty(len(f), np.float64)
    R = np.empty(len(f), np.float64)
    lib().orc_corr_reference(C.c_uint32(len(f)), P(f), P(H), P(R))
    return H, R


def corr_exact(f, L=1000):
    f = np.ascontiguousarray(f, np.uint32)
    T, dT = xlog2x(max(L, len(f) + 2))
    H = np.empty(len(f), np.float64)
    R = np.empty(len(f), np.float64)
    lib().orc_corr_exact(C.c_uint32(len(f)), P(f), P(T), P(dT), P(H), P(R))
    return H, R


def walks(V, row_ptr, col, thresh, alias_idx, flags, walkers, min_L, max_len, rnd, seed, threads=0,
          want_tokens=True, cn_data=None):
    """One round of walks.  cn_data (edge data per CSR entry) selects the
    reference's common-neighbour transition instead of the alias step."""
    flags = None if flags is None else np.ascontiguousarray(flags, np.uint8)
    cn_data = None if cn_data is None else np.ascontiguousarray(cn_data, np.float32)
    L = lib()
    nt = C.c_uint64()
    ns = C.c_uint64()
    args = (C.c_uint32(V), P(np.ascontiguousarray(row_ptr, np.uint64)), P(np.ascontiguousarray(col, np.uint32)),
            P(np.ascontiguousarray(thresh, np.uint32)), P(np.ascontiguousarray(alias_idx, np.uint32)), P(cn_data),
            P(flags),
            C.c_uint32(walkers), C.c_uint32(min_L), C.c_uint32(max_len), C.c_uint32(rnd), C.c_uint64(seed))
    off = np.empty(walkers + 1, np.uint64)
    wid = np.empty(max(walkers, 1), np.uint32)
    n = L.orc_walks(*args, P(off), None, P(wid), C.byref(nt), C.byref(ns), C.c_int(threads))
    if not want_tokens:
        return off[:n + 1], None, wid[:n], int(ns.value)
    tok = np.empty(max(nt.value, 1), np.uint32)
    n = L.orc_walks(*args, P(off), P(tok), P(wid), C.byref(nt), C.byref(ns), C.c_int(threads))
    return off[:n + 1], tok[:nt.value], wid[:n], int(ns.value)


def tanh_restated(x):
    return lib().orc_tanh_restated(float(x))


def vocab(deg, min_count):
    deg = np.ascontiguousarray(deg, np.uint32)
    V = len(deg)
    order = np.empty(V, np.uint32)
    rank = np.empty(V, np.uint32)
    inv = np.empty(V, np.uint8)
    vs = C.c_uint64()
    tw = C.c_uint64()
    lib().orc_vocab(C.c_uint32(V), P(deg), C.c_uint32(min_count), P(order), P(rank), P(inv), C.byref(vs),
                    C.byref(tw))
    return order, rank, inv, int(vs.value), int(tw.value)


def keep_thresholds(deg, sample, train_words):
    deg = np.ascontiguousarray(deg, np.uint32)
    keep = np.empty(len(deg), np.uint32)
    lib().orc_keep_thresholds(C.c_uint32(len(deg)), P(deg), C.c_float(sample), C.c_uint64(train_words), P(keep))
    return keep


def unigram(deg, in_vocab):
    deg = np.ascontiguousarray(deg, np.uint32)
    inv = np.ascontiguousarray(in_vocab, np.uint8)
    V = len(deg)
    t = np.empty(V, np.uint32)
    a = np.empty(V, np.uint32)
    lib().orc_unigram(C.c_uint32(V), P(deg), P(inv), P(t), P(a))
    return t, a


def init_syn0(V, d, rank):
    out = np.empty((V, d), np.float32)
    lib().orc_init_syn0(C.c_uint32(V), C.c_uint32(d), P(np.ascontiguousarray(rank, np.uint32)), P(out))
    return out


class SgnsParams(C.Structure):
    _fields_ = [("V", C.c_uint32), ("d", C.c_uint32), ("window", C.c_uint32), ("negative", C.c_uint32),
                ("bgroup", C.c_uint32), ("rmax", C.c_uint32), ("compat_break", C.c_int), ("alpha", C.c_float),
                ("round", C.c_uint32),
                ("seed", C.c_uint64), ("in_vocab", C.c_void_p), ("vocab_rank", C.c_void_p),
                ("keep_thr", C.c_void_p), ("uni_thr", C.c_void_p), ("uni_alias", C.c_void_p)]


def group_centres(batch, window, negative):
    """Centres per SGNS group (the kernel's tile limits: <= 32 outputs, <= 64 inputs)."""
    b = 2 * window + 1 if batch == 0 else batch
    b = min(b, 32 - negative, 64 - 2 * window)
    return b


def tile_rows(centres, window):
    """Input positions a tile may span: 48 when one group's windows fit, else 64."""
    return 48 if centres + 2 * window <= 48 else 64


def sgns_groups(tables, offsets, tokens, syn0, syn1, *, d, window, negative, batch, alpha, rnd, seed,
                compat_break=False, mode=0, max_groups=4096):
    """FP64 SGNS oracle.  tables = dict(in_vocab, vocab_rank, keep_thr, uni_thr, uni_alias)."""
    keep = [np.ascontiguousarray(tables[k], np.uint8 if k == "in_vocab" else np.uint32)
            for k in ("in_vocab", "vocab_rank", "keep_thr", "uni_thr", "uni_alias")]
    V = len(keep[0])
    B = group_centres(batch, window, negative)
    p = SgnsParams(V, d, window, negative, B, tile_rows(B, window), int(compat_break), alpha, rnd, seed,
                   *[k.ctypes.data for k in keep])
    s0 = np.ascontiguousarray(syn0, np.float64).copy()
    s1 = np.ascontiguousarray(syn1, np.float64).copy()
    offsets = np.ascontiguousarray(offsets, np.uint64)
    tokens = np.ascontiguousarray(tokens, np.uint32)
    G = max_groups if mode == 0 else 0
    groups = np.zeros((max(G, 1), 5), np.uint32)
    rows = np.zeros((max(G, 1), 96), np.uint32)
    logits = np.zeros((max(G, 1), 64, 32), np.float64)
    dv = np.zeros((max(G, 1), 64, d), np.float64)
    du = np.zeros((max(G, 1), 32, d), np.float64)
    pairs = C.c_uint64()
    n = lib().orc_sgns_groups(C.byref(p), P(offsets), P(tokens), C.c_uint64(len(offsets) - 1), P(s0), P(s1),
                              C.c_int(mode), C.c_uint32(G), P(groups), P(rows), P(logits), P(dv), P(du),
                              C.byref(pairs))
    k = min(n, G)
    return {"n_groups": n, "pairs": int(pairs.value), "groups": groups[:k], "rows": rows[:k],
            "logits": logits[:k], "dv": dv[:k], "du": du[:k], "syn0": s0, "syn1": s1, "keep": keep}


def w2v_train(syn0, syn1, offsets, tokens, deg, order, rank, in_vocab, vocab_size, train_words, *, window, negative,
              sample, alpha, use_lut=True, compat_break=False, table_size=10_000_000, threads=0):
    """word2vec skip-gram restatement; trains syn0/syn1 (f32, [V, d]) in place."""
    V, d = syn0.shape
    pairs = C.c_uint64()
    words = lib().orc_w2v_train(
        C.c_uint32(V), C.c_uint32(d), P(syn0), P(syn1), P(np.ascontiguousarray(offsets, np.uint64)),
        P(np.ascontiguousarray(tokens, np.uint32)), C.c_uint64(len(offsets) - 1), P(np.ascontiguousarray(deg, np.uint32)),
        P(np.ascontiguousarray(order, np.uint32)), P(np.ascontiguousarray(rank, np.uint32)),
        P(np.ascontiguousarray(in_vocab, np.uint8)), C.c_uint64(vocab_size), C.c_uint64(train_words),
        C.c_uint32(window), C.c_uint32(negative), C.c_float(sample), C.c_float(alpha), C.c_int(int(use_lut)),
        C.c_int(int(compat_break)), C.c_uint32(table_size), C.c_int(threads), C.byref(pairs))
    return int(words), int(pairs.value)


def evaluate(syn0, row_ptr, col, in_vocab, flags, rnd, seed):
    V, d = syn0.shape
    flags = np.ascontiguousarray(flags, np.uint8).copy()
    scores = np.zeros(V, np.float32)
    eva = lib().orc_eval(C.c_uint32(V), C.c_uint32(d), P(np.ascontiguousarray(syn0, np.float32)),
                         P(np.ascontiguousarray(row_ptr, np.uint64)), P(np.ascontiguousarray(col, np.uint32)),
                         P(np.ascontiguousarray(in_vocab, np.uint8)), P(flags), P(scores), C.c_uint32(rnd),
                         C.c_uint64(seed))
    return int(eva), flags, scores


def auc(pos, neg):
    pos = np.ascontiguousarray(pos, np.float32)
    neg = np.ascontiguousarray(neg, np.float32)
    return float(lib().orc_auc(C.c_uint64(len(pos)), P(pos), C.c_uint64(len(neg)), P(neg)))


def w2v_unigram_table(cn_by_rank, table_size):
    """word2vec's unigram table (vocabulary indices) over counts in rank order."""
    cn = np.ascontiguousarray(cn_by_rank, np.uint32)
    table = np.empty(table_size, np.int32)
    lib().orc_w2v_unigram_table(C.c_uint64(len(cn)), P(cn), C.c_uint32(table_size), P(table))
    return table


def w2v_exp_table():
    t = np.empty(1001, np.float32)
    lib().orc_w2v_exp_table(P(t))
    return t


def w2v_lut_index(f):
    lib().orc_w2v_lut_index.argtypes = [C.c_float]
    return int(lib().orc_w2v_lut_index(C.c_float(f)))


def w2v_negative(next_random, table, vocab_size):
    """One negative draw: returns (vocabulary index, advanced LCG state)."""
    nr = C.c_uint64(next_random)
    table = np.ascontiguousarray(table, np.int32)
    lib().orc_w2v_negative.restype = C.c_uint32
    r = lib().orc_w2v_negative(C.byref(nr), P(table), C.c_uint32(len(table)), C.c_uint64(vocab_size))
    return int(r), int(nr.value)


def eval_subset(deg, v, rnd, seed):
    """Neighbour columns the evaluation scores for vertex v of degree deg."""
    out = np.empty(30, np.uint32)
    lib().orc_eval_subset.restype = C.c_uint32
    n = lib().orc_eval_subset(C.c_uint32(deg), C.c_uint32(v), C.c_uint32(rnd), C.c_uint64(seed), P(out))
    return out[:n].copy()


def common_neighbours(V, row_ptr, col, u, v):
    u = np.ascontiguousarray(u, np.uint32)
    v = np.ascontiguousarray(v, np.uint32)
    out = np.empty(max(len(u), 1), np.int32)
    lib().orc_common_neighbours(C.c_uint32(V), P(np.ascontiguousarray(row_ptr, np.uint64)),
                                P(np.ascontiguousarray(col, np.uint32)), C.c_uint64(len(u)), P(u), P(v), P(out))
    return out[:len(u)]


def kmeans_classes(syn0, order, vocab_size, classes, iters=10):
    """-classes k-means restated; classes of the vocabulary in vocabulary order."""
    V, d = syn0.shape
    cl = np.empty(max(vocab_size, 1), np.int32)
    lib().orc_kmeans_classes(C.c_uint32(d), P(np.ascontiguousarray(syn0, np.float32)),
                             P(np.ascontiguousarray(order, np.uint32)), C.c_uint64(vocab_size), C.c_uint32(classes),
                             C.c_uint32(iters), P(cl))
    return cl[:vocab_size]


W2V_SKIPGRAM, W2V_REUSE, W2V_CBOW = 0, 1, 2


def w2v_train_mode(syn0, syn1, offsets, tokens, deg, order, rank, in_vocab, vocab_size, train_words, *, mode,
                   window, negative, sample, alpha, use_lut=True, compat_break=False, table_size=10_000_000,
                   threads=0, seed=0, rnd=0):
    """word2vec restatement with the trainer chosen by mode (W2V_*)."""
    V, d = syn0.shape
    pairs = C.c_uint64()
    lib().orc_w2v_train_mode.restype = C.c_uint64
    words = lib().orc_w2v_train_mode(
        C.c_uint32(V), C.c_uint32(d), P(syn0), P(syn1), P(np.ascontiguousarray(offsets, np.uint64)),
        P(np.ascontiguousarray(tokens, np.uint32)), C.c_uint64(len(offsets) - 1),
        P(np.ascontiguousarray(deg, np.uint32)), P(np.ascontiguousarray(order, np.uint32)),
        P(np.ascontiguousarray(rank, np.uint32)), P(np.ascontiguousarray(in_vocab, np.uint8)),
        C.c_uint64(vocab_size), C.c_uint64(train_words), C.c_uint32(window), C.c_uint32(negative),
        C.c_float(sample), C.c_float(alpha), C.c_int(int(use_lut)), C.c_int(int(compat_break)),
        C.c_uint32(table_size), C.c_int(threads), C.byref(pairs), C.c_int(mode), C.c_uint64(seed), C.c_uint32(rnd))
    return int(words), int(pairs.value)
