"""Python binding of libfelog_hip.so (the C ABI declared in include/felog.h).

The binding is a thin ctypes layer: every method calls one ``felog_*`` entry
point and raises :class:`FelogError` with the library's message on a negative
status.  There is no fallback: if the shared library is missing the import of
:func:`load_library` raises, so nothing silently runs on the CPU.

Reference interfaces mirrored (paths relative to the reference repository):
``WalkEngine::random_walk`` (include/walk.hpp:529-542), ``train_corpus_cuda``
(src/word2vec.cu:2193), ``GraphEngine::load_graph`` (include/graph.hpp:712).
"""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfelog_hip.so")
HEADER_PATH = os.path.join(os.path.dirname(HERE), "include", "felog.h")

FP32, BF16 = 0, 1
UPDATE_DEFAULT, UPDATE_ATOMIC, UPDATE_WRITEBACK = 0, 1, 2
TRAIN_HOGBATCH, TRAIN_SKIPGRAM, TRAIN_SKIPGRAM_REUSE, TRAIN_CBOW = 0, 1, 2, 3
STATUS = {0: "OK", -1: "ERR_ARG", -2: "ERR_IO", -3: "ERR_HIP", -4: "ERR_OOM",
          -5: "ERR_INVARIANT", -6: "ERR_STATE", -7: "ERR_CAPACITY"}


class FelogError(RuntimeError):
    def __init__(self, status, where, msg):
        super().__init__(f"{where}: {STATUS.get(status, status)}: {msg}")
        self.status = status


class Options(C.Structure):
    _fields_ = [
        ("v_num", C.c_uint32), ("walkers", C.c_uint32), ("min_L", C.c_uint32),
        ("min_R", C.c_uint32), ("eval_batch", C.c_uint32), ("make_undirected", C.c_int32),
        ("size", C.c_uint32), ("window", C.c_uint32), ("negative", C.c_uint32),
        ("batch_size", C.c_uint32), ("sample", C.c_float), ("alpha", C.c_float),
        ("min_count", C.c_uint32), ("binary", C.c_int32), ("precision", C.c_int32),
        ("max_walk_len", C.c_uint32), ("overflow_capacity", C.c_uint64), ("seed", C.c_uint64),
        ("compat_lr", C.c_int32), ("compat_vocab_break", C.c_int32), ("max_rounds", C.c_uint32),
        ("device", C.c_int32), ("update_mode", C.c_int32), ("compat_walk", C.c_int32),
        ("trainer", C.c_int32), ("w2v_table_size", C.c_uint32),
    ]


class Corpus(C.Structure):
    _fields_ = [("round", C.c_uint32), ("max_len", C.c_uint32), ("n_walks", C.c_uint64),
                ("n_tokens", C.c_uint64), ("n_steps", C.c_uint64)]


class Stats(C.Structure):
    _fields_ = [
        ("v_num", C.c_uint64), ("e_num", C.c_uint64), ("vocab_size", C.c_uint64),
        ("rounds_done", C.c_uint32), ("n_unconverged", C.c_uint32),
        ("walk_steps", C.c_uint64), ("walk_tokens", C.c_uint64), ("train_words", C.c_uint64),
        ("train_kept", C.c_uint64), ("train_pairs", C.c_uint64), ("train_groups", C.c_uint64),
        ("overflow_used", C.c_uint64),
        ("build_ms", C.c_double), ("walk_kernel_ms", C.c_double), ("compact_kernel_ms", C.c_double),
        ("train_kernel_ms", C.c_double), ("eval_kernel_ms", C.c_double),
        ("walk_ms_total", C.c_double), ("train_ms_total", C.c_double), ("eval_ms_total", C.c_double),
        ("sgns_prep_ms", C.c_double), ("sgns_tile_ms", C.c_double), ("graph_ms_total", C.c_double),
        ("walk_tokens_total", C.c_uint64), ("train_rows_written", C.c_uint64), ("train_rows_lost", C.c_uint64),
    ]


BUDGET_FEATURES, BUDGET_PARAMS = 8, 4866


class BudgetModel(C.Structure):
    """felog_budget_model: MLP parameters, feature standardisation and box,
    target standardisation."""
    _fields_ = [("params", C.c_float * BUDGET_PARAMS),
                ("feat_mean", C.c_float * 8), ("feat_std", C.c_float * 8),
                ("feat_lo", C.c_float * 8), ("feat_hi", C.c_float * 8),
                ("targ_mean", C.c_float * 2), ("targ_std", C.c_float * 2)]


class TileShape(C.Structure):
    """felog_tile_shape: SGNS tile rows, waves per SIMD, workgroups per CU,
    centre windows per wavefront, dim split (columns per wavefront item of
    the update products) and (from the autotuner) the tile-kernel time."""
    _fields_ = [("rows", C.c_uint32), ("waves_per_simd", C.c_uint32), ("groups_per_cu", C.c_uint32),
                ("windows_per_wave", C.c_uint32), ("dim_split", C.c_uint32), ("tile_ms", C.c_float)]


SHAPE_FIELDS = ("rows", "waves_per_simd", "groups_per_cu", "windows_per_wave", "dim_split")


_lib = None


def header_symbols(path=HEADER_PATH):
    """Names of every entry point include/felog.h declares."""
    text = re.sub(r"/\*.*?\*/", "", open(path).read(), flags=re.S)
    return sorted(set(re.findall(r"\b(felog_\w+)\s*\(", text)))


def load_library(path=LIB_PATH):
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is missing: build it with `make -C info-centric_ge_amd`")
    lib = C.CDLL(path)
    P = C.c_void_p
    sig = {
        "felog_default_options": (None, [C.POINTER(Options)]),
        "felog_create": (C.c_int32, [C.POINTER(P), C.POINTER(Options)]),
        "felog_destroy": (C.c_int32, [P]),
        "felog_last_error": (C.c_char_p, [P]),
        "felog_load_edges": (C.c_int32, [P, C.c_char_p, C.c_uint32, C.c_int]),
        "felog_set_edges": (C.c_int32, [P, P, P, P, C.c_uint64, C.c_uint32, C.c_int]),
        "felog_build": (C.c_int32, [P]),
        "felog_walk_round": (C.c_int32, [P, C.c_uint32, C.POINTER(Corpus)]),
        "felog_train_round": (C.c_int32, [P, C.POINTER(Corpus)]),
        "felog_train_corpus": (C.c_int32, [P, P, P, C.c_uint64, C.c_uint32]),
        "felog_train_corpus_noalloc": (C.c_int32, [P, P, P, C.c_uint64, C.c_uint32]),
        "felog_evaluate": (C.c_int32, [P, C.POINTER(C.c_uint32), C.POINTER(C.c_int)]),
        "felog_run": (C.c_int32, [P]),
        "felog_corpus_copy_out": (C.c_int32, [P, P, P, P]),
        "felog_csr_copy_out": (C.c_int32, [P, P, P, P]),
        "felog_alias_copy_out": (C.c_int32, [P, P, P]),
        "felog_vocab_copy_out": (C.c_int32, [P, P, P]),
        "felog_unigram_copy_out": (C.c_int32, [P, P, P]),
        "felog_flags_copy_out": (C.c_int32, [P, P, P]),
        "felog_w2v_table_copy_out": (C.c_int32, [P, P]),
        "felog_kmeans_classes": (C.c_int32, [P, C.c_uint32, C.c_uint32, P]),
        "felog_read_edges_txt": (C.c_int32, [C.c_char_p, C.c_int, P, P, P, C.c_uint64, C.POINTER(C.c_uint64)]),
        "felog_load_edges_txt": (C.c_int32, [P, C.c_char_p, C.c_uint32, C.c_int, C.c_int]),
        "felog_write_edges_txt": (C.c_int32, [C.c_char_p, P, P, P, C.c_uint64]),
        "felog_write_edges_bin": (C.c_int32, [C.c_char_p, P, P, P, C.c_uint64]),
        "felog_common_neighbours": (C.c_int32, [P, P, P, C.c_uint64, P]),
        "felog_use_common_neighbours": (C.c_int32, [P, C.c_int]),
        "felog_cn_copy_out": (C.c_int32, [P, P]),
        "felog_link_auc": (C.c_int32, [P, P, P, C.c_uint64, P, C.POINTER(C.c_double)]),
        "felog_write_classes": (C.c_int32, [P, C.c_char_p, C.c_uint32]),
        "felog_embeddings_copy_out": (C.c_int32, [P, P, P]),
        "felog_embeddings_copy_in": (C.c_int32, [P, P, P]),
        "felog_write_walks": (C.c_int32, [P, C.c_char_p]),
        "felog_write_embeddings": (C.c_int32, [P, C.c_char_p, C.c_int]),
        "felog_stats_get": (C.c_int32, [P, C.POINTER(Stats)]),
        "felog_debug_sgns": (C.c_int32, [P, P, P, C.c_uint64, C.c_uint32, C.c_float, C.c_uint32,
                                         C.POINTER(C.c_uint32), P, P, P, P, P]),
        "felog_budget_features": (C.c_int32, [P, P]),
        "felog_budget_grad": (C.c_int32, [P, P, P, C.c_uint64, P, P, C.POINTER(C.c_double)]),
        "felog_budget_train": (C.c_int32, [P, P, P, C.c_uint64, C.c_uint32, C.c_uint32, C.c_float, C.c_uint64,
                                           P, P]),
        "felog_budget_set": (C.c_int32, [P, C.POINTER(BudgetModel)]),
        "felog_budget_copy_out": (C.c_int32, [P, P, P, C.POINTER(C.c_uint32)]),
        "felog_run_budget": (C.c_int32, [P]),
        "felog_budget_recheck": (C.c_int32, [P, C.c_int]),
        "felog_run_graph": (C.c_int32, [P, C.c_uint32]),
        "felog_corpus_get": (C.c_int32, [P, C.POINTER(Corpus)]),
        "felog_tile_shape_set": (C.c_int32, [P, C.POINTER(TileShape)]),
        "felog_tile_shape_get": (C.c_int32, [P, C.POINTER(TileShape)]),
        "felog_tune_sgns": (C.c_int32, [P, C.c_uint64, C.c_uint32, C.POINTER(TileShape), C.c_uint32,
                                        C.POINTER(C.c_uint32)]),
        "felog_visits_reset": (C.c_int32, [P]),
        "felog_visit_stats": (C.c_int32, [P, P]),
        "felog_track_lost_updates": (C.c_int32, [P, C.c_int32]),
        "felog_probe_random_reads": (C.c_int32, [C.c_int, C.c_uint64, C.c_uint64, C.c_uint32,
                                                 C.POINTER(C.c_double)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def read_edges_txt(path, columns=3):
    """Text edge list (src dst [data]) through the library's reader; no
    context needed.  Returns (src, dst, w) with w None for two columns."""
    lib = load_library()
    m = C.c_uint64()
    st = lib.felog_read_edges_txt(path.encode(), columns, None, None, None, 0, C.byref(m))
    if st != 0:
        raise FelogError(st, "felog_read_edges_txt", path)
    src = np.empty(max(m.value, 1), np.uint32)
    dst = np.empty(max(m.value, 1), np.uint32)
    w = np.empty(max(m.value, 1), np.float32) if columns == 3 else None
    st = lib.felog_read_edges_txt(path.encode(), columns, _ptr(src), _ptr(dst), _ptr(w), m.value, C.byref(m))
    if st != 0:
        raise FelogError(st, "felog_read_edges_txt", path)
    n = m.value
    return src[:n], dst[:n], (w[:n] if w is not None else None)


def write_edges(path, src, dst, w=None, binary=True):
    """Binary graph file (12- or 8-byte records) or text edge list."""
    lib = load_library()
    src = np.ascontiguousarray(src, np.uint32)
    dst = np.ascontiguousarray(dst, np.uint32)
    w = None if w is None else np.ascontiguousarray(w, np.float32)
    fn = lib.felog_write_edges_bin if binary else lib.felog_write_edges_txt
    st = fn(path.encode(), _ptr(src), _ptr(dst), _ptr(w), len(src))
    if st != 0:
        raise FelogError(st, "felog_write_edges", path)


def probe_random_reads(table_bytes, reads, read_bytes=32, device=0):
    """Random reads per second the device returns for rows of read_bytes at
    independent random rows of a table of table_bytes (felog_probe_random_reads):
    the ceiling the walk's one 32-byte record per step is measured against."""
    lib = load_library()
    r = C.c_double()
    st = lib.felog_probe_random_reads(device, int(table_bytes), int(reads), int(read_bytes), C.byref(r))
    if st != 0:
        raise FelogError(st, "felog_probe_random_reads", "")
    return r.value


def default_options():
    lib = load_library()
    o = Options()
    lib.felog_default_options(C.byref(o))
    return o


class Felog:
    """One context: one GPU's copy of the graph, tables and round state."""

    def __init__(self, **overrides):
        self.lib = load_library()
        self.opt = default_options()
        for k, v in overrides.items():
            if not hasattr(self.opt, k):
                raise KeyError(f"unknown option {k}")
            setattr(self.opt, k, v)
        self.ctx = C.c_void_p()
        s = self.lib.felog_create(C.byref(self.ctx), C.byref(self.opt))
        if s != 0:
            raise FelogError(s, "felog_create", "invalid options or no usable device")
        self.V = 0
        self.E = 0

    # -- helpers ---------------------------------------------------------
    def _check(self, status, where):
        if status != 0:
            raise FelogError(status, where, self.lib.felog_last_error(self.ctx).decode())

    def close(self):
        if self.ctx:
            self.lib.felog_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- graph -----------------------------------------------------------
    def set_edges(self, src, dst, w, v_num, make_undirected=True):
        src = np.ascontiguousarray(src, dtype=np.uint32)
        dst = np.ascontiguousarray(dst, dtype=np.uint32)
        w = np.ascontiguousarray(w, dtype=np.float32)
        self._check(self.lib.felog_set_edges(self.ctx, _ptr(src), _ptr(dst), _ptr(w), len(src), v_num,
                                             int(make_undirected)), "felog_set_edges")
        self.V = v_num
        self.E = len(src) * (2 if make_undirected else 1)

    def load_edges(self, path, v_num, make_undirected=True):
        self._check(self.lib.felog_load_edges(self.ctx, path.encode(), v_num, int(make_undirected)),
                    "felog_load_edges")
        self.V = v_num
        self.E = (os.path.getsize(path) // 12) * (2 if make_undirected else 1)

    def build(self):
        self._check(self.lib.felog_build(self.ctx), "felog_build")

    # -- hot path --------------------------------------------------------
    def walk_round(self, rnd):
        c = Corpus()
        self._check(self.lib.felog_walk_round(self.ctx, rnd, C.byref(c)), "felog_walk_round")
        return {"round": c.round, "max_len": c.max_len, "n_walks": c.n_walks,
                "n_tokens": c.n_tokens, "n_steps": c.n_steps}

    def train_round(self):
        c = Corpus()
        self._check(self.lib.felog_corpus_get(self.ctx, C.byref(c)), "felog_corpus_get")
        self._check(self.lib.felog_train_round(self.ctx, C.byref(c)), "felog_train_round")

    def train_corpus(self, offsets, tokens, rnd):
        offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
        tokens = np.ascontiguousarray(tokens, dtype=np.uint32)
        self._check(self.lib.felog_train_corpus(self.ctx, _ptr(offsets), _ptr(tokens), len(offsets) - 1, rnd),
                    "felog_train_corpus")

    def train_corpus_noalloc(self, offsets, tokens, rnd):
        """train_corpus without device allocation (the library's static
        scratch); a corpus beyond it raises with status ERR_CAPACITY."""
        offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
        tokens = np.ascontiguousarray(tokens, dtype=np.uint32)
        self._check(self.lib.felog_train_corpus_noalloc(self.ctx, _ptr(offsets), _ptr(tokens), len(offsets) - 1,
                                                        rnd), "felog_train_corpus_noalloc")

    def evaluate(self):
        nu = C.c_uint32()
        stop = C.c_int()
        self._check(self.lib.felog_evaluate(self.ctx, C.byref(nu), C.byref(stop)), "felog_evaluate")
        return nu.value, bool(stop.value)

    def run(self):
        self._check(self.lib.felog_run(self.ctx), "felog_run")

    # -- learned walk budget ---------------------------------------------
    def budget_features(self):
        out = np.empty((self.V, BUDGET_FEATURES), np.float32)
        self._check(self.lib.felog_budget_features(self.ctx, _ptr(out)), "felog_budget_features")
        return out

    def budget_grad(self, X, T, params):
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        params = np.ascontiguousarray(params, np.float32)
        grad = np.empty(BUDGET_PARAMS, np.float32)
        loss = C.c_double()
        self._check(self.lib.felog_budget_grad(self.ctx, _ptr(X), _ptr(T), len(X), _ptr(params), _ptr(grad),
                                               C.byref(loss)), "felog_budget_grad")
        return grad, loss.value

    def budget_train(self, X, T, params, epochs, batch=4096, lr=1e-3, seed=1):
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        p = np.ascontiguousarray(params, np.float32).copy()
        losses = np.zeros(epochs, np.float64)
        self._check(self.lib.felog_budget_train(self.ctx, _ptr(X), _ptr(T), len(X), epochs, batch, lr, seed,
                                                _ptr(p), _ptr(losses)), "felog_budget_train")
        return p, losses

    def budget_set(self, model):
        self._check(self.lib.felog_budget_set(self.ctx, None if model is None else C.byref(model)),
                    "felog_budget_set")

    def budget(self):
        cnt = np.empty(self.V, np.uint16)
        ln = np.empty(self.V, np.uint16)
        r = C.c_uint32()
        self._check(self.lib.felog_budget_copy_out(self.ctx, _ptr(cnt), _ptr(ln), C.byref(r)),
                    "felog_budget_copy_out")
        return cnt, ln, r.value

    def run_budget(self):
        self._check(self.lib.felog_run_budget(self.ctx), "felog_run_budget")

    def budget_recheck(self, on=True):
        """Score the budgeted vertices whose count ends at min_R once there;
        those not converged walk one more round."""
        self._check(self.lib.felog_budget_recheck(self.ctx, int(on)), "felog_budget_recheck")

    def run_graph(self, max_rounds=0):
        """run() / run_budget() with each round replayed from a hipGraph;
        at most max_rounds more rounds (0: until the stop rule)."""
        self._check(self.lib.felog_run_graph(self.ctx, max_rounds), "felog_run_graph")

    # -- SGNS tile shape ---------------------------------------------------
    @staticmethod
    def _shape_dict(t):
        return {"rows": t.rows, "waves_per_simd": t.waves_per_simd, "groups_per_cu": t.groups_per_cu,
                "windows_per_wave": t.windows_per_wave, "dim_split": t.dim_split, "tile_ms": t.tile_ms}

    def tile_shape(self):
        t = TileShape()
        self._check(self.lib.felog_tile_shape_get(self.ctx, C.byref(t)), "felog_tile_shape_get")
        return self._shape_dict(t)

    def set_tile_shape(self, rows=0, waves_per_simd=0, groups_per_cu=0, windows_per_wave=0, dim_split=0):
        """Zero fields take the defaults; set_tile_shape() restores all."""
        t = TileShape(rows, waves_per_simd, groups_per_cu, windows_per_wave, dim_split, 0.0)
        self._check(self.lib.felog_tile_shape_set(self.ctx, C.byref(t)), "felog_tile_shape_set")

    def tune_sgns(self, sample=0, reps=2):
        """Time every tile shape on the current walk corpus (first `sample`
        sentences, 0: all) and install the fastest; returns every shape
        tried with its mean tile-kernel time."""
        res = (TileShape * 128)()
        n = C.c_uint32()
        self._check(self.lib.felog_tune_sgns(self.ctx, sample, reps, res, 128, C.byref(n)), "felog_tune_sgns")
        return [self._shape_dict(res[i]) for i in range(min(n.value, 128))]

    def visits_reset(self):
        self._check(self.lib.felog_visits_reset(self.ctx), "felog_visits_reset")

    def track_lost_updates(self, on=True):
        """Count the tile trainer's write-backs that replace another tile's
        update (stats: train_rows_written, train_rows_lost)."""
        self._check(self.lib.felog_track_lost_updates(self.ctx, int(bool(on))), "felog_track_lost_updates")

    def visit_stats(self):
        out = np.zeros(4, np.float64)
        self._check(self.lib.felog_visit_stats(self.ctx, _ptr(out)), "felog_visit_stats")
        return {"kl": out[0], "entropy": out[1], "unvisited": int(out[2]), "visits": int(out[3])}

    # -- read-back -------------------------------------------------------
    def corpus(self):
        c = Corpus()
        self._check(self.lib.felog_corpus_get(self.ctx, C.byref(c)), "felog_corpus_get")
        off = np.empty(c.n_walks + 1, np.uint64)
        tok = np.empty(max(c.n_tokens, 1), np.uint32)
        wid = np.empty(max(c.n_walks, 1), np.uint32)
        self._check(self.lib.felog_corpus_copy_out(self.ctx, _ptr(off), _ptr(tok), _ptr(wid)),
                    "felog_corpus_copy_out")
        return off, tok[:c.n_tokens], wid[:c.n_walks]

    def csr(self, weights=True):
        """row_ptr, col and (unless weights=False) w of the CSR."""
        rp = np.empty(self.V + 1, np.uint64)
        col = np.empty(max(self.E, 1), np.uint32)
        w = np.empty(max(self.E, 1), np.float32) if weights else None
        self._check(self.lib.felog_csr_copy_out(self.ctx, _ptr(rp), _ptr(col), _ptr(w)), "felog_csr_copy_out")
        return rp, col[:self.E], (w[:self.E] if weights else None)

    def alias(self):
        t = np.empty(max(self.E, 1), np.uint32)
        a = np.empty(max(self.E, 1), np.uint32)
        self._check(self.lib.felog_alias_copy_out(self.ctx, _ptr(t), _ptr(a)), "felog_alias_copy_out")
        return t[:self.E], a[:self.E]

    def vocab(self):
        order = np.empty(self.V, np.uint32)
        deg = np.empty(self.V, np.uint32)
        self._check(self.lib.felog_vocab_copy_out(self.ctx, _ptr(order), _ptr(deg)), "felog_vocab_copy_out")
        return order, deg

    def unigram(self):
        t = np.empty(self.V, np.uint32)
        a = np.empty(self.V, np.uint32)
        self._check(self.lib.felog_unigram_copy_out(self.ctx, _ptr(t), _ptr(a)), "felog_unigram_copy_out")
        return t, a

    def flags(self):
        f = np.empty(self.V, np.uint8)
        s = np.empty(self.V, np.float32)
        self._check(self.lib.felog_flags_copy_out(self.ctx, _ptr(f), _ptr(s)), "felog_flags_copy_out")
        return f, s

    def load_edges_txt(self, path, v_num, make_undirected=True, weighted=True):
        self._check(self.lib.felog_load_edges_txt(self.ctx, path.encode(), v_num, int(make_undirected),
                                                  int(weighted)), "felog_load_edges_txt")
        self.V = v_num
        m = C.c_uint64()
        self.lib.felog_read_edges_txt(path.encode(), 3 if weighted else 2, None, None, None, 0, C.byref(m))
        self.E = m.value * (2 if make_undirected else 1)

    def common_neighbours(self, src, dst):
        src = np.ascontiguousarray(src, np.uint32)
        dst = np.ascontiguousarray(dst, np.uint32)
        out = np.empty(max(len(src), 1), np.int32)
        self._check(self.lib.felog_common_neighbours(self.ctx, _ptr(src), _ptr(dst), len(src), _ptr(out)),
                    "felog_common_neighbours")
        return out[:len(src)]

    def use_common_neighbours(self, on=True):
        self._check(self.lib.felog_use_common_neighbours(self.ctx, int(on)), "felog_use_common_neighbours")

    def cn(self):
        out = np.empty(max(self.E, 1), np.float32)
        self._check(self.lib.felog_cn_copy_out(self.ctx, _ptr(out)), "felog_cn_copy_out")
        return out[:self.E]

    def link_auc(self, pairs, labels):
        """(AUC, scores) of n vertex pairs (n x 2) with 0/1 labels."""
        pairs = np.ascontiguousarray(pairs, np.uint32).reshape(-1, 2)
        labels = np.ascontiguousarray(labels, np.uint8)
        scores = np.empty(len(pairs), np.float32)
        auc = C.c_double()
        self._check(self.lib.felog_link_auc(self.ctx, _ptr(pairs), _ptr(labels), len(pairs), _ptr(scores),
                                            C.byref(auc)), "felog_link_auc")
        return auc.value, scores

    def kmeans_classes(self, classes, iters=10):
        """k-means class of every vocabulary entry (vocabulary order)."""
        vs = self.stats()["vocab_size"]
        out = np.empty(max(vs, 1), np.int32)
        self._check(self.lib.felog_kmeans_classes(self.ctx, classes, iters, _ptr(out)), "felog_kmeans_classes")
        return out[:vs]

    def write_classes(self, path, classes):
        self._check(self.lib.felog_write_classes(self.ctx, path.encode(), classes), "felog_write_classes")

    def w2v_table(self):
        """The per-pair trainers' unigram table (vocabulary indices)."""
        n = self.opt.w2v_table_size or 100_000_000
        t = np.empty(n, np.int32)
        self._check(self.lib.felog_w2v_table_copy_out(self.ctx, _ptr(t)), "felog_w2v_table_copy_out")
        return t

    def embeddings(self):
        d = self.opt.size
        s0 = np.empty((self.V, d), np.float32)
        s1 = np.empty((self.V, d), np.float32)
        self._check(self.lib.felog_embeddings_copy_out(self.ctx, _ptr(s0), _ptr(s1)), "felog_embeddings_copy_out")
        return s0, s1

    def set_embeddings(self, syn0=None, syn1=None):
        a = None if syn0 is None else np.ascontiguousarray(syn0, dtype=np.float32)
        b = None if syn1 is None else np.ascontiguousarray(syn1, dtype=np.float32)
        self._check(self.lib.felog_embeddings_copy_in(self.ctx, _ptr(a), _ptr(b)), "felog_embeddings_copy_in")

    def write_walks(self, prefix):
        self._check(self.lib.felog_write_walks(self.ctx, prefix.encode()), "felog_write_walks")

    def write_embeddings(self, path, binary=False):
        self._check(self.lib.felog_write_embeddings(self.ctx, path.encode(), int(binary)),
                    "felog_write_embeddings")

    def stats(self):
        s = Stats()
        self._check(self.lib.felog_stats_get(self.ctx, C.byref(s)), "felog_stats_get")
        return {name: getattr(s, name) for name, _ in Stats._fields_}

    def debug_sgns(self, offsets, tokens, rnd, alpha, max_groups):
        offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
        tokens = np.ascontiguousarray(tokens, dtype=np.uint32)
        d = self.opt.size
        G = max_groups
        groups = np.zeros((G, 5), np.uint32)
        rows = np.zeros((G, 96), np.uint32)
        logits = np.zeros((G, 64, 32), np.float32)
        dv = np.zeros((G, 64, d), np.float32)
        du = np.zeros((G, 32, d), np.float32)
        n = C.c_uint32()
        self._check(self.lib.felog_debug_sgns(self.ctx, _ptr(offsets), _ptr(tokens), len(offsets) - 1, rnd,
                                              alpha, G, C.byref(n), _ptr(groups), _ptr(rows), _ptr(logits),
                                              _ptr(dv), _ptr(du)), "felog_debug_sgns")
        live = np.nonzero(groups[: n.value, 2] > 0)[0]  # slots of sentences without a tile stay empty
        return {"groups": groups[live], "rows": rows[live], "logits": logits[live], "dv": dv[live], "du": du[live]}
