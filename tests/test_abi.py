"""The C-ABI library and its host-side surface, without a GPU.

Loads libfelog_hip.so, checks that every entry point include/felog.h
declares is exported, that the option defaults are the reference's, that
errors come back as status codes (no GPU here: creating a context must fail
cleanly, not abort), and the on-disk formats of the generators.
"""
import ctypes
import os
import struct
import subprocess

import numpy as np
import pytest

import felog
import graphgen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    lib = felog.load_library()
    names = felog.header_symbols()
    assert len(names) >= 50
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", felog.LIB_PATH], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    assert set(names) <= exported
    # and the library exports no felog_ entry point the header leaves out
    assert {n for n in exported if n.startswith("felog_")} == set(names)


def test_default_options_are_the_reference_defaults():
    o = felog.default_options()
    # include/walk.hpp:416-417 (20, 5); src/word2vec.cu:162-168 (size 100,
    # window 5, negative 5, sample 1e-3, min-count 5);
    # include/option_helper.hpp:138 (--batch_size 16384).  alpha defaults to
    # 0.0025, a tenth of word2vec's 0.025: tiles sum their pair gradients.
    assert (o.min_L, o.min_R, o.size, o.window, o.negative, o.min_count, o.eval_batch) == (20, 5, 100, 5, 5, 5, 16384)
    assert abs(o.sample - 1e-3) < 1e-9 and abs(o.alpha - 0.0025) < 1e-9
    assert o.max_walk_len == 1000 and o.batch_size == 0 and o.precision == felog.FP32


def test_options_struct_layout_matches_header():
    assert ctypes.sizeof(felog.Options) == 112
    assert felog.Options.seed.offset == 72 and felog.Options.device.offset == 92
    assert felog.Options.update_mode.offset == 96 and felog.Options.compat_walk.offset == 100
    assert ctypes.sizeof(felog.TileShape) == 24 and felog.TileShape.tile_ms.offset == 20
    assert ctypes.sizeof(felog.Stats) == 8 * 3 + 4 * 2 + 8 * 7 + 8 * 11 + 8 * 3


@pytest.mark.parametrize("bad", [dict(size=0), dict(size=600), dict(window=0), dict(window=40),
                                 dict(negative=30), dict(precision=7), dict(alpha=0.0)])
def test_invalid_options_are_rejected(bad):
    with pytest.raises(felog.FelogError) as e:
        felog.Felog(**bad)
    assert e.value.status in (-1, -3)


def test_null_context_calls_return_errors():
    lib = felog.load_library()
    assert lib.felog_build(None) == -1
    assert lib.felog_destroy(None) == -1
    assert lib.felog_stats_get(None, None) == -1
    assert lib.felog_last_error(None) == b"null context"


def test_random_read_probe_rejects_bad_arguments():
    # argument checks run before any HIP call, so no GPU is needed
    lib = felog.load_library()
    r = ctypes.c_double()
    assert lib.felog_probe_random_reads(0, 1 << 20, 1000, 24, ctypes.byref(r)) == -1
    assert lib.felog_probe_random_reads(0, 16, 1000, 32, ctypes.byref(r)) == -1
    assert lib.felog_probe_random_reads(0, 1 << 20, 0, 32, ctypes.byref(r)) == -1
    assert lib.felog_probe_random_reads(0, 1 << 40, 1000, 16, ctypes.byref(r)) == -1
    assert lib.felog_probe_random_reads(0, 1 << 20, 1000, 32, None) == -1


def test_binary_edge_format_roundtrip(tmp_path):
    s, d, w = graphgen.erdos_renyi(300, 1000, seed=2)
    p = tmp_path / "g.bin"
    graphgen.write_bin(str(p), s, d, w)
    raw = p.read_bytes()
    assert len(raw) == 12 * len(s)  # include/storage.hpp: sizeof(Edge<float>) == 12, no header
    a, b, c = struct.unpack_from("<IIf", raw, 12 * 5)
    assert (a, b) == (s[5], d[5]) and np.float32(c) == w[5]
    s2, d2, w2 = graphgen.read_bin(str(p))
    assert np.array_equal(s, s2) and np.array_equal(d, d2) and np.array_equal(w, w2)


@pytest.mark.parametrize("gen", ["erdos_renyi", "rmat", "chung_lu"])
def test_generators_are_simple_and_seeded(gen):
    fn = getattr(graphgen, gen)
    s, d, w = fn(5000, 20000, seed=3)
    s2, d2, w2 = fn(5000, 20000, seed=3)
    assert np.array_equal(s, s2) and np.array_equal(w, w2)
    assert np.all(s != d)
    key = np.minimum(s, d).astype(np.int64) * 5000 + np.maximum(s, d)
    assert len(np.unique(key)) == len(key)
    deg = np.bincount(s, minlength=5000) + np.bincount(d, minlength=5000)
    assert deg.min() >= 1
    assert np.all((w > 0) & (w <= 1))
    assert abs(len(s) - 20000) < 0.02 * 20000


def test_lognormal_weights_and_lp_split():
    s, d, w = graphgen.rmat(3000, 12000, seed=4, weights="lognormal")
    assert np.all(w > 0) and w.max() > 1
    ts, td, tw, pos, neg = graphgen.lp_split(s, d, w, 3000, frac=0.1, seed=1)
    assert len(pos) == len(neg) == int(0.1 * len(s))
    assert len(ts) + len(pos) == len(s)
    deg = np.bincount(ts, minlength=3000) + np.bincount(td, minlength=3000)
    orig = np.bincount(s, minlength=3000) + np.bincount(d, minlength=3000)
    assert np.all(deg[orig > 0] >= 1)  # held-out edges never isolate a vertex
    ek = set((np.minimum(s, d).astype(np.int64) * 3000 + np.maximum(s, d)).tolist())
    nk = (np.minimum(neg[:, 0], neg[:, 1]).astype(np.int64) * 3000 + np.maximum(neg[:, 0], neg[:, 1])).tolist()
    assert not (set(nk) & ek) and len(set(nk)) == len(nk) and np.all(neg[:, 0] != neg[:, 1])


def test_cli_usage_without_arguments():
    exe = os.path.join(ROOT, "info-centric_ge_amd", "felog")
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 1 and "usage" in r.stderr
