"""CPU checks of the graph tools either side of the path (src/tools/ and the
text edge I/O of include/storage.hpp): the oracle's common-neighbour count,
the library's text reader and writers, gconverter, gtrain's split and the
lpauc scorer."""
import os
import subprocess

import numpy as np
import pytest

import felog
import graphgen
import orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TOOLS = os.path.join(ROOT, "info-centric_ge_amd")


def tool(name, *args):
    r = subprocess.run([os.path.join(TOOLS, name), *map(str, args)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    return r.stdout


def test_common_neighbours_oracle():
    """Distinct common neighbours, the pair's own vertices excluded: a hand
    example, then a random multigraph with self loops against Python sets."""
    # 0-1, 0-2, 1-2, 1-3, 2-3, 0-3: every pair of {0..3} shares the other two
    src = np.array([0, 0, 1, 1, 2, 0], np.uint32)
    dst = np.array([1, 2, 2, 3, 3, 3], np.uint32)
    rp, col, _ = orc.build_csr(5, src, dst, np.ones(6, np.float32), True)
    assert orc.common_neighbours(5, rp, col, [0, 1, 0, 4], [1, 3, 4, 4]).tolist() == [2, 2, 0, 0]
    rng = np.random.default_rng(3)
    V = 60
    src = rng.integers(0, V, 700).astype(np.uint32)
    dst = rng.integers(0, V, 700).astype(np.uint32)
    rp, col, _ = orc.build_csr(V, src, dst, np.ones(700, np.float32), True)
    nb = [set(col[rp[v]:rp[v + 1]].tolist()) for v in range(V)]
    u = rng.integers(0, V, 500)
    v = rng.integers(0, V, 500)
    want = [len((nb[a] & nb[b]) - {a, b}) for a, b in zip(u, v)]
    assert orc.common_neighbours(V, rp, col, u, v).tolist() == want


def test_text_edges_roundtrip(tmp_path):
    src, dst, w = graphgen.erdos_renyi(500, 2000, seed=4)
    p = str(tmp_path / "e.txt")
    felog.write_edges(p, src, dst, w, binary=False)
    s2, d2, w2 = felog.read_edges_txt(p, 3)
    assert np.array_equal(s2, src) and np.array_equal(d2, dst)
    # "%f" keeps six decimals; strtof rounds that text to the nearest float
    assert np.array_equal(w2, np.array([float(f"{x:.6f}") for x in w], np.float32))
    felog.write_edges(p, src, dst, None, binary=False)
    s3, d3, w3 = felog.read_edges_txt(p, 2)
    assert np.array_equal(s3, src) and np.array_equal(d3, dst) and w3 is None
    # reading stops at the first record that does not parse, as fscanf does
    open(p, "w").write("1 2 0.5\n3 4 0.25\n5 x 1\n6 7 1\n")
    s4, d4, w4 = felog.read_edges_txt(p, 3)
    assert s4.tolist() == [1, 3] and d4.tolist() == [2, 4] and w4.tolist() == [0.5, 0.25]


@pytest.mark.parametrize("weighted", [True, False])
def test_gconverter_matches_the_binary_format(tmp_path, weighted):
    src, dst, w = graphgen.rmat(2000, 8000, seed=5)
    txt, out = str(tmp_path / "g.txt"), str(tmp_path / "g.bin")
    with open(txt, "w") as f:
        for a, b, x in zip(src, dst, w):
            f.write(f"{a} {b} {x:.6f}\n" if weighted else f"{a} {b}\n")
    msg = tool("gconverter", "-i", txt, "-o", out, "-s", "weighted" if weighted else "unweighted")
    assert f"{len(src)} edges are converted" in msg
    raw = open(out, "rb").read()
    if weighted:
        ref = np.array([float(f"{x:.6f}") for x in w], np.float32)
        rec = np.frombuffer(raw, graphgen.EDGE_DTYPE)
        assert np.array_equal(rec["src"], src) and np.array_equal(rec["dst"], dst)
        assert np.array_equal(rec["data"], ref)
    else:
        rec = np.frombuffer(raw, np.uint32).reshape(-1, 2)
        assert len(raw) == 8 * len(src)
        assert np.array_equal(rec[:, 0], src) and np.array_equal(rec[:, 1], dst)


def test_gtrain_split(tmp_path):
    """gtrain's files: the held-out edges and the train edges are the input,
    every held-out edge leaves both endpoints a neighbour, negatives are
    distinct non-edges, labels positives first; the same seed gives the
    same files."""
    n = 3000
    src, dst, w = graphgen.rmat(n, 12_000, seed=6)
    data = (np.arange(len(src)) % 7).astype(np.uint32)
    txt = str(tmp_path / "in.txt")
    with open(txt, "w") as f:
        for a, b, x in zip(src, dst, data):
            f.write(f"{a}\t{b}\t{x}\n")
    outs = []
    for rep in range(2):
        lp, te, tl = (str(tmp_path / f"{k}{rep}.txt") for k in ("train", "edges", "labels"))
        tool("gtrain", "-i", txt, "-o", lp, "-e", te, "-l", tl, "-s", "0.1", "-v", n, "-t", 4, "--seed", 11)
        outs.append([open(p).read() for p in (lp, te, tl)])
    assert outs[0] == outs[1]
    train = np.array([l.split() for l in outs[0][0].splitlines()], np.int64)
    pairs = np.array([l.split() for l in outs[0][1].splitlines()], np.int64)
    labels = np.array([int(l) for l in outs[0][2].splitlines()])
    k = int(np.float32(0.1 * len(src)))
    assert len(pairs) == len(labels) == 2 * k and labels[:k].all() and not labels[k:].any()
    assert len(train) == len(src) - k
    inp = sorted(zip(src.tolist(), dst.tolist(), data.tolist()))
    pos = pairs[:k]
    # train triples plus the held-out pairs (their data looked up) = input
    got = sorted([tuple(r) for r in train.tolist()])
    key = {(a, b): x for a, b, x in inp}
    got = sorted(got + [(a, b, key[(a, b)]) for a, b in pos.tolist()])
    assert got == inp
    deg = np.bincount(train[:, 0], minlength=n) + np.bincount(train[:, 1], minlength=n)
    assert (deg[pos[:, 0]] >= 1).all() and (deg[pos[:, 1]] >= 1).all()
    edges = set(zip(src.tolist(), dst.tolist())) | set(zip(dst.tolist(), src.tolist()))
    neg = [tuple(x) for x in pairs[k:].tolist()]
    assert len(set(neg)) == k and not any(p in edges for p in neg)


@pytest.mark.parametrize("binary", [False, True])
def test_lpauc_scores_like_the_oracle(tmp_path, binary):
    rng = np.random.default_rng(7)
    V, d = 400, 16
    emb = rng.standard_normal((V, d)).astype(np.float32)
    ids = rng.permutation(V)
    path = str(tmp_path / "emb.txt")
    with open(path, "wb") as f:
        f.write(f"{V} {d}\n".encode())
        for v in ids:
            f.write(f"{v} ".encode())
            if binary:
                f.write(emb[v].tobytes())
            else:
                f.write(" ".join(f"{x:f}" for x in emb[v]).encode() + b" ")
            f.write(b"\n")
    pairs = rng.integers(0, V, (600, 2))
    pairs[:50] = pairs[50:100]  # tied scores
    labels = (rng.random(600) < 0.4).astype(int)
    te, tl = str(tmp_path / "te.txt"), str(tmp_path / "tl.txt")
    np.savetxt(te, pairs, fmt="%d")
    np.savetxt(tl, labels, fmt="%d")
    out = tool("lpauc", "-emb", path, "-e", te, "-l", tl, "-binary", int(binary))
    got = float(out.split()[1])
    e = emb.astype(np.float64) if binary else np.array([[float(f"{x:f}") for x in r] for r in emb])
    a, b = e[pairs[:, 0]], e[pairs[:, 1]]
    s = (a * b).sum(1) / np.sqrt((a * a).sum(1) * (b * b).sum(1))
    ref = orc.auc(s[labels == 1].astype(np.float32), s[labels == 0].astype(np.float32))
    assert abs(got - ref) < 1e-4
