"""The text -> CSR converter with the reference's semantics
(xbfs_graph_from_tuples and the text_to_bin command line,
CSRGenerator/tuple_text_to_bin.cpp:59-426): header lines skipped, ids read as
(src, dst) pairs, rows in file order, duplicates and self-loops kept, the
reversed entry appended when asked.  Pinned by a hand-computed edge list, by
the reference's own converter compiled from its source (oracle/_ref, where
the reference tree is present) on seeded edge lists, and by both loaders."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from xbfs_amd import Graph, XbfsError

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "xbfs_amd", "text_to_bin")
REF_CONV = os.path.join(ROOT, "oracle", "_ref", "ref_text_to_bin")

HAND = "# header line 1\n% header 2\n0 1\n1 2\n0 1\n2 2\n3 0\n"


def test_hand_computed_not_reversed(tmp_path):
    """Pairs (0,1) (1,2) (0,1) (2,2) (3,0): n = 4, m = 5; row 0 = [1, 1] (the
    duplicate kept), row 2 = [2] (the self-loop kept)."""
    p = tmp_path / "e.txt"
    p.write_text(HAND)
    g, w = Graph.from_tuples(str(p), reverse=False, skip_lines=2)
    assert g.n == 4 and g.m == 5
    assert g.beg_pos.tolist() == [0, 2, 3, 4, 5]
    assert g.adj.tolist() == [1, 1, 2, 2, 0]
    assert ((w >= 1) & (w <= 63)).all()


def test_hand_computed_reversed():
    """reverse = 1 appends src to dst's row after each pair, in file order:
    (0,1): r0+=1 r1+=0; (1,2): r1+=2 r2+=1; (0,1): r0+=1 r1+=0; (2,2): r2+=2
    twice; (3,0): r3+=0 r0+=3 -> rows [1,1,3] [0,2,0] [1,2,2] [0], m = 10."""
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "e.txt")
        open(p, "w").write(HAND)
        g, w = Graph.from_tuples(p, reverse=True, skip_lines=2)
    assert g.n == 4 and g.m == 10
    assert g.beg_pos.tolist() == [0, 3, 6, 9, 10]
    assert g.adj.tolist() == [1, 1, 3, 0, 2, 0, 1, 2, 2, 0]
    # a reversed pair carries one weight on both of its entries
    assert w[0] == w[3] and w[4] == w[6] and w[7] == w[8] and w[9] == w[2]


def test_errors(tmp_path):
    p = tmp_path / "e.txt"
    p.write_text(HAND)
    with pytest.raises(XbfsError) as e:
        Graph.from_tuples(str(p), skip_lines=8)  # more header lines than the file has
    assert e.value.code == 5
    q = tmp_path / "bad.txt"
    q.write_text("0 1\n1 x\n")
    with pytest.raises(XbfsError) as e:
        Graph.from_tuples(str(q))
    assert e.value.code == 5
    r = tmp_path / "neg.txt"
    r.write_text("0 -1\n")
    with pytest.raises(XbfsError):
        Graph.from_tuples(str(r))
    with pytest.raises(XbfsError) as e:
        Graph.from_tuples(str(tmp_path / "missing.txt"))
    assert e.value.code == 4


def run_cli(*args, cwd=None):
    return subprocess.run([CLI, *map(str, args)], capture_output=True, text=True, timeout=60, cwd=cwd)


def test_cli_argument_count_and_outputs(tmp_path):
    r = run_cli("only", "three", "args")
    assert r.returncode == 1 and "Wrong input" in r.stdout
    p = tmp_path / "e.txt"
    p.write_text(HAND)
    r = run_cli(p, 1, 2, 4)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "edge count: 10" in r.stdout and "vert count: 4" in r.stdout
    beg = np.fromfile(str(p) + "_beg_pos.bin", np.int64)
    adj = np.fromfile(str(p) + "_csr.bin", np.int32)
    wt = np.fromfile(str(p) + "_weight.bin", np.int32)
    assert beg.tolist() == [0, 3, 6, 9, 10]
    assert adj.tolist() == [1, 1, 3, 0, 2, 0, 1, 2, 2, 0]
    assert wt.shape == (10,) and ((wt >= 1) & (wt <= 63)).all()
    # the loader reads the three files back (weights narrowed to int8)
    g = Graph.load(str(p) + "_beg_pos.bin", str(p) + "_csr.bin", str(p) + "_weight.bin")
    assert g.beg_pos.tolist() == beg.tolist() and g.adj.tolist() == adj.tolist()


def random_tuples(seed, pairs=400, nv=60):
    rng = np.random.default_rng(seed)
    e = rng.integers(0, nv, size=(pairs, 2))
    e[::17, 1] = e[::17, 0]          # self-loops
    e[5::23] = e[4::23][: len(e[5::23])]  # duplicates
    lines = ["% a header", "# another"] + [f"{a}\t{b}" if i % 3 else f"{a} {b}" for i, (a, b) in enumerate(e)]
    return "\n".join(lines) + "\n"


@pytest.mark.parametrize("reverse", [0, 1])
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_matches_reference_converter(tmp_path, reverse, seed):
    """Row for row equal to the reference's converter (its int64 beg_pos and
    int64 adjacency) on seeded edge lists with header lines, tabs, duplicates
    and self-loops."""
    if not os.path.exists(REF_CONV):
        pytest.skip("reference converter not built (reference tree absent)")
    text = random_tuples(seed)
    d_ref, d_our = tmp_path / "ref", tmp_path / "our"
    d_ref.mkdir()
    d_our.mkdir()
    (d_ref / "t.txt").write_text(text)
    (d_our / "t.txt").write_text(text)
    r = subprocess.run([REF_CONV, str(d_ref / "t.txt"), str(reverse), "2", "1"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    r2 = run_cli(d_our / "t.txt", reverse, 2, 1)
    assert r2.returncode == 0, r2.stdout + r2.stderr
    rb = np.fromfile(str(d_ref / "t.txt") + "_beg_pos.bin", np.int64)
    ra = np.fromfile(str(d_ref / "t.txt") + "_csr.bin", np.int64)
    ob = np.fromfile(str(d_our / "t.txt") + "_beg_pos.bin", np.int64)
    oa = np.fromfile(str(d_our / "t.txt") + "_csr.bin", np.int32)
    assert rb.tolist() == ob.tolist()
    assert ra.tolist() == oa.astype(np.int64).tolist()
    # the reference loader reads our int32 files
    if os.path.exists(O.REF_LOADER):
        n, m, *_ = O.ref_loader(str(d_our / "t.txt") + "_beg_pos.bin", str(d_our / "t.txt") + "_csr.bin")
        assert (n, m) == (ob.shape[0] - 1, oa.shape[0])
