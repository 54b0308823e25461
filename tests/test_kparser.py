"""The .k mesh parser (cudafemproject_amd.kparser) against the reference's
parser output (golden/kmesh/*/ref_*.txt, written by golden/make_kmesh_golden.py
from the reference's mesh_parser_lib) and against the legacy prepared-mesh
fixtures the reference's own tests hold (golden/9task_3, golden/test_rect_pcg)."""
import os

import numpy as np
import pytest

from conftest import GOLDEN
from cudafemproject_amd.kparser import KMesh, run_parser

FILES = ["nodes.txt", "elements.txt", "constraints.txt", "loads.txt", "stress.txt"]


@pytest.mark.parametrize("name,dim", [("small_test", 3), ("bulk_small", 3), ("test_rect_pcg", 2)])
def test_byte_identical_to_reference_parser(tmp_path, name, dim):
    d = os.path.join(GOLDEN, "kmesh", name)
    run_parser(str(tmp_path), str(tmp_path), os.path.join(d, name + ".k"), name, dim)
    for f in FILES:
        with open(tmp_path / f) as a, open(os.path.join(d, "ref_" + f)) as b:
            assert a.read() == b.read(), f


def _nums(path):
    with open(path) as f:
        return np.array([float(t) for t in f.read().split()])


@pytest.mark.parametrize("name", ["9task_3", "test_rect_pcg"])
def test_matches_legacy_prepared_fixture(tmp_path, name):
    """The fixtures predate the current stress.txt layout (no interior-node
    column, unit normals); every other file must agree in value."""
    raw = os.path.join(GOLDEN, name, name + ".k") if name == "9task_3" else \
        os.path.join(GOLDEN, "kmesh", name, name + ".k")
    m = KMesh(raw, 2)
    m.write(str(tmp_path))
    fx = os.path.join(GOLDEN, name)
    for f in ["nodes.txt", "elements.txt", "constraints.txt", "loads.txt"]:
        np.testing.assert_array_equal(_nums(tmp_path / f), _nums(os.path.join(fx, f)), err_msg=f)
    new = [ln.split() for ln in open(tmp_path / "stress.txt").read().splitlines()[1:]]
    old = [ln.split() for ln in open(os.path.join(fx, "stress.txt")).read().splitlines()[1:]]
    assert len(new) == len(old)
    for a, b in zip(new, old):
        assert a[:2] == b[:2] and a[3] == b[2]          # segment nodes, adjacent element
        n_new = np.array([float(v) for v in a[4:6]])
        n_old = np.array([float(v) for v in b[3:5]])
        np.testing.assert_allclose(n_new / np.linalg.norm(n_new), n_old / np.linalg.norm(n_old),
                                   atol=1e-12)
        assert float(a[6]) == float(b[5])


def test_normals_point_away_from_interior_node(tmp_path):
    m = KMesh(os.path.join(GOLDEN, "kmesh", "small_test", "small_test.k"), 3)
    m.write(str(tmp_path))
    X = _nums(tmp_path / "nodes.txt")[1:].reshape(-1, 3)
    for ln in open(tmp_path / "stress.txt").read().splitlines()[1:]:
        t = ln.split()
        a, b, c, inner = (int(v) for v in t[:4])
        n = np.array([float(v) for v in t[5:8]])
        assert np.dot(n, X[inner] - X[a]) < 0
        # un-normalised cross product of the segment edges: |n| = 2 * area
        assert np.isclose(np.linalg.norm(n), np.linalg.norm(np.cross(X[b] - X[a], X[c] - X[a])))


def test_segment_without_element_raises(tmp_path):
    src = open(os.path.join(GOLDEN, "kmesh", "test_rect_pcg", "test_rect_pcg.k")).read()
    bad = tmp_path / "bad.k"
    i = src.index("*LOAD_SEGMENT")
    j = src.index("\n", src.index("\n", i) + 1)
    line = src[src.index("\n", i) + 1:j].split()
    line[3], line[4] = "999991", "999992"
    bad.write_text(src[:src.index("\n", i) + 1] + "  ".join(line) + src[j:])
    with pytest.raises((ValueError, KeyError)):
        KMesh(str(bad), 2).write(str(tmp_path / "out"))
