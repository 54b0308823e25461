"""CPU tests of the host side: the C-ABI library exports, the mesh planner,
the mesh generators and the prepared-mesh format."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from cudafemproject_amd import problem as P
from cudafemproject_amd._lib import LIB_PATH, SIGNATURES


def header_functions():
    with open(os.path.join(ROOT, "include", "femgpu.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(femgpu_\w+)\s*\(", text)))


def test_library_exports_every_header_symbol():
    assert os.path.exists(LIB_PATH), "run __graft_entry__.build() first"
    lib = ctypes.CDLL(LIB_PATH)
    names = header_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    # the Python binding covers the whole header
    assert set(names) == set(SIGNATURES)


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_create_without_gpu_fails_cleanly():
    import cudafemproject_amd as F
    with pytest.raises(F.FemGpuError):
        F.Context(0)


@pytest.mark.parametrize("args", [(3, 5, 512, 256), (3, 16, 512, 256), (3, 16, 256, 256),
                                  (3, 12, 1024, 512), (2, 12, 512, 256), (2, 64, 1024, 512),
                                  (3, 40, 768, 512), (3, 24, 3072, 1024), (3, 40, 3072, 1024), (3, 40, 2048, 512),
                                  (2, 96, 3072, 1024)])
def test_plan_indexing_emulation(tmp_path, args):
    exe = tmp_path / "plan_check"
    src = os.path.join(ROOT, "tests", "cpp", "plan_check.cpp")
    plan = os.path.join(ROOT, "cudafemproject_amd", "csrc", "mesh_plan.cpp")
    subprocess.run(["g++", "-O2", "-std=c++17", "-include", "numeric",
                    "-I" + os.path.join(ROOT, "cudafemproject_amd", "csrc"), src, plan, "-o",
                    str(exe)], check=True)
    r = subprocess.run([str(exe), *map(str, args)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "max err" in r.stdout


def test_kuhn_mesh_conforming_equal_volumes():
    nodes, elems = P.box_tet(3, 2, 4, 1.5, 1.0, 2.0)
    X = nodes[elems]
    vol = np.einsum("ij,ij->i", X[:, 1] - X[:, 0], np.cross(X[:, 2] - X[:, 0], X[:, 3] - X[:, 0])) / 6
    h3 = 0.5 * 0.5 * 0.5
    assert np.all(vol > 0) and np.allclose(vol, h3 / 6)
    assert np.isclose(vol.sum(), 1.5 * 1.0 * 2.0)
    faces = np.sort(np.concatenate([elems[:, [1, 2, 3]], elems[:, [0, 2, 3]], elems[:, [0, 1, 3]],
                                    elems[:, [0, 1, 2]]]), axis=1)
    _, cnt = np.unique(faces, axis=0, return_counts=True)
    assert cnt.max() == 2                       # conforming: no face shared 3 times
    n_boundary = (cnt == 1).sum()
    assert n_boundary == 2 * 2 * (3 * 2 + 2 * 4 + 3 * 4)   # 2 triangles per boundary quad


@pytest.mark.parametrize("axis,side", [(0, 0), (0, 1), (1, 1), (2, 0), (2, 1)])
def test_box_face_tris_outward(axis, side):
    n = (2, 3, 4)
    nodes, elems = P.box_tet(*n)
    fn, fa, fi = P.box_face_tris(*n, axis, side)
    others = [n[i] for i in range(3) if i != axis]
    assert len(fn) == 2 * others[0] * others[1]
    L = 1.0
    assert np.allclose(nodes[fn][:, :, axis], side * L)
    for f, e in zip(fn, fa):
        assert set(f.tolist()) <= set(elems[e].tolist())
    nr = P.outward_normals(nodes, fn, fi)
    assert np.all(np.sign(nr[:, axis]) == (1 if side else -1))


def test_plate_mesh_ccw_and_area():
    nodes, elems = P.plate_tri(5, 3, 1.0, 0.5)
    X = nodes[elems]
    a = 0.5 * ((X[:, 1, 0] - X[:, 0, 0]) * (X[:, 2, 1] - X[:, 0, 1]) -
               (X[:, 1, 1] - X[:, 0, 1]) * (X[:, 2, 0] - X[:, 0, 0]))
    assert np.all(a > 0) and np.isclose(a.sum(), 0.5)


def test_prepared_round_trip(tmp_path):
    pb = P.cube_problem(2)
    pb.load_node = np.array([3, 7], np.int32)
    pb.load_vec = np.array([[1.0, 2.0, 3.0], [-1.0, 0.5, 0.25]])
    P.write_prepared(pb, str(tmp_path))
    q = P.read_prepared(str(tmp_path), 3, E=pb.E, nu=pb.nu)
    assert np.array_equal(q.nodes, pb.nodes) and np.array_equal(q.elems, pb.elems)
    assert np.array_equal(q.constr_node, pb.constr_node)
    assert np.array_equal(q.face_nodes, pb.face_nodes) and np.array_equal(q.face_adj, pb.face_adj)
    assert np.allclose(q.face_normal, pb.face_normal) and np.allclose(q.face_p, pb.face_p)
    assert np.array_equal(q.face_interior, pb.face_interior)
    assert np.array_equal(q.load_node, pb.load_node) and np.allclose(q.load_vec, pb.load_vec)


def test_legacy_stress_format_reads():
    from helpers import fixture_problem
    pb = fixture_problem("test_rect_pcg")
    assert pb.n_nodes == 6 and pb.n_elems == 4
    assert list(pb.face_adj) == [1, 1]
    assert np.allclose(pb.face_p, 7000.0)
    assert np.all(pb.face_interior == -1)
