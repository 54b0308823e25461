"""GPU: the assembled load vectors the solves use, against closed forms and the
oracle.

* Reference convention with nodal loads (loads.txt): the reference gathers
  each nodal load per element slot and the element-slot sum overwrites the
  pressure loads there (cpu_fem_utils.cpp:588-596, :522-530); the oracle's
  restatement of that path is bitwise equal to the reference's CPU build, so
  the GPU vector is compared with it.
* Physical convention (F = -p A/npf n per face node, constrained DOFs zero):
  the right-hand sides of the cube, column and cantilever workloads are
  checked against what the geometry fixes: the resultant -p A n, the first
  moments of a uniform pressure, and the per-node share of a face node on a
  uniformly triangulated face (h^2 for a node with six incident triangles,
  h^2/2 on an edge).
Tolerances are stated per test."""
import numpy as np
import pytest

import cudafemproject_amd as F
import oracle as O
from cudafemproject_amd.problem import (cantilever_problem, column_problem, cube_problem,
                                        plate_problem)

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a, float).ravel(), np.asarray(b, float).ravel()
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _nodal_loads(pb, seed, n=8):
    rng = np.random.default_rng(seed)
    free = np.setdiff1d(np.arange(pb.n_nodes), pb.constr_node)
    pb.load_node = np.sort(rng.choice(free, n, replace=False)).astype(np.int32)
    pb.load_vec = rng.uniform(-1e3, 1e3, (n, pb.dim))
    return pb


@pytest.mark.parametrize("name", ["plate", "cube", "plate_loads_on_face"])
def test_reference_nodal_loads_vs_oracle(ctx, name):
    """loads.txt nodal loads in the reference convention: GPU RHS equals the
    oracle's (which equals the reference build bit for bit) to 1e-14."""
    if name == "plate":
        pb = _nodal_loads(plate_problem(20, 10), 1)
    elif name == "cube":
        pb = _nodal_loads(cube_problem(8, convention=F.LOADS_REFERENCE), 2)
    else:
        # nodal loads on nodes that also carry pressure: they overwrite it
        pb = plate_problem(20, 10)
        right = np.unique(pb.face_nodes)
        pb.load_node = right[:5].astype(np.int32)
        pb.load_vec = np.random.default_rng(3).uniform(-1e3, 1e3, (5, 2))
    ctx.set_problem(pb)
    f = ctx.rhs()
    fo = O.rhs(pb, np.float64)
    assert rel(f, fo) < 1e-14
    # the loaded nodes carry a load, and it is not the pressure alone
    assert np.abs(f[pb.load_node]).max() > 0


def test_physical_nodal_loads_added_once(ctx):
    """Physical convention, nodal loads only: f equals the listed loads at
    their nodes (each added once) and zero elsewhere."""
    pb = _nodal_loads(cube_problem(8, convention=F.LOADS_PHYSICAL), 4)
    pb.face_p = np.zeros_like(pb.face_p)
    ctx.set_problem(pb)
    f = ctx.rhs()
    want = np.zeros_like(f)
    want[pb.load_node] = pb.load_vec
    assert np.array_equal(f, want)


def _top_face_checks(f, nodes, top, p, h, lx, ly, zmax):
    """Uniform pressure p on the face z = zmax (outward normal +z) of a box
    triangulated with equal diagonals (six triangles at an inner face node)."""
    assert np.all(f[:, :2] == 0.0)
    off = np.abs(nodes[:, 2] - zmax) > 1e-12
    assert np.all(f[off] == 0.0)
    fz = f[:, 2]
    area = lx * ly
    assert abs(fz.sum() + p * area) <= 1e-12 * p * area
    # first moments of a uniform load: the A/3 vertex rule is exact for
    # linear functions
    assert abs((fz * nodes[:, 0]).sum() + p * area * lx / 2) <= 1e-12 * p * area * lx
    assert abs((fz * nodes[:, 1]).sum() + p * area * ly / 2) <= 1e-12 * p * area * ly
    x, y = nodes[top, 0], nodes[top, 1]
    inner = (x > 1e-12) & (x < lx - 1e-12) & (y > 1e-12) & (y < ly - 1e-12)
    edge = ~inner & ~(((x < 1e-12) | (x > lx - 1e-12)) & ((y < 1e-12) | (y > ly - 1e-12)))
    np.testing.assert_allclose(fz[top[inner]], -p * h * h, rtol=1e-13)
    np.testing.assert_allclose(fz[top[edge]], -p * h * h / 2, rtol=1e-13)


@pytest.mark.parametrize("n", [12, 64])
def test_physical_cube_rhs_closed_form(ctx, n):
    """The cube workload's load vector (n = 64 is the benchmark's 64^3 cube,
    pressure 1e3 on z = 1): resultant -p, first moments -p/2, inner face
    nodes -p h^2, edge nodes -p h^2/2, zero off the face and in x, y."""
    pb = cube_problem(n, convention=F.LOADS_PHYSICAL)
    ctx.set_problem(pb)
    f = ctx.rhs()
    top = np.nonzero(np.abs(pb.nodes[:, 2] - 1.0) <= 1e-12)[0]
    _top_face_checks(f, pb.nodes, top, 1e3, 1.0 / n, 1.0, 1.0, 1.0)


def test_physical_column_rhs_closed_form(ctx):
    """The column workload's load vector (16 x 16 x 256, unit pressure on the
    top face z = 1.6, 0.1 x 0.1): same closed forms."""
    pb = column_problem()
    ctx.set_problem(pb)
    f = ctx.rhs()
    top = np.nonzero(np.abs(pb.nodes[:, 2] - 1.6) <= 1e-12)[0]
    _top_face_checks(f, pb.nodes, top, 1.0, 0.1 / 16, 0.1, 0.1, 1.6)


def test_physical_cantilever_rhs_closed_form(ctx):
    """The cantilever workload's tip load (x = 3 end, faces with centroid
    z > 0.875, pressure 1e3 applied along -z): resultant -p times the strip's
    area, its first moments, zero in x and y and away from the tip."""
    pb = cantilever_problem()
    ctx.set_problem(pb)
    f = ctx.rhs()
    assert np.all(f[:, :2] == 0.0)
    X = pb.nodes
    tip = np.abs(X[:, 0] - 3.0) <= 1e-12
    assert np.all(f[~tip] == 0.0)
    # strip: the four top cell rows of the end face (triangle centroids lie
    # 1/3 or 2/3 of a cell below a row's top), z in [0.875, 1], y in [0, 1]
    area, p = 0.125, 1e3
    assert abs(f[:, 2].sum() + p * area) <= 1e-12 * p * area
    assert abs((f[:, 2] * X[:, 1]).sum() + p * area * 0.5) <= 1e-12 * p * area
    assert abs((f[:, 2] * X[:, 2]).sum() + p * area * 0.9375) <= 1e-12 * p * area


def test_physical_plate_rhs_closed_form(ctx):
    """The 64 x 32 plate (traction t = 1e3 along +x on the right edge x = 1,
    0 <= y <= 0.5): resultant t * 0.5, inner edge nodes t h, end nodes t h/2,
    y components zero, the fixed left edge zero."""
    pb = plate_problem(64, 32, convention=F.LOADS_PHYSICAL)
    ctx.set_problem(pb)
    f = ctx.rhs()
    t, h = 1e3, 0.5 / 32
    assert np.all(f[:, 1] == 0.0)
    right = np.abs(pb.nodes[:, 0] - 1.0) <= 1e-12
    assert np.all(f[~right] == 0.0)
    assert abs(f[:, 0].sum() - t * 0.5) <= 1e-12 * t
    y = pb.nodes[right, 1]
    fx = f[right, 0]
    ends = (y < 1e-12) | (y > 0.5 - 1e-12)
    np.testing.assert_allclose(fx[~ends], t * h, rtol=1e-13)
    np.testing.assert_allclose(fx[ends], t * h / 2, rtol=1e-13)
