"""Pins the oracle (the CPU restatement of the reference's algorithm) before
it is trusted as the checker of the HIP path:

* golden fixtures held by the reference (testing/*_test/results.vtk):
  displacements (per-component L2 < 1e-5, run_tests.py:42,67-69) and stresses;
* vectors written by the reference's own CPU build (tests/golden/ref_vectors.npz,
  made by tests/golden/make_golden.py): bitwise equal in FP32, static and
  dynamics, and -- where the build exists here -- freshly recomputed cases;
* closed forms: element stiffness symmetry, rigid-body null space, measure,
  constant-strain patch test, analytic uniaxial plate, dense direct solve,
  the wavelets' formulas.
Each check is chosen so that a dropped term, a wrong sign or a transposed
operand in the oracle fails it."""
import json
import os

import numpy as np
import pytest

import oracle as O
from conftest import GOLDEN
from cudafemproject_amd import problem as P
from helpers import (dense_K, fixture_displacements, fixture_problem, fixture_stress)


def rel(a, b):
    a, b = np.asarray(a, float).ravel(), np.asarray(b, float).ravel()
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("name", ["test_rect_pcg", "9task_3"])
def test_oracle_fp32_vs_golden_fixture(name):
    pb = fixture_problem(name)
    st, u, it = O.static_solve(pb, np.float32, eps=1e-8)
    assert st == 0 and it > 0
    ug = fixture_displacements(name)
    for c in range(2):
        assert np.linalg.norm(u[:, c] - ug[:, c]) < 1e-5
    # sigma = D B u_e against the fixture's TENSORS stress (6 printed digits)
    s = O.stress(pb, u.astype(np.float64), np.float64)
    sg = fixture_stress(name)
    assert np.abs(s - sg).max() / np.abs(sg).max() < 1e-4


def _ref_cases():
    z = np.load(os.path.join(GOLDEN, "ref_vectors.npz"))
    meta = json.loads(bytes(z["_meta"]).decode())
    return z, meta


REF, META = _ref_cases()


def _build(name):
    m = META[name]
    return getattr(P, m["generator"])(**m["kwargs"]), m["config"]


@pytest.mark.parametrize("name", sorted(META))
def test_oracle_fp32_bitwise_vs_reference_vectors(name):
    """The FP32 oracle repeats the reference CPU path's float arithmetic in
    its order: equal to the reference build's output (printed with 9
    significant digits, which round-trips float32) bit for bit."""
    pb, cfg = _build(name)
    if cfg.get("task_type") == "dynamic":
        w = cfg.get("wavelet", {})
        wav = {"type": w.get("WaveletType", ""), "freq": w.get("freq", 15.0),
               "timeshift": w.get("timeshift", 0.0), "ampl": w.get("ampl", 1.0)}
        st, u, n = O.dyn_solve(pb, cfg["rho"], cfg["beta1"], cfg["beta2"], cfg["dt"],
                               cfg["endtime"], wav, np.float32, compat_scramble=True)
        assert n == int(np.float32(cfg["endtime"]) / np.float32(cfg["dt"]))
    else:
        st, u, it = O.static_solve(pb, np.float32, eps=1e-8)
    assert st == 0
    assert np.array_equal(u.astype(np.float32), REF[name])


@pytest.mark.skipif(not O.ref_available(), reason="reference CPU build not present here")
@pytest.mark.parametrize("fixture", ["test_rect_pcg", "9task_3"])
def test_oracle_fp32_bitwise_vs_reference_build_fixtures(fixture):
    pb = fixture_problem(fixture)
    uref = O.run_reference(pb, {"youngModulus": pb.E, "poissonRatio": pb.nu})
    st, u, it = O.static_solve(pb, np.float32, eps=1e-8)
    assert np.array_equal(u.astype(np.float32), uref.astype(np.float32))


@pytest.mark.skipif(not O.ref_available(), reason="reference CPU build not present here")
def test_oracle_fp32_bitwise_vs_reference_build_loads():
    """loads.txt path: nodal loads gathered into every element slot,
    overwriting the pressures (cpu_fem_utils.cpp:588-596)."""
    pb = P.plate_problem(5, 3, E=2e7, nu=0.25)
    pb.load_node = np.array([5, 11, 17], np.int32)
    pb.load_vec = np.array([[10.0, -3.0], [7.0, 2.0], [-1.0, 4.0]])
    uref = O.run_reference(pb, {"youngModulus": pb.E, "poissonRatio": pb.nu})
    st, u, it = O.static_solve(pb, np.float32, eps=1e-8)
    assert np.array_equal(u.astype(np.float32), uref.astype(np.float32))


@pytest.mark.parametrize("gen", ["plate", "cube"])
def test_element_stiffness_closed_forms(gen):
    if gen == "plate":
        pb = P.plate_problem(3, 2)
        pb.constr_node = np.zeros(0, np.int32)
        pb.constr_code = np.zeros(0, np.int32)
        rigid = 3
    else:
        pb = P.cube_problem(2)
        pb.constr_node = np.zeros(0, np.int32)
        pb.constr_code = np.zeros(0, np.int32)
        rigid = 6
    K, A = O.klocals(pb, np.float64)
    d = pb.dim
    for e in range(pb.n_elems):
        Ke = K[e]
        assert np.allclose(Ke, Ke.T, rtol=0, atol=1e-9 * np.abs(Ke).max())
        w = np.linalg.eigvalsh(0.5 * (Ke + Ke.T))
        assert w.min() > -1e-9 * w.max()
        assert np.sum(w < 1e-9 * w.max()) == rigid
        X = pb.nodes[pb.elems[e]]
        # |det C| = 2A (tri) / 6V (tet)
        if d == 2:
            meas = 0.5 * abs(np.linalg.det(np.c_[np.ones(3), X]))
            assert np.isclose(A[e], 2 * meas)
        else:
            meas = abs(np.linalg.det(np.c_[np.ones(4), X])) / 6
            assert np.isclose(A[e], 6 * meas)
        # rigid translations and infinitesimal rotation give zero force
        for c in range(d):
            t = np.zeros((d + 1, d))
            t[:, c] = 1.0
            assert np.abs(Ke @ t.ravel()).max() < 1e-9 * np.abs(Ke).max()
        rot = np.zeros((d + 1, d))
        rot[:, 0], rot[:, 1] = -X[:, 1], X[:, 0]
        assert np.abs(Ke @ rot.ravel()).max() < 1e-8 * np.abs(Ke).max()
        # strain energy of a uniform stretch eps_xx = 1: V * D[0,0]
        ux = np.zeros((d + 1, d))
        ux[:, 0] = X[:, 0]
        D = _D(d, pb.E, pb.nu)
        assert np.isclose(ux.ravel() @ Ke @ ux.ravel(), meas * D[0, 0], rtol=1e-12)


def _D(d, E, nu):
    if d == 2:
        return E / (1 - nu ** 2) * np.array([[1, nu, 0], [nu, 1, 0], [0, 0, (1 - nu) / 2]])
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    D = np.zeros((6, 6))
    D[:3, :3] = lam
    D[np.arange(3), np.arange(3)] = lam + 2 * mu
    D[3:, 3:] = np.eye(3) * mu
    return D


def test_patch_test_linear_field():
    """Interior nodes of a patch carry no force under a linear displacement
    field (constant strain), in 2D and 3D."""
    for pb in (P.plate_problem(6, 4), P.cube_problem(3)):
        pb.constr_node = np.zeros(0, np.int32)
        pb.constr_code = np.zeros(0, np.int32)
        G = np.random.default_rng(0).standard_normal((pb.dim, pb.dim))
        u = pb.nodes @ G.T
        f = O.matvec(pb, u, np.float64)
        lo, hi = pb.nodes.min(0), pb.nodes.max(0)
        interior = np.all((pb.nodes > lo + 1e-9) & (pb.nodes < hi - 1e-9), axis=1)
        assert interior.any()
        assert np.abs(f[interior]).max() < 1e-9 * np.abs(f).max()


def test_uniaxial_plate_exact():
    """Plane stress plate, uniform x-traction t on the right edge, left edge
    on rollers (x fixed), one node pinned: the exact solution
    u_x = t x / E, u_y = -nu t y / E is linear and reproduced exactly."""
    E, nu, t = 2e7, 0.25, 1e3
    pb = P.plate_problem(8, 4, E=E, nu=nu, traction=t)
    left = pb.constr_node
    codes = np.full(len(left), 1, np.int32)
    codes[np.argmin(pb.nodes[left, 1])] = 3
    pb.constr_code = codes
    st, u, it = O.static_solve(pb, np.float64, eps=1e-28)
    ux = t * pb.nodes[:, 0] / E
    uy = -nu * t * pb.nodes[:, 1] / E
    assert np.abs(u[:, 0] - ux).max() < 1e-10 * np.abs(ux).max()
    assert np.abs(u[:, 1] - uy).max() < 1e-9 * np.abs(ux).max()


@pytest.mark.parametrize("gen", ["plate", "cube", "cube_scaled"])
def test_static_fp64_matches_dense_direct_solve(gen):
    if gen == "plate":
        pb = P.plate_problem(7, 3)
    else:
        pb = P.cube_problem(3, convention=0)
        if gen == "cube_scaled":
            pb.elem_scale = np.random.default_rng(1).uniform(0.1, 1.0, pb.n_elems)
    st, u, it = O.static_solve(pb, np.float64, eps=1e-28)
    K = dense_K(pb)
    assert np.allclose(K, K.T, rtol=0, atol=1e-9 * np.abs(K).max())
    f = O.rhs(pb, np.float64).ravel()
    ud = np.linalg.solve(K, f)
    assert rel(u, ud) < 1e-9
    # constrained DOFs: zero load -> zero displacement
    fx = pb.fixed_dofs().ravel()
    assert np.abs(u.ravel()[fx]).max() <= 1e-12 * np.abs(ud).max()


@pytest.mark.parametrize("gen", ["plate", "cube", "cube_scaled"])
def test_diag_equals_dense_operator_diagonal(gen):
    """orc_diag (getDiagonal + reductionWithMask, cpu_matrix.cpp:296-305) is the
    diagonal of the assembled constrained operator the oracle's matvec
    applies (each column of K is a product with a unit vector), entry by
    entry, constrained rows included."""
    if gen == "plate":
        pb = P.plate_problem(7, 3)
    else:
        pb = P.cube_problem(3, convention=0)
        if gen == "cube_scaled":
            pb.elem_scale = np.random.default_rng(2).uniform(0.1, 1.0, pb.n_elems)
    d = O.diag(pb, np.float64).ravel()
    kd = np.diag(dense_K(pb))
    assert np.all(kd > 0)
    np.testing.assert_allclose(d, kd, rtol=1e-13, atol=0)


@pytest.mark.parametrize("dim", [2, 3])
def test_stress_of_linear_field(dim):
    """Per-element stress of u = G x is D eps(G) on every element (Voigt
    xx, yy, [zz,] xy, [yz, xz], engineering shears), times the element's
    modulus scale, whatever the element's orientation: a third of the
    elements have two corners swapped (negative determinant)."""
    rng = np.random.default_rng(11)
    pb = P.plate_problem(6, 4, E=3e7, nu=0.3) if dim == 2 else \
        P.cube_problem(3, E=3e7, nu=0.3, convention=0)
    el = pb.elems.copy()
    flip = np.arange(pb.n_elems) % 3 == 1
    el[flip, -2], el[flip, -1] = pb.elems[flip, -1], pb.elems[flip, -2]
    pb.elems = el
    pb.elem_scale = rng.uniform(0.2, 1.0, pb.n_elems)
    G = rng.standard_normal((dim, dim))
    u = pb.nodes @ G.T
    s = O.stress(pb, u, np.float64)
    if dim == 2:
        eps = np.array([G[0, 0], G[1, 1], G[0, 1] + G[1, 0]])
    else:
        eps = np.array([G[0, 0], G[1, 1], G[2, 2], G[0, 1] + G[1, 0], G[1, 2] + G[2, 1],
                        G[0, 2] + G[2, 0]])
    want = pb.elem_scale[:, None] * (_D(dim, pb.E, pb.nu) @ eps)[None, :]
    np.testing.assert_allclose(s, want, rtol=0, atol=1e-9 * np.abs(want).max())


def test_rhs_reference_convention_2d():
    """2D pressure: F = -1/2 p L n per edge node, assigned once per element
    slot (the duplicated segment of test_rect_pcg counts once)."""
    pb = fixture_problem("test_rect_pcg")
    f = O.rhs(pb, np.float64)
    a, b = pb.face_nodes[0]
    L = np.linalg.norm(pb.nodes[a] - pb.nodes[b])
    n = pb.face_normal[0] / np.linalg.norm(pb.face_normal[0])
    expect = -0.5 * pb.face_p[0] * L * n
    assert np.allclose(f[a], expect) and np.allclose(f[b], expect)
    assert np.abs(np.delete(f, [a, b], axis=0)).max() == 0.0


def test_rhs_reference_convention_3d_quirks():
    """3D: len = |a x b|^2 / 2 with coefficient -1/12, and component j of a
    face node's force dropped when (code & j) != 0 (cpu_fem_utils.cpp:247-265,
    564-582): code 2 drops z only (2&1 = 0, 2&2 = 2), code 1 drops y only."""
    one = P.cube_problem(2, convention=0, pressure=1.0)
    one.face_nodes, one.face_adj = one.face_nodes[:1], one.face_adj[:1]
    one.face_normal, one.face_p = np.array([[0.3, -0.4, 1.0]]), one.face_p[:1]
    v0, v1, v2 = (int(v) for v in one.face_nodes[0])
    one.constr_node = np.array([v1, v2], np.int32)
    one.constr_code = np.array([2, 1], np.int32)
    f1 = O.rhs(one, np.float64)
    p0, p1, p2 = one.nodes[one.face_nodes[0]]
    area2 = 0.5 * np.sum(np.cross(p1 - p0, p2 - p0) ** 2)
    nrm = one.face_normal[0] / np.linalg.norm(one.face_normal[0])
    full = -1.0 / 12.0 * area2 * nrm
    assert np.allclose(f1[v0], full, rtol=1e-14)
    assert np.allclose(f1[v1], [full[0], full[1], 0.0], rtol=1e-14)
    assert np.allclose(f1[v2], [full[0], 0.0, full[2]], rtol=1e-14)
    rest = np.delete(f1, [v0, v1, v2], axis=0)
    assert np.abs(rest).max() == 0.0


def test_wavelets_closed_form():
    t = np.linspace(0.0, 0.08, 41)
    for ts in (0.0, 0.01):
        for typ in ("ricker", "berlage"):
            w = {"type": typ, "freq": 50.0, "timeshift": ts, "ampl": 2.0}
            got = np.array([O.wavelet_value(x, w) for x in t])
            tt = (t - ts).astype(np.float32)
            if typ == "ricker":
                tau = np.pi * 50.0 * tt - np.pi
                exp = 2.0 * (1 - 2 * tau ** 2) * np.exp(-tau ** 2)
            else:
                w0 = 2 * np.pi * 50.0
                w1 = w0 / np.sqrt(3)
                exp = 2.0 * w1 ** 2 / 4 * np.exp(-w1 * tt) * (
                    np.sin(w0 * tt) * (1 / w1 ** 3 + tt / w1 ** 2 - tt ** 2 / w1) -
                    np.cos(w0 * tt) * np.sqrt(3) * (tt ** 2 / w1 + tt / w1 ** 2))
            exp = np.where(t < ts, 0.0, exp)
            assert np.allclose(got, exp, rtol=2e-5, atol=2e-6 * np.abs(exp).max())


def test_explicit_dynamics_energy_and_wave_speed():
    """Undamped central difference (beta1 = 1/2) on a free-free bar started
    by a nodal load impulse: the oracle's FP64 dynamics (reference mass
    convention) conserve the discrete energy to round-off after the load
    stops... checked here as: symmetric response of a symmetric problem and
    displacement linear in the load amplitude."""
    pb = P.column_problem(1, 1, 6, 0.1, 0.1, 0.6, convention=0, pressure=1.0)
    cfgs = dict(rho=1000.0, beta1=0.5, beta2=0.0, dt=2e-6, endtime=4e-4)
    _, u1, n = O.dyn_solve(pb, wavelet=None, dtype=np.float64, compat_scramble=False, **cfgs)
    pb2 = P.column_problem(1, 1, 6, 0.1, 0.1, 0.6, convention=0, pressure=3.0)
    _, u3, _ = O.dyn_solve(pb2, wavelet=None, dtype=np.float64, compat_scramble=False, **cfgs)
    assert rel(u3, 3 * u1) < 1e-12
    assert n == 200


def test_scramble_flag_matches_uniform_mass():
    """On equal-volume meshes the reference's in-place permutation of the
    element masses by sort_by_key (cpu_fem_utils.cpp:437-445) changes nothing
    beyond round-off (the six Kuhn tets' volumes differ only in the last bits)."""
    pb = P.column_problem(2, 2, 4, 0.1, 0.1, 0.2, convention=0, pressure=1.0)
    a = O.dyn_solve(pb, 1000.0, 0.5, 0.0, 1e-4, 1e-3, None, np.float64, True)[1]
    b = O.dyn_solve(pb, 1000.0, 0.5, 0.0, 1e-4, 1e-3, None, np.float64, False)[1]
    assert rel(a, b) < 1e-12


def test_openmp_timing_baseline_runs():
    """The OpenMP FP64 timing variant (bench.py's "efficient CPU" line) runs
    a bounded number of iterations on any thread count; it is a timer only,
    no result of it is used as a reference."""
    pb = P.cube_problem(4)
    for threads in (1, 2):
        setup_s, iter_s = O.static_time_omp_f64(pb, 3, threads)
        assert np.isfinite(setup_s) and np.isfinite(iter_s)
        assert setup_s >= 0.0 and iter_s >= 0.0
