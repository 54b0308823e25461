"""GPU: the whole run_fem.py -> fem_main -> results.vtk flow on the golden
fixtures, and the argv form run_fem.py of the reference builds."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from conftest import GOLDEN, ROOT
from helpers import fixture_displacements, fixture_problem, fixture_stress, read_vtk_field
from test_femproject_cpu import workspace

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("fixture", ["test_rect_pcg", "9task_3"])
def test_run_fem_static_fixture(tmp_path, fixture, fem_main):
    cfg = workspace(tmp_path, fixture, youngModulus=2e7, poissonRatio=0.25)
    r = subprocess.run(["python", os.path.join(ROOT, "run_fem.py"), cfg], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    vtk = tmp_path / "fem_results" / "2D" / fixture / "results.vtk"
    u = read_vtk_field(str(vtk), "displacements").reshape(-1, 3)[:, :2]
    ug = fixture_displacements(fixture)
    for c in range(2):
        assert np.linalg.norm(u[:, c] - ug[:, c]) < 1e-5     # run_tests.py:42,67-69
    s = read_vtk_field(str(vtk), "stress").reshape(-1, 3, 3)
    sg = fixture_stress(fixture)
    got = np.stack([s[:, 0, 0], s[:, 1, 1], s[:, 0, 1]], 1)
    assert np.abs(got - sg).max() / np.abs(sg).max() < 1e-4
    # every other field of the fixture (legacy results layout, 6 printed
    # digits there): per-element strain, its xy component, the von Mises
    # values, the xy stress, and SmoothStress (zeros: the reference's main
    # runs without smoothing)
    fx = os.path.join(GOLDEN, fixture, "results.vtk")
    for name in ("deformation", "XY_DEFORM", "XY_STRESS", "mises_stress", "mises_deformation"):
        a, b = read_vtk_field(str(vtk), name), read_vtk_field(fx, name)
        assert a.shape == b.shape, name
        assert np.abs(a - b).max() / np.abs(b).max() < 1e-4, name
    assert np.array_equal(read_vtk_field(str(vtk), "SmoothStress"),
                          read_vtk_field(fx, "SmoothStress"))


def test_fem_main_argv_form_dynamic(tmp_path, fem_main):
    """run_fem.py:80-84 argv (task dim proj prep res nu E verbose 1 rho da db dt
    endtime beta1 beta2): explicit dynamics on the test_rect_pcg mesh vs the
    FP64 oracle."""
    cfg = workspace(tmp_path, "test_rect_pcg")
    prep = tmp_path / "prepared_meshes" / "2D" / "test_rect_pcg"
    res = tmp_path / "res"
    res.mkdir()
    exe = fem_main
    r = subprocess.run([exe, "test_rect_pcg", "2", ROOT, str(prep), str(res), "0.25", "2e7", "0",
                        "1", "2400", "0", "0", "1e-6", "2e-4", "0.5", "0"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    u = read_vtk_field(str(res / "results.vtk"), "displacements").reshape(-1, 3)[:, :2]
    pb = fixture_problem("test_rect_pcg")
    _, uo, n = O.dyn_solve(pb, 2400.0, 0.5, 0.0, 1e-6, 2e-4, None, np.float64,
                           compat_scramble=False)
    assert np.linalg.norm(u - uo) / np.linalg.norm(uo) < 1e-7   # 9 printed digits


def test_run_fem_from_raw_k_mesh(tmp_path, fem_main):
    """The full reference flow from the raw LS-DYNA mesh: run_fem.py parses
    raw_meshes/2D/9task_3.k, solves, and results.vtk matches the golden
    results the reference's tests hold (run_tests.py:42,67-69 criterion)."""
    import json
    import shutil
    from conftest import GOLDEN
    raw = tmp_path / "proj" / "raw_meshes" / "2D"
    raw.mkdir(parents=True)
    shutil.copy(os.path.join(GOLDEN, "9task_3", "9task_3.k"), raw / "9task_3.k")
    p = tmp_path / "c.json"
    p.write_text(json.dumps({"DIM": 2, "working_path": str(tmp_path / "work") + "/",
                             "CudaFEMproject_path": str(tmp_path / "proj") + "/",
                             "task": ["9task_3"], "youngModulus": 2e7, "poissonRatio": 0.25}))
    r = subprocess.run(["python", os.path.join(ROOT, "run_fem.py"), str(p)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    vtk = tmp_path / "work" / "fem_results" / "2D" / "9task_3" / "results.vtk"
    u = read_vtk_field(str(vtk), "displacements").reshape(-1, 3)[:, :2]
    ug = fixture_displacements("9task_3")
    for c in range(2):
        assert np.linalg.norm(u[:, c] - ug[:, c]) < 1e-5


def test_smooth_stress_projection(tmp_path, fem_main):
    """--smooth: SmoothStress is the L2 projection of sigma_xx onto the nodal
    basis.  Uniaxial plate (rollers on the left edge, one pinned node,
    traction t on the right): sigma_xx = t in every element, so the
    projection is t at every node; and on the 9task_3 fixture the projection
    keeps the integral of sigma_xx (the nodal basis sums to one)."""
    from cudafemproject_amd.problem import plate_problem, write_prepared
    t = 1e3
    pb = plate_problem(8, 4, E=2e7, nu=0.25, traction=t)
    codes = np.full(len(pb.constr_node), 1, np.int32)
    codes[np.argmin(pb.nodes[pb.constr_node, 1])] = 3
    pb.constr_code = codes
    write_prepared(pb, str(tmp_path / "prepared_meshes" / "2D" / "uni"))
    import json
    cfg = tmp_path / "c.json"
    cfg.write_text(json.dumps({"working_path": str(tmp_path) + "/", "task": ["uni"], "DIM": 2,
                               "youngModulus": 2e7, "poissonRatio": 0.25}))
    r = subprocess.run([fem_main, str(cfg), "uni", "--smooth", "--tol", "1e-26"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    vtk = str(tmp_path / "fem_results" / "2D" / "uni" / "results.vtk")
    np.testing.assert_allclose(read_vtk_field(vtk, "SmoothStress"), t, rtol=1e-6)
    # the reference convention's loads give sigma_xx = traction here
    np.testing.assert_allclose(read_vtk_field(vtk, "stress").reshape(-1, 3, 3)[:, 0, 0], t,
                               rtol=1e-6)

    cfg2 = workspace(tmp_path / "f", "9task_3", youngModulus=2e7, poissonRatio=0.25)
    r = subprocess.run([fem_main, cfg2, "9task_3", "--smooth"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    vtk = str(tmp_path / "f" / "fem_results" / "2D" / "9task_3" / "results.vtk")
    S = read_vtk_field(vtk, "SmoothStress")
    sxx = read_vtk_field(vtk, "stress").reshape(-1, 3, 3)[:, 0, 0]
    pbf = fixture_problem("9task_3")
    X = pbf.nodes
    A = np.array([0.5 * abs(np.linalg.det(np.c_[np.ones(3), X[e]])) for e in pbf.elems])
    lhs = (A * S[pbf.elems].mean(1)).sum()          # integral of the nodal field
    rhs = (A * sxx).sum()                           # integral of sigma_xx
    assert abs(lhs - rhs) <= 1e-6 * np.abs(A * sxx).sum()
