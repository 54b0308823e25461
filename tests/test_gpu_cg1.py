"""GPU: the single-reduction PCG (FEMGPU_CG_SINGLE_REDUCTION, cg1.hip) against
the standard recurrence, the oracle and the dense solve.  Both recurrences
are the same Jacobi-PCG (equal iterates in exact arithmetic), so at a tight
tolerance they give the same solution and at the reference's rule they take
the same number of iterations up to rounding.  Small patches (set_tuning)
put most nodes on patch interfaces, where every touching patch recomputes
the node's update."""
import numpy as np
import pytest

import cudafemproject_amd as F
import oracle as O
from cudafemproject_amd.problem import cantilever_problem, cube_problem, plate_problem
from helpers import dense_K, fixture_problem

pytestmark = pytest.mark.gpu

SINGLE = dict(recurrence=F.CG_SINGLE_REDUCTION)


def rel(a, b):
    a, b = np.asarray(a, float).ravel(), np.asarray(b, float).ravel()
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _problems():
    p3s = cube_problem(4, convention=F.LOADS_REFERENCE)
    p3s.elem_scale = np.random.default_rng(3).uniform(0.05, 1.0, p3s.n_elems) ** 3
    return {"plate": plate_problem(12, 7, E=2e7, nu=0.25),
            "cube": cube_problem(5, convention=F.LOADS_REFERENCE),
            "cube_scaled": p3s,
            "rect_pcg": fixture_problem("test_rect_pcg")}


PROBLEMS = _problems()
# (elements per patch, threads per patch): one patch, and many small patches
TUNINGS = [(0, 0), (48, 256), (200, 512)]


@pytest.fixture
def tuned(ctx):
    yield ctx
    ctx.set_tuning(0, 0)


@pytest.mark.parametrize("tuning", TUNINGS)
@pytest.mark.parametrize("name", list(PROBLEMS))
def test_single_reduction_converged_vs_dense(tuned, name, tuning):
    """Tight FP64 solve equals the dense direct solve of the oracle's operator
    (same bound as the standard recurrence's test)."""
    pb = PROBLEMS[name]
    print(f"[cg1 dense] {name} tuning={tuning}", flush=True)
    tuned.set_tuning(*tuning)
    tuned.set_problem(pb)
    u, st = tuned.solve(F.FP64, tol=1e-13, tol_kind=F.TOL_REL_RES, **SINGLE)
    K = dense_K(pb)
    uo = np.linalg.solve(K, O.rhs(pb, np.float64).ravel())
    assert st["status"] == 0
    assert rel(u, uo) < 1e-10
    assert st["rel_res"] < 1e-12


@pytest.mark.parametrize("tuning", TUNINGS)
@pytest.mark.parametrize("name", list(PROBLEMS))
def test_single_reduction_reference_rule(tuned, name, tuning):
    """gamma < 1e-8 gamma0 (femfunc.cpp:20,92): iteration count of the
    oracle's standard PCG within two, result as close to the oracle as the
    oracle is to the converged solution (the bound of the standard
    recurrence's test)."""
    pb = PROBLEMS[name]
    tuned.set_tuning(*tuning)
    tuned.set_problem(pb)
    u, st = tuned.solve(F.FP64, tol=1e-8, tol_kind=F.TOL_REF_GAMMA, **SINGLE)
    s, uo, it = O.static_solve(pb, np.float64, eps=1e-8)
    assert s == 0 and st["status"] == 0
    assert abs(st["iterations"] - it) <= 2
    _, uc, _ = O.static_solve(pb, np.float64, eps=1e-26)
    assert rel(u, uo) <= max(1e-10, 1.0 * rel(uo, uc))
    assert st["gamma_ratio"] < 1e-8


@pytest.mark.parametrize("tuning", TUNINGS)
@pytest.mark.parametrize("name", ["plate", "cube", "cube_scaled"])
def test_single_reduction_mixed_and_fp32(tuned, name, tuning):
    pb = PROBLEMS[name]
    tuned.set_tuning(*tuning)
    tuned.set_problem(pb)
    u64, s64 = tuned.solve(F.FP64, tol=1e-10, tol_kind=F.TOL_REL_RES)
    um, sm = tuned.solve(F.MIXED, tol=1e-10, tol_kind=F.TOL_REL_RES, **SINGLE)
    print(f"[cg1 mixed] {name} {tuning}: it {sm['iterations']} (fp64 std {s64['iterations']}), "
          f"outer {sm['outer_iterations']}, rel_res {sm['rel_res']:.2e}", flush=True)
    assert sm["status"] == 0
    assert sm["rel_res"] <= 1e-10 * 1.0001
    assert rel(um, u64) < 1e-7
    u32, s32 = tuned.solve(F.FP32, tol=1e-5, tol_kind=F.TOL_REL_RES, **SINGLE)
    assert s32["status"] == 0
    assert rel(u32, u64) < 1e-3


def test_single_reduction_iterations_match_standard(ctx):
    """Same iterates in exact arithmetic: at a tight rule both recurrences
    stop within a few iterations of each other and agree to 1e-9."""
    pb = cube_problem(8, convention=F.LOADS_PHYSICAL)
    ctx.set_problem(pb)
    ua, sa = ctx.solve(F.FP64, tol=1e-11, tol_kind=F.TOL_REL_RES)
    ub, sb = ctx.solve(F.FP64, tol=1e-11, tol_kind=F.TOL_REL_RES, **SINGLE)
    print(f"iterations standard {sa['iterations']} single-reduction {sb['iterations']}", flush=True)
    assert abs(sa["iterations"] - sb["iterations"]) <= 3
    assert rel(ub, ua) < 1e-9


def test_single_reduction_deterministic(ctx):
    pb = PROBLEMS["cube_scaled"]
    ctx.set_problem(pb)
    a, _ = ctx.solve(F.FP64, tol=1e-8, **SINGLE)
    b, _ = ctx.solve(F.FP64, tol=1e-8, **SINGLE)
    assert np.array_equal(a, b)
    c, _ = ctx.solve(F.MIXED, tol=1e-9, tol_kind=F.TOL_REL_RES, **SINGLE)
    d, _ = ctx.solve(F.MIXED, tol=1e-9, tol_kind=F.TOL_REL_RES, **SINGLE)
    assert np.array_equal(c, d)


def test_single_reduction_zero_rhs_and_max_iter(ctx):
    pb = plate_problem(6, 3)
    pb.face_p = np.zeros_like(pb.face_p)
    ctx.set_problem(pb)
    u, st = ctx.solve(F.FP64, tol=1e-8, **SINGLE)
    assert np.all(u == 0) and st["iterations"] == 0 and st["status"] == 0
    ctx.set_problem(PROBLEMS["cube"])
    u, st = ctx.solve(F.FP64, tol=1e-30, tol_kind=F.TOL_REL_RES, max_iter=7,
                      raise_on_error=False, **SINGLE)
    assert st["status"] != 0 and st["iterations"] == 7
    u2, st2 = ctx.solve(F.FP64, tol=1e-30, tol_kind=F.TOL_REL_RES, max_iter=7,
                        raise_on_error=False)
    assert rel(u, u2) < 1e-8          # 7 iterations of the same method


def test_single_reduction_autograd_and_compliance(ctx):
    pb = cube_problem(4, convention=F.LOADS_PHYSICAL)
    pb.elem_scale = np.random.default_rng(5).uniform(0.1, 1.0, pb.n_elems)
    ctx.set_problem(pb)
    c1, g1, u1, _ = ctx.compliance_grad(F.FP64, tol=1e-12)
    c2, g2, u2, _ = ctx.compliance_grad(F.FP64, tol=1e-12, **SINGLE)
    assert abs(c2 - c1) <= 1e-10 * abs(c1)
    assert rel(g2, g1) < 1e-9


def test_single_reduction_mixed_restarts_after_breakdown(ctx):
    """Cantilever 96x32x32 (589,824 tets, E_e = E0 rho_e^3): the first
    replacement that keeps p is followed by a breakdown of the FP32
    single-reduction recurrence; the solve restarts from the FP64 residual
    instead of failing."""
    ctx.set_problem(cantilever_problem())
    u64, s64 = ctx.solve(F.FP64, tol=1e-8, tol_kind=F.TOL_REL_RES)
    for rec in (F.CG_SINGLE_REDUCTION, F.CG_STANDARD):
        um, sm = ctx.solve(F.MIXED, tol=1e-8, tol_kind=F.TOL_REL_RES, recurrence=rec)
        print(f"[mixed cantilever] recurrence {rec}: it {sm['iterations']} (fp64 "
              f"{s64['iterations']}), outer {sm['outer_iterations']}, rel_res {sm['rel_res']:.2e}",
              flush=True)
        assert sm["status"] == 0
        assert sm["rel_res"] <= 1e-8 * 1.0001
        assert rel(um, u64) < 1e-6
