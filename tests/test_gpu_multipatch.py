"""GPU parity on meshes that plan to many patches: the interface path.

Every mesh here is cut into several patches at the tuning it runs with, so a
large share of its nodes are interface nodes: their sums go through the
per-patch partial-sum slots (the matvec's gather launch, the PCG update's
interface branch, the first-slot constrained rows, the explicit stepper's
late correction of interface nodes).  Each case's patch count and interface
node count are part of its test id (``-np<patches>-if<interface nodes>``).

The oracle is the FP64 restatement of the reference's CPU algorithm
(femfunc.cpp:25-128 for the PCG, :130-381 for the GN22 steppers,
cpu_fem_utils.cpp:338-596 for the element operators and loads).
Tolerances are stated per test."""
import numpy as np
import pytest

import cudafemproject_amd as F
import oracle as O
from cudafemproject_amd.problem import column_problem, cube_problem, plate_problem
from helpers import dense_K, fixture_problem

pytestmark = pytest.mark.gpu

REF = F.LOADS_REFERENCE


def rel(a, b):
    a, b = np.asarray(a, float).ravel(), np.asarray(b, float).ravel()
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _scaled(pb, seed):
    pb.elem_scale = np.random.default_rng(seed).uniform(0.05, 1.0, pb.n_elems) ** 3
    return pb


def _with_nodal_loads(pb, seed):
    """Adds nodal loads (loads.txt lines) on a few free nodes."""
    rng = np.random.default_rng(seed)
    free = np.setdiff1d(np.arange(pb.n_nodes), pb.constr_node)
    pb.load_node = np.sort(rng.choice(free, 6, replace=False)).astype(np.int32)
    pb.load_vec = rng.uniform(-1e3, 1e3, (6, pb.dim))
    return pb


# (name, problem factory, (elements per patch, threads per patch)); (0, 0) is
# the default tuning (768 elements, 256 threads)
CASES = [
    ("cube12", lambda: cube_problem(12, convention=REF), (0, 0)),
    ("cube16", lambda: cube_problem(16, convention=REF), (0, 0)),
    ("column4x4x32", lambda: column_problem(4, 4, 32, 0.1, 0.1, 0.8, pressure=1e3,
                                            convention=REF), (0, 0)),
    ("plate64x32", lambda: plate_problem(64, 32, E=2e7, nu=0.25), (0, 0)),
    ("cube5_t48", lambda: cube_problem(5, convention=REF), (48, 256)),
    ("cube12_scaled_t200", lambda: _scaled(cube_problem(12, convention=REF), 4), (200, 512)),
    ("plate12x7_t48", lambda: plate_problem(12, 7, E=2e7, nu=0.25), (48, 256)),
    ("cube8_loads_t48", lambda: _with_nodal_loads(cube_problem(8, convention=REF), 5), (48, 256)),
    ("9task_3_t16", lambda: fixture_problem("9task_3"), (16, 256)),
]

_BUILT = {}


def _problem(name):
    if name not in _BUILT:
        fac, tuning = {n: (f, t) for n, f, t in CASES}[name]
        _BUILT[name] = (fac(), tuning)
    return _BUILT[name]


def _case_id(name):
    pb, tuning = _problem(name)
    try:
        info = F.plan_info(pb.nodes, pb.elems, *tuning)
        return f"{name}-np{info['patches']}-if{info['interface']}"
    except Exception:                      # library not built: ids only
        return name


NAMES = [c[0] for c in CASES]
IDS = [_case_id(n) for n in NAMES]


@pytest.fixture
def tuned(ctx):
    yield ctx
    ctx.set_tuning(0, 0)


def _setup(ctx, name):
    pb, tuning = _problem(name)
    info = F.plan_info(pb.nodes, pb.elems, *tuning)
    print(f"[multipatch] {name} tuning={tuning} plan={info}", flush=True)
    assert info["patches"] > 1 and info["interface"] > 0
    ctx.set_tuning(*tuning)
    ctx.set_problem(pb)
    return pb


def test_cases_are_multi_patch():
    for n in NAMES:
        pb, tuning = _problem(n)
        info = F.plan_info(pb.nodes, pb.elems, *tuning)
        assert info["patches"] > 1 and info["interface"] > 0, n


@pytest.mark.parametrize("name", NAMES, ids=IDS)
def test_matvec_vs_oracle(tuned, name):
    """FP64 EbE matvec <= 1e-12 (summation order only), FP32 <= 2e-6."""
    pb = _setup(tuned, name)
    x = np.random.default_rng(0).standard_normal((pb.n_nodes, pb.dim))
    yo = O.matvec(pb, x, np.float64)
    assert rel(tuned.matvec(x, F.FP64), yo) < 1e-12
    assert rel(tuned.matvec(x, F.FP32), yo) < 2e-6


@pytest.mark.parametrize("name", NAMES, ids=IDS)
def test_diag_and_rhs_vs_oracle(tuned, name):
    """Assembled diagonal <= 1e-13 and reference-convention RHS (pressures and
    nodal loads, cpu_fem_utils.cpp:564-596) <= 1e-14."""
    pb = _setup(tuned, name)
    assert rel(tuned.diag(), O.diag(pb, np.float64)) < 1e-13
    assert rel(tuned.rhs(), O.rhs(pb, np.float64)) < 1e-14


@pytest.mark.parametrize("name", NAMES, ids=IDS)
def test_static_fp64_reference_rule(tuned, name):
    """gamma < 1e-8 gamma0 (femfunc.cpp:20,92): iteration count within one of
    the oracle's, and the result as close to the oracle as the oracle is to
    the converged solution."""
    pb = _setup(tuned, name)
    u, st = tuned.solve(F.FP64, tol=1e-8, tol_kind=F.TOL_REF_GAMMA)
    s, uo, it = O.static_solve(pb, np.float64, eps=1e-8)
    assert s == 0 and st["status"] == 0
    assert abs(st["iterations"] - it) <= 1
    _, uc, _ = O.static_solve(pb, np.float64, eps=1e-26)
    assert rel(u, uo) <= max(1e-10, rel(uo, uc))


@pytest.mark.parametrize("name", NAMES, ids=IDS)
def test_static_converged_vs_oracle(tuned, name):
    """Reference rule tightened to gamma < 1e-24 gamma0 on both sides:
    displacements <= 1e-6 relative (FP64 mode), stresses <= 1e-6; both
    recurrences and the mixed solve (rel-res 1e-10) against the same oracle
    solution."""
    pb = _setup(tuned, name)
    s, uo, _ = O.static_solve(pb, np.float64, eps=1e-24)
    assert s == 0
    u, st = tuned.solve(F.FP64, tol=1e-24, tol_kind=F.TOL_REF_GAMMA)
    assert st["status"] == 0 and rel(u, uo) < 1e-6
    assert rel(tuned.stress(u), O.stress(pb, uo, np.float64)) < 1e-6
    u1, s1 = tuned.solve(F.FP64, tol=1e-24, tol_kind=F.TOL_REF_GAMMA,
                         recurrence=F.CG_SINGLE_REDUCTION)
    assert s1["status"] == 0 and rel(u1, uo) < 1e-6
    for rec in (F.CG_STANDARD, F.CG_SINGLE_REDUCTION):
        um, sm = tuned.solve(F.MIXED, tol=1e-10, tol_kind=F.TOL_REL_RES, recurrence=rec)
        assert sm["status"] == 0 and sm["rel_res"] <= 1e-10 * 1.0001
        assert rel(um, uo) < 1e-7, rec


@pytest.mark.parametrize("name", ["plate64x32"], ids=[IDS[NAMES.index("plate64x32")]])
def test_plate_cfg1_vs_dense(tuned, name):
    """The 64 x 32 plate itself (4,096 triangles, 4,290 DOF): tight FP64 solve
    equals the dense direct solve of the oracle's operator (1e-10)."""
    pb = _setup(tuned, name)
    u, st = tuned.solve(F.FP64, tol=1e-13, tol_kind=F.TOL_REL_RES)
    uo = np.linalg.solve(dense_K(pb), O.rhs(pb, np.float64).ravel())
    assert st["status"] == 0 and rel(u, uo) < 1e-10


DYN_NAMES = ["cube12", "column4x4x32", "plate64x32", "cube5_t48", "plate12x7_t48"]
DYN_IDS = [IDS[NAMES.index(n)] for n in DYN_NAMES]
RICKER = {"type": "ricker", "freq": 200.0, "timeshift": 0.0, "ampl": 1000.0}


def _dyn_params(pb):
    # dt well inside the stability limit of each mesh (checked: halving dt
    # does not change the oracle's trajectory's size)
    return (1e-5, 2e-3) if pb.dim == 3 else (1e-6, 2e-4)


@pytest.mark.parametrize("name", DYN_NAMES, ids=DYN_IDS)
def test_explicit_one_and_two_launch_vs_oracle(tuned, name):
    """Explicit central difference (GN22 beta2 = 0) in FP64 state: the
    one-launch step (interface nodes corrected one launch late) and the
    two-launch step (forced by a receiver) both <= 1e-11 from the oracle's
    restatement of gpuCalculateFEM_DYN; FP32 state <= 1e-4."""
    pb = _setup(tuned, name)
    dt, T = _dyn_params(pb)
    s, uo, n = O.dyn_solve(pb, 1000.0, 0.5, 0.0, dt, T, RICKER, np.float64,
                           compat_scramble=False)
    assert s == 0 and np.abs(uo).max() > 0
    kw = dict(rho=1000.0, beta1=0.5, beta2=0.0, dt=dt, endtime=T,
              wavelet_type=F.WAVELET_RICKER, freq=RICKER["freq"], ampl=RICKER["ampl"])
    u1, st1 = tuned.dyn_run(**kw)
    assert st1["steps"] == n
    assert rel(u1, uo) < 1e-11
    tuned.set_receivers(np.array([pb.n_nodes - 1]))
    u2, st2 = tuned.dyn_run(**kw)
    tuned.set_receivers(np.zeros(0, np.int32))
    assert st2["steps"] == n
    assert rel(u2, uo) < 1e-11
    u3, _ = tuned.dyn_run(precision=F.FP32, **kw)
    assert rel(u3, uo) < 1e-4


@pytest.mark.parametrize("name", ["column4x4x32", "cube5_t48"],
                         ids=[IDS[NAMES.index("column4x4x32")], IDS[NAMES.index("cube5_t48")]])
def test_implicit_reference_mode_vs_oracle(tuned, name):
    """Implicit GN22 in the reference's form with the inner PCG tightened to
    gamma < 1e-24 gamma0 on both sides: <= 1e-9 relative."""
    pb = _setup(tuned, name)
    s, uo, n = O.dyn_solve(pb, 1000.0, 0.5, 0.55, 1e-4, 1e-3, RICKER, np.float64,
                           compat_scramble=False, pcg_eps=1e-24)
    assert s == 0
    u, st = tuned.dyn_run(rho=1000.0, beta1=0.5, beta2=0.55, dt=1e-4, endtime=1e-3,
                          wavelet_type=F.WAVELET_RICKER, freq=RICKER["freq"],
                          ampl=RICKER["ampl"], pcg_eps=1e-24)
    assert st["steps"] == n
    assert rel(u, uo) < 1e-9
