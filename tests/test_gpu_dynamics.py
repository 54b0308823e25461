"""GPU parity of the GN22 steppers (gpuCalculateFEM_DYN, femfunc.cpp:130-381):
explicit central difference (beta2 = 0) and implicit Newmark, FP64 and FP32
state, against the oracle's restatement and the reference-build vectors.
Tolerances are stated per test."""
import json
import os

import numpy as np
import pytest

import cudafemproject_amd as F
import oracle as O
from conftest import GOLDEN
from cudafemproject_amd import problem as P
from helpers import dense_K

pytestmark = pytest.mark.gpu

WAV = {"ricker": F.WAVELET_RICKER, "berlage": F.WAVELET_BERLAGE, "": F.WAVELET_NONE}


def rel(a, b):
    a, b = np.asarray(a, float).ravel(), np.asarray(b, float).ravel()
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _ref():
    z = np.load(os.path.join(GOLDEN, "ref_vectors.npz"))
    return z, json.loads(bytes(z["_meta"]).decode())


REF, META = _ref()
DYN = sorted(k for k, m in META.items() if m["config"].get("task_type") == "dynamic")


def _case(name):
    m = META[name]
    pb = getattr(P, m["generator"])(**m["kwargs"])
    c = m["config"]
    w = c.get("wavelet", {})
    wav = {"type": w.get("WaveletType", ""), "freq": w.get("freq", 15.0),
           "timeshift": w.get("timeshift", 0.0), "ampl": w.get("ampl", 1.0)}
    return pb, c, wav


def _gpu(ctx, pb, c, wav, precision=F.FP64, pcg_eps=1e-4, **kw):
    ctx.set_problem(pb)
    return ctx.dyn_run(rho=c["rho"], beta1=c["beta1"], beta2=c["beta2"], dt=c["dt"],
                       endtime=c["endtime"], wavelet_type=WAV[wav["type"]], freq=wav["freq"],
                       timeshift=wav["timeshift"], ampl=wav["ampl"], precision=precision,
                       pcg_eps=pcg_eps, **kw)


@pytest.mark.parametrize("name", [n for n in DYN if "explicit" in n])
def test_explicit_fp64_vs_oracle(ctx, name):
    """Explicit steps are a fixed sequence of products and axpys: FP64 GPU
    vs FP64 oracle <= 1e-11 relative; vs the reference build (FP32) <= 1e-5."""
    pb, c, wav = _case(name)
    u, st = _gpu(ctx, pb, c, wav)
    s, uo, n = O.dyn_solve(pb, c["rho"], c["beta1"], c["beta2"], c["dt"], c["endtime"], wav,
                           np.float64, compat_scramble=False)
    assert s == 0 and st["steps"] == n
    assert rel(u, uo) < 1e-11
    assert rel(u, REF[name]) < 1e-5


@pytest.mark.parametrize("name", [n for n in DYN if "explicit" in n])
def test_explicit_fp32_state(ctx, name):
    pb, c, wav = _case(name)
    u, st = _gpu(ctx, pb, c, wav, precision=F.FP32)
    s, uo, n = O.dyn_solve(pb, c["rho"], c["beta1"], c["beta2"], c["dt"], c["endtime"], wav,
                           np.float64, compat_scramble=False)
    assert rel(u, uo) < 1e-4


@pytest.mark.parametrize("graph_steps", [1, 7, 64])
def test_explicit_graph_replays_match_eager(ctx, graph_steps):
    """Steps captured in HIP graphs of several lengths (plus an eager tail)
    give bitwise the same trajectory."""
    pb = P.column_problem(2, 2, 12, 0.1, 0.1, 1.2, convention=F.LOADS_PHYSICAL, pressure=1e3)
    ctx.set_problem(pb)
    kw = dict(rho=1000.0, dt=5e-6, endtime=5e-4, wavelet_type=F.WAVELET_RICKER, freq=500.0,
              ampl=1.0)
    a, _ = ctx.dyn_run(steps_per_graph=graph_steps, **kw)
    b, _ = ctx.dyn_run(steps_per_graph=100000, **kw)   # all steps eager
    assert np.array_equal(a, b)


def test_implicit_reference_mode_vs_oracle_tight(ctx):
    """Implicit GN22 in the reference's form (K a = F - K d, preconditioner
    cK diag K + diag M_consistent), inner PCG tightened to gamma < 1e-24
    gamma0 on both sides: <= 1e-9 relative."""
    name = [n for n in DYN if "implicit" in n][0]
    pb, c, wav = _case(name)
    u, st = _gpu(ctx, pb, c, wav, pcg_eps=1e-24)
    s, uo, n = O.dyn_solve(pb, c["rho"], c["beta1"], c["beta2"], c["dt"], c["endtime"], wav,
                           np.float64, compat_scramble=False, pcg_eps=1e-24)
    assert s == 0
    assert rel(u, uo) < 1e-9


def test_implicit_reference_mode_default_eps(ctx):
    """At the reference's inner tolerance (1e-4) both solvers stop early;
    the GPU result is as close to the oracle as the oracle is to the
    tightly solved trajectory."""
    name = [n for n in DYN if "implicit" in n][0]
    pb, c, wav = _case(name)
    u, st = _gpu(ctx, pb, c, wav, pcg_eps=1e-4)
    _, uo, _ = O.dyn_solve(pb, c["rho"], c["beta1"], c["beta2"], c["dt"], c["endtime"], wav,
                           np.float64, compat_scramble=False, pcg_eps=1e-4)
    _, ut, _ = O.dyn_solve(pb, c["rho"], c["beta1"], c["beta2"], c["dt"], c["endtime"], wav,
                           np.float64, compat_scramble=False, pcg_eps=1e-24)
    assert rel(u, uo) <= max(1e-8, 2.0 * rel(uo, ut))
    assert st["pcg_iterations"] > 0


def _dense_physical(pb, rho, beta1, beta2, dt, nsteps, f):
    """Dense restatement of the physical GN22 stepper: lumped mass rho V/4
    (3D) / rho A/3 (2D); implicit operator M + beta2 dt^2/2 K."""
    K = dense_K(pb)
    d = pb.dim
    m = np.zeros(pb.n_dof)
    for e in pb.elems:
        X = pb.nodes[e]
        if d == 3:
            v = abs(np.linalg.det(np.c_[np.ones(4), X])) / 6 / 4
        else:
            v = 0.5 * abs(np.linalg.det(np.c_[np.ones(3), X])) / 3
        for a in e:
            m[a * d:(a + 1) * d] += rho * v
    x, v, a = np.zeros(pb.n_dof), np.zeros(pb.n_dof), np.zeros(pb.n_dof)
    for _ in range(nsteps):
        x = x + dt * v + 0.5 * (1 - beta2) * dt * dt * a
        v = v + (1 - beta1) * dt * a
        r = f - K @ x
        if beta2 == 0:
            a = r / m
        else:
            a = np.linalg.solve(np.diag(m) + 0.5 * beta2 * dt * dt * K, r)
        v = v + beta1 * dt * a
        x = x + 0.5 * beta2 * dt * dt * a
    return x.reshape(-1, d)


@pytest.mark.parametrize("beta2", [0.0, 0.6])
def test_physical_modes_vs_dense(ctx, beta2):
    pb = P.column_problem(1, 1, 4, 0.1, 0.1, 0.4, convention=F.LOADS_PHYSICAL, pressure=1e4)
    ctx.set_problem(pb)
    f = ctx.rhs().ravel()
    rho, dt, T = 1000.0, 2e-6, 2e-4
    u, st = ctx.dyn_run(rho=rho, beta1=0.5, beta2=beta2, dt=dt, endtime=T,
                        mass_convention=F.MASS_PHYSICAL, implicit_mode=F.IMPLICIT_PHYSICAL,
                        pcg_eps=1e-26)
    nsteps = int(T / dt)
    ud = _dense_physical(pb, rho, 0.5, beta2, dt, nsteps, f)
    assert st["steps"] == nsteps
    assert rel(u, ud) < 1e-8
