"""GPU: receiver traces, SEG-Y output, Rayleigh damping and relaxation of the
GN22 steppers (the parts of gpuCalculateFEM_DYN that are switched off in the
reference, femfunc.cpp:137-151,213-243,349-366).  Each is checked against an
independent consequence: traces against separate runs and the central
difference identities of GN22 with beta1 = 1/2; damping against the modal
decay law of mass-proportional damping and against the explicit stepper;
relaxation against the static solve."""
import json
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

import cudafemproject_amd as F
from conftest import GOLDEN, ROOT
from cudafemproject_amd import problem as P

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a, float).ravel(), np.asarray(b, float).ravel()
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def column(nz=16):
    return P.column_problem(2, 2, nz, 0.1, 0.1, 0.1 * nz / 2, convention=F.LOADS_PHYSICAL,
                            pressure=1.0)


KW = dict(rho=1000.0, beta1=0.5, beta2=0.0, dt=2e-5, wavelet_type=F.WAVELET_RICKER, freq=500.0,
          ampl=1e3, mass_convention=F.MASS_PHYSICAL)


def test_traces_equal_separate_runs(ctx):
    pb = column()
    ctx.set_problem(pb)
    recv = np.array([0, 5, pb.n_nodes // 2, pb.n_nodes - 1], np.int32)
    ctx.set_receivers(recv, F.RECV_DISPLACEMENT, 1)
    n = 80                                       # graph blocks of 64 and an eager tail
    ctx.dyn_run(endtime=(n + 0.5) * KW["dt"], **KW)
    tr = ctx.traces()
    assert tr.shape == (n, len(recv), 3)
    ctx.set_receivers([], F.RECV_DISPLACEMENT, 1)
    for k in (1, 10, 64, n):
        u, st = ctx.dyn_run(endtime=(k + 0.5) * KW["dt"], **KW)
        assert st["steps"] == k
        np.testing.assert_allclose(tr[k - 1], u[recv], rtol=1e-13, atol=1e-30)


@pytest.mark.parametrize("alpha", [0.0, 30.0])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("shape", ["column", "plate"])
def test_one_launch_step_matches_two_launch_step(ctx, shape, precision, alpha):
    """Without receivers an explicit step is one patch launch whose interface
    nodes are corrected one launch late; with receivers it is the predictor
    launch plus the corrector launch.  Both give the same displacements
    (meshes of 12 and 6 patches, so interface nodes are involved; 150 steps:
    two graph replays and an eager tail; with and without mass damping)."""
    pb = (P.column_problem(8, 8, 24, 0.4, 0.4, 1.2, convention=F.LOADS_PHYSICAL, pressure=1.0)
          if shape == "column" else P.plate_problem(64, 32, convention=F.LOADS_PHYSICAL))
    ctx.set_problem(pb)
    n = 150
    kw = dict(KW, precision=F.FP64 if precision == "fp64" else F.FP32, damping_alpha=alpha,
              endtime=(n + 0.5) * KW["dt"])
    u1, st1 = ctx.dyn_run(**kw)
    ctx.set_receivers(np.array([0], np.int32), F.RECV_VELOCITY, 1)
    try:
        u2, st2 = ctx.dyn_run(**kw)
    finally:
        ctx.set_receivers([], F.RECV_DISPLACEMENT, 1)
    assert st1["steps"] == st2["steps"] == n
    assert np.linalg.norm(u2) > 0
    assert rel(u1, u2) < (1e-12 if precision == "fp64" else 1e-5)


@pytest.mark.parametrize("tuning", [(1536, 512), (3072, 1024)])
def test_one_launch_step_other_tunings(ctx, tuning):
    """The one-launch explicit step with 512- and 1024-thread patches (the
    FP64 kernels of the larger workgroups spill registers) against the
    two-launch step."""
    pb = P.column_problem(8, 8, 48, 0.4, 0.4, 2.4, convention=F.LOADS_PHYSICAL, pressure=1.0)
    ctx.set_tuning(*tuning)
    try:
        ctx.set_problem(pb)
        n = 70
        kw = dict(KW, endtime=(n + 0.5) * KW["dt"])
        u1, _ = ctx.dyn_run(**kw)
        ctx.set_receivers(np.array([0], np.int32), F.RECV_DISPLACEMENT, 1)
        try:
            u2, _ = ctx.dyn_run(**kw)
        finally:
            ctx.set_receivers([], F.RECV_DISPLACEMENT, 1)
    finally:
        ctx.set_tuning(0, 0)
    assert np.linalg.norm(u2) > 0
    assert rel(u1, u2) < 1e-12


def test_velocity_acceleration_central_differences(ctx):
    """GN22 with beta1 = 1/2, beta2 = 0: v_n = (d_{n+1} - d_{n-1}) / (2 dt) and
    a_n = (d_{n+1} - 2 d_n + d_{n-1}) / dt^2 exactly (to rounding)."""
    pb = column()
    ctx.set_problem(pb)
    recv = np.arange(0, pb.n_nodes, 7, dtype=np.int32)
    out = {}
    for q in (F.RECV_DISPLACEMENT, F.RECV_VELOCITY, F.RECV_ACCELERATION):
        ctx.set_receivers(recv, q, 1)
        ctx.dyn_run(endtime=100.5 * KW["dt"], **KW)
        out[q] = ctx.traces()
    d, v, a = out[F.RECV_DISPLACEMENT], out[F.RECV_VELOCITY], out[F.RECV_ACCELERATION]
    dt = KW["dt"]
    vc = (d[2:] - d[:-2]) / (2 * dt)
    ac = (d[2:] - 2 * d[1:-1] + d[:-2]) / dt ** 2
    assert np.abs(vc - v[1:-1]).max() <= 1e-8 * np.abs(v).max()
    assert np.abs(ac - a[1:-1]).max() <= 1e-6 * np.abs(a).max()


def test_sampling_period(ctx):
    pb = column()
    ctx.set_problem(pb)
    recv = np.array([3, 17], np.int32)
    ctx.set_receivers(recv, F.RECV_VELOCITY, 1)
    ctx.dyn_run(endtime=90.5 * KW["dt"], **KW)
    t1 = ctx.traces()
    ctx.set_receivers(recv, F.RECV_VELOCITY, 7)
    ctx.dyn_run(endtime=90.5 * KW["dt"], **KW)
    t7 = ctx.traces()
    assert t7.shape[0] == 90 // 7
    np.testing.assert_array_equal(t7, t1[6::7][:t7.shape[0]])


def lumped_mass(pb, rho):
    X = pb.nodes[pb.elems]
    V = np.abs(np.einsum("ei,ei->e", X[:, 1] - X[:, 0],
                         np.cross(X[:, 2] - X[:, 0], X[:, 3] - X[:, 0]))) / 6
    m = np.zeros(pb.n_nodes)
    np.add.at(m, pb.elems.ravel(), np.repeat(rho * V / 4, 4))
    return m


def energy(ctx, pb, d, v, rho):
    m = lumped_mass(pb, rho)
    return 0.5 * np.sum(m[:, None] * v * v) + 0.5 * np.sum(d * ctx.matvec(d, F.FP64))


def run_energy(ctx, pb, T, **kw):
    everything = np.arange(pb.n_nodes, dtype=np.int32)
    steps = int(T / kw["dt"])
    ctx.set_receivers(everything, F.RECV_VELOCITY, steps)
    d, st = ctx.dyn_run(endtime=(steps + 0.5) * kw["dt"], **kw)
    v = ctx.traces()[-1]
    ctx.set_receivers([], F.RECV_VELOCITY, 1)
    return energy(ctx, pb, d, v, kw["rho"]), d


def test_zero_damping_is_undamped(ctx):
    pb = column()
    ctx.set_problem(pb)
    a, _ = ctx.dyn_run(endtime=50.5 * KW["dt"], **KW)
    b, _ = ctx.dyn_run(endtime=50.5 * KW["dt"], damping_alpha=0.0, damping_beta=0.0, **KW)
    np.testing.assert_array_equal(a, b)


def test_mass_damping_decay_law(ctx):
    """M a + alpha M v + K d = 0 after the pulse: every mode's energy decays as
    exp(-alpha t) (up to terms of relative size alpha / (2 omega))."""
    pb = column()
    ctx.set_problem(pb)
    T, alpha, t_pulse = 0.03, 40.0, 1.0 / KW["freq"]     # the Ricker peak
    e0, _ = run_energy(ctx, pb, T, **KW)
    ea, _ = run_energy(ctx, pb, T, damping_alpha=alpha, **KW)
    ratio = ea / e0
    expect = np.exp(-alpha * (T - t_pulse))
    print(f"energy ratio {ratio:.4f} expected {expect:.4f}", flush=True)
    assert abs(ratio / expect - 1) < 0.03


def test_implicit_mass_damping_matches_explicit(ctx):
    """Same damped problem through the implicit Newmark operator
    ((1 + alpha beta1 dt) M + beta2 dt^2/2 K) and the explicit stepper: both
    converge to the same motion as dt -> 0."""
    pb = column(8)
    ctx.set_problem(pb)
    kw = dict(KW, dt=2e-6, freq=400.0)
    T = 4e-3
    ue, _ = ctx.dyn_run(endtime=T, damping_alpha=200.0, **kw)
    ui, st = ctx.dyn_run(endtime=T, damping_alpha=200.0, implicit_mode=F.IMPLICIT_PHYSICAL,
                         pcg_eps=1e-20, **dict(kw, beta2=0.5))
    print(f"implicit vs explicit (damped) rel diff {rel(ui, ue):.2e}", flush=True)
    assert rel(ui, ue) < 2e-4
    u0, _ = ctx.dyn_run(endtime=T, implicit_mode=F.IMPLICIT_PHYSICAL, pcg_eps=1e-20,
                        **dict(kw, beta2=0.5))
    assert rel(u0, ue) > 5 * rel(ui, ue)        # damping visibly changes the motion


def test_implicit_stiffness_damping_dissipates(ctx):
    pb = column(8)
    ctx.set_problem(pb)
    kw = dict(KW, beta2=0.5, dt=1e-5, implicit_mode=F.IMPLICIT_PHYSICAL, pcg_eps=1e-16)
    T = 5e-3
    e = [run_energy(ctx, pb, T, damping_beta=b, **kw)[0] for b in (0.0, 1e-6, 1e-5)]
    print("energies", e, flush=True)
    assert e[0] > e[1] > e[2] > 0


def test_dynamic_relaxation_reaches_static_solution(ctx):
    """Constant load, mass damping near critical, relaxation until max|v| < eps:
    the displacement is the static solution."""
    pb = P.cube_problem(4, E=2e7, nu=0.25, pressure=1e3, convention=F.LOADS_PHYSICAL)
    ctx.set_problem(pb)
    us, _ = ctx.solve(F.FP64, tol=1e-12, tol_kind=F.TOL_REL_RES)
    u, st = ctx.dyn_run(rho=1000.0, beta1=0.5, beta2=0.0, dt=2e-5, endtime=-1e-9,
                        mass_convention=F.MASS_PHYSICAL, damping_alpha=500.0,
                        relax_reference=False, max_steps=400000, steps_per_graph=256)
    print(f"relaxation: {st['steps']} steps, rel diff to static {rel(u, us):.2e}", flush=True)
    assert st["status"] == 0 and st["steps"] < 400000
    assert rel(u, us) < 1e-5


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_implicit_relaxation_equals_fixed_run(ctx, precision):
    """Implicit stepping with relaxation (max|v| checked after every step,
    each step's end launched before the check) ends at the static solution
    and equals, bitwise, a run of the same number of steps without the check
    (each step's end captured with the next step's start)."""
    pb = P.cube_problem(4, E=2e7, nu=0.25, pressure=1e3, convention=F.LOADS_PHYSICAL)
    ctx.set_problem(pb)
    us, _ = ctx.solve(F.FP64, tol=1e-12, tol_kind=F.TOL_REL_RES)
    prec = F.FP64 if precision == "fp64" else F.FP32
    kw = dict(rho=1000.0, beta1=0.5, beta2=0.5, dt=2e-4, mass_convention=F.MASS_PHYSICAL,
              implicit_mode=F.IMPLICIT_PHYSICAL, pcg_eps=1e-10 if prec == F.FP64 else 1e-6,
              damping_alpha=500.0, precision=prec)
    # the distance to the static displacement is about max|v| at the stop
    # over the slowest mode's decay rate (~40/s here): 1.4e-5 relative at
    # max|v| < 1e-8, so 1e-10 (FP64) and 1e-7 (FP32)
    eps = 1e-10 if prec == F.FP64 else 1e-7
    u, st = ctx.dyn_run(endtime=-eps, relax_reference=False, max_steps=20000, **kw)
    n = st["steps"]
    print(f"implicit relaxation {precision}: {n} steps, {st['pcg_iterations']} PCG iterations, "
          f"rel diff to static {rel(u, us):.2e}", flush=True)
    assert st["status"] == 0 and 1 < n < 20000
    assert rel(u, us) < (1e-5 if prec == F.FP64 else 1e-3)
    u2, st2 = ctx.dyn_run(endtime=(n + 0.5) * kw["dt"], **kw)
    assert st2["steps"] == n and st2["pcg_iterations"] == st["pcg_iterations"]
    assert np.array_equal(u, u2)


def test_implicit_step_graph_release(ctx):
    """Implicit runs under many option sets (one dt each) leave their step
    graphs cached until more than 256 belong to other option sets; the run
    that releases them, and every run after, gives bitwise the results of the
    same options in a fresh context."""
    pb = column(8)
    ctx.set_problem(pb)
    kw = dict(KW, beta2=0.5, implicit_mode=F.IMPLICIT_PHYSICAL, pcg_eps=1e-8)

    def run(c, dt):
        u, st = c.dyn_run(**dict(kw, dt=dt, endtime=3.5 * dt))
        assert st["steps"] == 3
        return u

    u_first = run(ctx, 4e-5)
    for k in range(140):                         # 2 to 3 step graphs each
        run(ctx, 4e-5 * (1.0 + 0.001 * (k + 1)))
    u_again = run(ctx, 4e-5)
    u_other = run(ctx, 4e-5 * 1.0005)
    with F.Context(0) as c2:
        c2.set_problem(pb)
        u_ref = run(c2, 4e-5)
        u_ref_other = run(c2, 4e-5 * 1.0005)
    assert np.array_equal(u_first, u_ref)
    assert np.array_equal(u_again, u_ref)
    assert np.array_equal(u_other, u_ref_other)


def test_relaxation_reference_stub_is_one_step(ctx):
    pb = column()
    ctx.set_problem(pb)
    _, st = ctx.dyn_run(endtime=-1.0, **KW)
    assert st["steps"] == 1


def test_fem_main_writes_receivers_and_segy(tmp_path, fem_main):
    from test_femproject_cpu import workspace
    cfg = workspace(tmp_path, "test_rect_pcg", task_type="dynamic", rho=2400, dt=1e-5,
                    endtime=2.05e-3, beta1=0.5, beta2=0.0,
                    wavelet={"WaveletType": "ricker", "freq": 1000, "ampl": 1e3})
    r = subprocess.run([fem_main, cfg, "test_rect_pcg",
                        "--receivers", "1,3,5", "--receiver-axis", "0", "--segy-every", "4"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    res = tmp_path / "fem_results" / "2D" / "test_rect_pcg"
    rows = (res / "receivers.txt").read_text().splitlines()
    assert rows[0].split() == ["3", "1", "3", "5"]
    vals = np.array([[float(x) for x in ln.split()] for ln in rows[1:]])
    assert vals.shape == (205 // 4, 4)
    np.testing.assert_allclose(vals[:, 0], 4e-5 * np.arange(1, 52), rtol=1e-6)
    b = (res / "receivers.sgy").read_bytes()
    ntr, ns = struct.unpack(">h", b[3212:3214])[0], struct.unpack(">h", b[3220:3222])[0]
    assert (ntr, ns) == (3, 51) and struct.unpack(">h", b[3216:3218])[0] == 40
    for t in range(3):
        o = 3600 + t * (240 + 4 * ns) + 240
        tr = np.frombuffer(b[o:o + 4 * ns], ">f4")
        np.testing.assert_allclose(tr, vals[:, 1 + t], rtol=1e-6, atol=1e-30)


# ---------------------------------------------------------------------------
# Restatements: receiver traces against the oracle's GN22 stepper, damped
# stepping against a dense GN22 with Rayleigh damping.

@pytest.mark.parametrize("tuning", [(0, 0), (48, 256)])
def test_traces_vs_oracle(ctx, tuning):
    """Displacement traces (every 4th step) equal the oracle's restatement of
    gpuCalculateFEM_DYN (femfunc.cpp:130-381) stopped at each sampled step
    (reference conventions, Ricker wavelet): <= 1e-11 relative per sample,
    on one patch and on 48-element patches (receivers on interface nodes)."""
    import oracle as O
    pb = P.column_problem(2, 2, 8, 0.1, 0.1, 0.8, pressure=1.0, convention=F.LOADS_REFERENCE)
    rho, dt, every, n = 1000.0, 1e-4, 4, 24
    wav = {"type": "ricker", "freq": 50.0, "timeshift": 0.0, "ampl": 1000.0}
    recv = np.array([3, 20, pb.n_nodes // 2, pb.n_nodes - 1], np.int32)
    ctx.set_tuning(*tuning)
    try:
        ctx.set_problem(pb)
        ctx.set_receivers(recv, F.RECV_DISPLACEMENT, every)
        ctx.dyn_run(rho=rho, beta1=0.5, beta2=0.0, dt=dt, endtime=(n + 0.5) * dt,
                    wavelet_type=F.WAVELET_RICKER, freq=50.0, ampl=1000.0)
        tr = ctx.traces()
        ctx.set_receivers([], F.RECV_DISPLACEMENT, 1)
    finally:
        ctx.set_tuning(0, 0)
    assert tr.shape == (n // every, len(recv), 3)
    for k in range(1, n // every + 1):
        s, uo, steps = O.dyn_solve(pb, rho, 0.5, 0.0, dt, (k * every + 0.5) * dt, wav,
                                   np.float64, compat_scramble=False)
        assert s == 0 and steps == k * every
        assert rel(tr[k - 1], uo[recv]) < 1e-11, k


def _dense_damped(pb, rho, beta1, beta2, dt, nsteps, f, alpha, beta):
    """Dense GN22 with Rayleigh damping C = alpha M + beta K, lumped mass
    rho V/4 (3D): M a' + C v' + K d' = f with d~ = d + dt v + (1-beta2) dt^2/2 a,
    v~ = v + (1-beta1) dt a, v' = v~ + beta1 dt a', d' = d~ + beta2 dt^2/2 a'."""
    from helpers import dense_K
    K = dense_K(pb)
    m = np.zeros(pb.n_dof)
    for e in pb.elems:
        v = abs(np.linalg.det(np.c_[np.ones(4), pb.nodes[e]])) / 6 / 4
        for a in e:
            m[3 * a:3 * a + 3] += rho * v
    M = np.diag(m)
    C = alpha * M + beta * K
    A = M + beta1 * dt * C + 0.5 * beta2 * dt * dt * K
    x, v, a = np.zeros(pb.n_dof), np.zeros(pb.n_dof), np.zeros(pb.n_dof)
    for _ in range(nsteps):
        xt = x + dt * v + 0.5 * (1 - beta2) * dt * dt * a
        vt = v + (1 - beta1) * dt * a
        a = np.linalg.solve(A, f - C @ vt - K @ xt)
        v = vt + beta1 * dt * a
        x = xt + 0.5 * beta2 * dt * dt * a
    return x.reshape(-1, 3)


@pytest.mark.parametrize("beta2,alpha,beta", [(0.0, 300.0, 0.0), (0.6, 300.0, 0.0),
                                              (0.6, 0.0, 2e-5), (0.6, 300.0, 2e-5)])
def test_rayleigh_damping_vs_dense(ctx, beta2, alpha, beta):
    """Mass- and stiffness-proportional damping (damping_alpha, damping_beta,
    femfunc.cpp:137-150,190-200,266-274), explicit (beta = 0) and implicit
    physical mode, FP64 state: <= 1e-8 relative to the dense restatement."""
    pb = P.column_problem(1, 1, 4, 0.1, 0.1, 0.4, convention=F.LOADS_PHYSICAL, pressure=1e4)
    ctx.set_problem(pb)
    f = ctx.rhs().ravel()
    rho, dt, T = 1000.0, 2e-6, 2e-4
    u, st = ctx.dyn_run(rho=rho, beta1=0.5, beta2=beta2, dt=dt, endtime=T,
                        mass_convention=F.MASS_PHYSICAL, implicit_mode=F.IMPLICIT_PHYSICAL,
                        pcg_eps=1e-26, damping_alpha=alpha, damping_beta=beta)
    nsteps = int(T / dt)
    assert st["steps"] == nsteps
    ud = _dense_damped(pb, rho, 0.5, beta2, dt, nsteps, f, alpha, beta)
    und = _dense_damped(pb, rho, 0.5, beta2, dt, nsteps, f, 0.0, 0.0)
    assert rel(ud, und) > 1e-4          # the damping changes the trajectory
    assert rel(u, ud) < 1e-8
