"""GPU: FEMGPU_NO_GRAPHS=1 (PCG iteration blocks and time steps launched
directly instead of replayed as captured HIP graphs, for tracing tools) runs
the same kernels in the same order: bitwise the same displacements and
iteration counts, for both PCG recurrences, the mixed solve, and the
explicit and implicit time steppers in both precisions."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

import cudafemproject_amd as F
from conftest import ROOT
from cudafemproject_amd.problem import cube_problem

pytestmark = pytest.mark.gpu

CHILD = r'''
import json, sys
import numpy as np
sys.path.insert(0, sys.argv[1])
import cudafemproject_amd as F
from cudafemproject_amd.problem import cube_problem
pb = cube_problem(10, convention=F.LOADS_REFERENCE)
out = {}
with F.Context(0) as c:
    c.set_tuning(200, 512)
    c.set_problem(pb)
    for name, prec, rec in (("fp64", F.FP64, F.CG_STANDARD), ("cg1", F.FP64, F.CG_SINGLE_REDUCTION),
                            ("mixed", F.MIXED, F.CG_STANDARD)):
        u, s = c.solve(prec, tol=1e-10, tol_kind=F.TOL_REL_RES, recurrence=rec)
        out[name] = [u.ravel().tolist(), s["iterations"]]
from cudafemproject_amd.problem import column_problem
pc = column_problem(4, 4, 64, 0.025, 0.025, 0.4, convention=F.LOADS_PHYSICAL)
with F.Context(0) as c:
    c.set_problem(pc)
    for beta2, steps in ((0.0, 150), (0.5, 40)):
        for prec in (F.FP64, F.FP32):
            dt = 2e-6 if beta2 == 0 else 8e-6
            u, s = c.dyn_run(rho=1000.0, beta1=0.5, beta2=beta2, dt=dt, endtime=(steps + 0.5) * dt,
                             wavelet_type=F.WAVELET_RICKER, freq=2000.0, ampl=1e3, precision=prec,
                             mass_convention=F.MASS_PHYSICAL, implicit_mode=F.IMPLICIT_PHYSICAL,
                             pcg_eps=1e-8)
            out[f"dyn{beta2}/{prec}"] = [u.ravel().tolist(), s["steps"] * 100000 + s["pcg_iterations"]]
print(json.dumps(out))
'''


def _run(no_graphs):
    env = dict(os.environ)
    env.pop("FEMGPU_NO_GRAPHS", None)
    if no_graphs:
        env["FEMGPU_NO_GRAPHS"] = "1"
    r = subprocess.run([sys.executable, "-c", CHILD, ROOT], capture_output=True, text=True,
                       env=env, timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    return json.loads(r.stdout.strip().splitlines()[-1])


def test_no_graphs_matches_graphs():
    a, b = _run(False), _run(True)
    for k in a:
        print(f"[nograph] {k}: iterations {a[k][1]} / {b[k][1]}", flush=True)
        assert a[k][1] == b[k][1], k
        assert np.array_equal(np.array(a[k][0]), np.array(b[k][0])), k
