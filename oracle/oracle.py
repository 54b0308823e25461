"""ctypes wrapper of the C oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker.  The product never uses it.
Also runs the reference's own CPU build (oracle/_ref/ref_driver) where present.
"""
from __future__ import annotations

import ctypes as C
import json
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_DRIVER = os.path.join(HERE, "_ref", "ref_driver")

ORC_OK, ORC_E_ARG, ORC_E_NONFINITE, ORC_E_MAXIT, ORC_E_UNSUPPORTED = 0, -1, -2, -3, -4

_D = C.POINTER(C.c_double)
_F = C.POINTER(C.c_float)
_I = C.POINTER(C.c_int)


class OrcProblem(C.Structure):
    _fields_ = [("dim", C.c_int), ("n_nodes", C.c_int), ("n_elems", C.c_int),
                ("nodes", _D), ("elems", _I),
                ("n_constr", C.c_int), ("constr_node", _I), ("constr_code", _I),
                ("n_faces", C.c_int), ("face_nodes", _I), ("face_adj", _I),
                ("face_normal", _D), ("face_p", _D),
                ("n_loads", C.c_int), ("load_node", _I), ("load_vec", _D),
                ("E", C.c_double), ("nu", C.c_double), ("elem_scale", _D)]


class OrcWavelet(C.Structure):
    _fields_ = [("type", C.c_int), ("freq", C.c_double), ("timeshift", C.c_double),
                ("ampl", C.c_double)]


class OrcDyn(C.Structure):
    _fields_ = [("rho", C.c_double), ("beta1", C.c_double), ("beta2", C.c_double),
                ("dt", C.c_double), ("endtime", C.c_double), ("wavelet", OrcWavelet),
                ("compat_scramble", C.c_int), ("pcg_eps", C.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} not built (make -C oracle)")
        _lib = C.CDLL(LIB)
    return _lib


class _Pinned:
    """Keeps numpy arrays alive while the C struct points at them."""

    def __init__(self, pb):
        d = pb.dim
        self.nodes = np.ascontiguousarray(pb.nodes, np.float64)
        self.elems = np.ascontiguousarray(pb.elems, np.int32)
        self.cn = np.ascontiguousarray(pb.constr_node, np.int32)
        self.cc = np.ascontiguousarray(pb.constr_code, np.int32)
        self.fn = np.ascontiguousarray(pb.face_nodes, np.int32).reshape(-1, d)
        self.fa = np.ascontiguousarray(pb.face_adj, np.int32)
        self.fnorm = np.ascontiguousarray(pb.face_normal, np.float64).reshape(-1, d)
        self.fp = np.ascontiguousarray(pb.face_p, np.float64)
        self.ln = np.ascontiguousarray(pb.load_node, np.int32)
        self.lv = np.ascontiguousarray(pb.load_vec, np.float64).reshape(-1, d)
        self.es = None if pb.elem_scale is None else np.ascontiguousarray(pb.elem_scale, np.float64)
        s = OrcProblem()
        s.dim, s.n_nodes, s.n_elems = d, pb.n_nodes, pb.n_elems
        s.nodes = self.nodes.ctypes.data_as(_D)
        s.elems = self.elems.ctypes.data_as(_I)
        s.n_constr = len(self.cn)
        s.constr_node = self.cn.ctypes.data_as(_I)
        s.constr_code = self.cc.ctypes.data_as(_I)
        s.n_faces = len(self.fa)
        s.face_nodes = self.fn.ctypes.data_as(_I)
        s.face_adj = self.fa.ctypes.data_as(_I)
        s.face_normal = self.fnorm.ctypes.data_as(_D)
        s.face_p = self.fp.ctypes.data_as(_D)
        s.n_loads = len(self.ln)
        s.load_node = self.ln.ctypes.data_as(_I)
        s.load_vec = self.lv.ctypes.data_as(_D)
        s.E, s.nu = pb.E, pb.nu
        s.elem_scale = self.es.ctypes.data_as(_D) if self.es is not None else None
        self.s = s


def _sfx(dtype):
    return ("_f32", np.float32, _F) if np.dtype(dtype) == np.float32 else ("_f64", np.float64, _D)


def _call(name, *args):
    st = getattr(lib(), name)(*args)
    return st


def static_solve(pb, dtype=np.float32, eps=1e-8, max_iter=0):
    sfx, dt, P = _sfx(dtype)
    pin = _Pinned(pb)
    u = np.zeros((pb.n_nodes, pb.dim), dt)
    it = C.c_int(0)
    f = getattr(lib(), "orc_static_solve" + sfx)
    f.argtypes = [C.POINTER(OrcProblem), C.c_double, C.c_int, P, C.POINTER(C.c_int)]
    st = f(C.byref(pin.s), eps, max_iter, u.ctypes.data_as(P), C.byref(it))
    return st, u, it.value


def klocals(pb, dtype=np.float64):
    sfx, dt, P = _sfx(dtype)
    pin = _Pinned(pb)
    nde = pb.dim * (pb.dim + 1)
    K = np.zeros((pb.n_elems, nde, nde), dt)   # block stored column-major
    A = np.zeros(pb.n_elems, dt)
    f = getattr(lib(), "orc_klocals" + sfx)
    f.argtypes = [C.POINTER(OrcProblem), P, P]
    st = f(C.byref(pin.s), K.ctypes.data_as(P), A.ctypes.data_as(P))
    assert st == ORC_OK, st
    return K.transpose(0, 2, 1).copy(), A      # row-major blocks


def _vec_fn(name, pb, dtype, x=None):
    sfx, dt, P = _sfx(dtype)
    pin = _Pinned(pb)
    out = np.zeros((pb.n_nodes, pb.dim), dt)
    f = getattr(lib(), name + sfx)
    if x is None:
        f.argtypes = [C.POINTER(OrcProblem), P]
        st = f(C.byref(pin.s), out.ctypes.data_as(P))
    else:
        xx = np.ascontiguousarray(x, dt).reshape(pb.n_nodes, pb.dim)
        f.argtypes = [C.POINTER(OrcProblem), P, P]
        st = f(C.byref(pin.s), xx.ctypes.data_as(P), out.ctypes.data_as(P))
    assert st == ORC_OK, st
    return out


def rhs(pb, dtype=np.float64):
    return _vec_fn("orc_rhs", pb, dtype)


def diag(pb, dtype=np.float64):
    return _vec_fn("orc_diag", pb, dtype)


def matvec(pb, x, dtype=np.float64):
    return _vec_fn("orc_matvec", pb, dtype, x)


def stress(pb, u, dtype=np.float64):
    sfx, dt, P = _sfx(dtype)
    pin = _Pinned(pb)
    nv = 3 * (pb.dim - 1)
    uu = np.ascontiguousarray(u, dt).reshape(pb.n_nodes, pb.dim)
    s = np.zeros((pb.n_elems, nv), dt)
    f = getattr(lib(), "orc_stress" + sfx)
    f.argtypes = [C.POINTER(OrcProblem), P, P]
    st = f(C.byref(pin.s), uu.ctypes.data_as(P), s.ctypes.data_as(P))
    assert st == ORC_OK, st
    return s


WAVELETS = {"": 0, "none": 0, "ricker": 1, "berlage": 2}


def dyn_solve(pb, rho, beta1, beta2, dt, endtime, wavelet=None, dtype=np.float32,
              compat_scramble=True, pcg_eps=0.0):
    sfx, dtp, P = _sfx(dtype)
    pin = _Pinned(pb)
    d = OrcDyn()
    d.rho, d.beta1, d.beta2, d.dt, d.endtime = rho, beta1, beta2, dt, endtime
    w = wavelet or {}
    d.wavelet.type = WAVELETS.get(w.get("type", ""), 0)
    d.wavelet.freq = w.get("freq", 15.0)
    d.wavelet.timeshift = w.get("timeshift", 0.0)
    d.wavelet.ampl = w.get("ampl", 1.0)
    d.compat_scramble = int(compat_scramble)
    d.pcg_eps = pcg_eps
    u = np.zeros((pb.n_nodes, pb.dim), dtp)
    n = C.c_int(0)
    f = getattr(lib(), "orc_dyn_solve" + sfx)
    f.argtypes = [C.POINTER(OrcProblem), C.POINTER(OrcDyn), P, C.POINTER(C.c_int)]
    st = f(C.byref(pin.s), C.byref(d), u.ctypes.data_as(P), C.byref(n))
    return st, u, n.value


def static_time_f32(pb, n_iter):
    """(setup_s, iter_s) of the FP32 restatement on one core."""
    pin = _Pinned(pb)
    a, b = C.c_double(0), C.c_double(0)
    f = lib().orc_static_time_f32
    f.argtypes = [C.POINTER(OrcProblem), C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    st = f(C.byref(pin.s), n_iter, C.byref(a), C.byref(b))
    assert st == ORC_OK, st
    return a.value, b.value


def static_time_omp_f64(pb, n_iter, n_threads):
    """(setup_s, iter_s) of the FP64 restatement with OpenMP loops."""
    pin = _Pinned(pb)
    a, b = C.c_double(0), C.c_double(0)
    f = lib().orc_static_time_omp_f64
    f.argtypes = [C.POINTER(OrcProblem), C.c_int, C.c_int, C.POINTER(C.c_double),
                  C.POINTER(C.c_double)]
    st = f(C.byref(pin.s), n_iter, n_threads, C.byref(a), C.byref(b))
    assert st == ORC_OK, st
    return a.value, b.value


def wavelet_value(t, wavelet):
    f = lib().orc_wavelet_value
    f.restype = C.c_float
    f.argtypes = [C.c_double, C.POINTER(OrcWavelet)]
    w = OrcWavelet(WAVELETS.get(wavelet.get("type", ""), 0), wavelet.get("freq", 15.0),
                   wavelet.get("timeshift", 0.0), wavelet.get("ampl", 1.0))
    return f(t, C.byref(w))


# --------------------------------------------------------------------------
# the reference's own CPU build (oracle/_ref), only where it was built

def ref_available() -> bool:
    return os.path.exists(REF_DRIVER)


def run_reference(pb, config: dict | None = None, task="task"):
    """Write pb as a prepared mesh, run the reference CPU path on it, return
    the displacement array it computed (float32 precision, printed %.9g)."""
    from cudafemproject_amd.problem import write_prepared
    cfg = dict(config or {})
    with tempfile.TemporaryDirectory() as td:
        ws = td + "/"
        write_prepared(pb, os.path.join(ws, "prepared_meshes", f"{pb.dim}D", task))
        os.makedirs(os.path.join(ws, "fem_results", f"{pb.dim}D", task), exist_ok=True)
        cfg.setdefault("DIM", pb.dim)
        cfg.setdefault("youngModulus", pb.E)
        cfg.setdefault("poissonRatio", pb.nu)
        cfg["working_path"] = ws
        cpath = os.path.join(td, "cfg.json")
        with open(cpath, "w") as f:
            json.dump(cfg, f)
        out = os.path.join(td, "u.txt")
        subprocess.run([REF_DRIVER, cpath, task, out], check=True, capture_output=True,
                       timeout=600)
        with open(out) as f:
            n, d = map(int, f.readline().split())
            u = np.loadtxt(f).reshape(n, d)
    return u
