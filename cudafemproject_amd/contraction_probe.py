"""ctypes binding of libfemgpu_probe.so: measurement kernels of the element
contraction of the batched plate solves, F_e = K_e U_e for four load cases
per linear triangle, on the fly (variant 0, the form the solver uses) or
through K_e and v_mfma_f32_4x4x1f32 (variant 1).  Used by
``bench.py --workload contraction`` and its GPU test; the solver does not
load it.  Device arrays are passed as torch CUDA tensors (plumbing only).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libfemgpu_probe.so")
N_CASES = 4
_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} is not built (cudafemproject_amd.build.build_probe())")
        lib = C.CDLL(LIB)
        lib.femgpu_probe_contraction.restype = C.c_int
        lib.femgpu_probe_contraction.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                                 C.c_void_p, C.c_float, C.c_float, C.c_void_p,
                                                 C.c_int, C.c_void_p]
        lib.femgpu_probe_mfma_layout.restype = C.c_int
        lib.femgpu_probe_mfma_layout.argtypes = [C.c_void_p, C.c_void_p]
        _lib = lib
    return _lib


def flat_plates(insts):
    """The triangles of several plates as one mesh (node ids offset per
    plate): coordinates (n, 2) float32, elements (m, 3) int32."""
    coords, elems, off = [], [], 0
    for b in insts:
        coords.append(np.asarray(b.nodes, np.float32))
        elems.append(np.asarray(b.elems, np.int64) + off)
        off += len(b.nodes)
    return np.concatenate(coords), np.concatenate(elems).astype(np.int32)


def contract(variant, elems, coords, u, lam, mu, f, reps=1, stream=None):
    """Launch the variant `reps` times on `stream` (a torch CUDA stream,
    default the current one).  elems (m, 3) int32, coords (n, 2) float32,
    u (n, 4, 2) float32 are contiguous CUDA tensors; f (m, 4, 6) float32
    receives the forces, f (m, 4) float32 each load case's element energy
    u' F."""
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    for t, dt in ((elems, torch.int32), (coords, torch.float32), (u, torch.float32),
                  (f, torch.float32)):
        assert t.is_cuda and t.is_contiguous() and t.dtype == dt
    m = elems.shape[0]
    assert elems.shape == (m, 3) and f.shape in ((m, N_CASES, 6), (m, N_CASES))
    mode = 0 if f.dim() == 3 else 1
    assert u.shape == (coords.shape[0], N_CASES, 2) and coords.shape[1] == 2
    rc = load().femgpu_probe_contraction(int(variant), mode, int(m), elems.data_ptr(),
                                         coords.data_ptr(), u.data_ptr(), float(lam), float(mu),
                                         f.data_ptr(), int(reps), C.c_void_p(s.cuda_stream))
    if rc != 0:
        raise RuntimeError(f"femgpu_probe_contraction: HIP error {rc}")


def mfma_layout(out, stream=None):
    """One wave of v_mfma_f32_4x4x1f32 with A = lane, B = 100 + lane:
    out (256,) float32 CUDA tensor receives D[lane][register]."""
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    assert out.is_cuda and out.dtype == torch.float32 and out.numel() == 256
    rc = load().femgpu_probe_mfma_layout(out.data_ptr(), C.c_void_p(s.cuda_stream))
    if rc != 0:
        raise RuntimeError(f"femgpu_probe_mfma_layout: HIP error {rc}")
