"""CPU: the measurement library of the stacked-load-case contraction
(libfemgpu_probe.so) is built by build() and exports its two entry points;
flat_plates joins plates into one mesh with offset node ids."""
import numpy as np


def test_probe_library_exports():
    from cudafemproject_amd import contraction_probe as CP
    lib = CP.load()
    assert hasattr(lib, "femgpu_probe_contraction")
    assert hasattr(lib, "femgpu_probe_mfma_layout")


def test_flat_plates_offsets():
    from cudafemproject_amd import contraction_probe as CP
    from cudafemproject_amd.problem import plate_batch
    insts = plate_batch(3, seed=7)
    coords, elems = CP.flat_plates(insts)
    assert coords.dtype == np.float32 and elems.dtype == np.int32
    assert len(coords) == sum(len(b.nodes) for b in insts)
    assert len(elems) == sum(len(b.elems) for b in insts)
    off = 0
    k = 0
    for b in insts:
        e = elems[k:k + len(b.elems)]
        assert np.array_equal(e, np.asarray(b.elems) + off)
        assert np.allclose(coords[e], np.asarray(b.nodes, np.float32)[np.asarray(b.elems)])
        off += len(b.nodes)
        k += len(b.elems)
