"""SEG-Y output (femgpu_write_segy) against the SEG-Y file the reference
holds (configs/test.sgy, written by its convertToSEGY / CSegyOut path): the
same traces and sample interval must give the same bytes."""
import os
import struct

import numpy as np

from conftest import GOLDEN
from cudafemproject_amd._lib import write_segy

REF = os.path.join(GOLDEN, "segy", "test.sgy")


def read_segy(path):
    b = open(path, "rb").read()
    ntr, ns = struct.unpack(">h", b[3212:3214])[0], struct.unpack(">h", b[3220:3222])[0]
    assert struct.unpack(">h", b[3224:3226])[0] == 5          # IEEE float
    tr = np.zeros((ntr, ns), np.float32)
    for t in range(ntr):
        o = 3600 + t * (240 + 4 * ns) + 240
        tr[t] = np.frombuffer(b[o:o + 4 * ns], ">f4")
    return tr, b


def test_byte_identical_to_reference_file(tmp_path):
    tr, ref = read_segy(REF)
    assert tr.shape == (3, 5)
    out = tmp_path / "out.sgy"
    write_segy(str(out), tr, 0.045739, 0)      # interval 45739 us as in the file
    assert out.read_bytes() == ref


def test_layout_and_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    tr = rng.standard_normal((7, 33)).astype(np.float32)
    out = tmp_path / "r.sgy"
    write_segy(str(out), tr, 2e-4, 3)
    back, b = read_segy(str(out))
    assert len(b) == 3600 + 7 * (240 + 4 * 33)
    np.testing.assert_array_equal(back, tr)
    assert struct.unpack(">h", b[3216:3218])[0] == 200
    assert struct.unpack(">i", b[3600 + 8:3600 + 12])[0] == 3
    assert b[:3200] == bytes(3200)
