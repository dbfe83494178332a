"""Packed window cells: the host decoder against an independent encoder.

Window cells reach the host as 12-byte records (key and value in one 96-bit
integer, values past the field in a side list).  tests/native/cell_pack_harness.cpp
encodes random cells with 128-bit arithmetic at key widths from 1 to 60 bits,
with values at the edges of the field and of the escape threshold, and decodes
them with the engine's cell_view (keys, values, lower_bound, slices).  The
device kernel's encoding of the same format is checked end to end, bit for
bit against the oracle, by the GPU parity tests.
"""
import os

import pytest

from conftest import ROOT, pipe

HOST = os.path.join(ROOT, "maboss", "gpu_amd", "csrc", "host")


@pytest.mark.skipif(not os.path.isdir("/opt/rocm/include"), reason="needs the HIP headers")
def test_packed_cells_round_trip(native_bin):
    exe = native_bin("cell_pack_harness",
                     extra=[os.path.join(HOST, "hostmem.cpp"), f"-I{HOST}", "-I/opt/rocm/include",
                            "-D__HIP_PLATFORM_AMD__", "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"])
    lines = pipe(exe, "")
    widths = {int(l.split()[1]): (int(l.split()[2]), int(l.split()[3])) for l in lines if l.startswith("ok")}
    assert sorted(widths) == [1, 7, 12, 24, 31, 32, 33, 45, 57, 59, 60]
    # no side list while the field holds any 64-bit value; one from 32 key bits on
    assert all(widths[k][1] == 0 for k in (1, 7, 12, 24, 31))
    assert all(widths[k][1] > 0 for k in (32, 33, 45, 57, 59, 60))
