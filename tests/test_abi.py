"""CPU checks of the C-ABI library: it builds, loads, exports every declared entry point, agrees
with the ctypes struct layouts, computes the host-side protocol tables exactly as the oracle, and
refuses invalid shapes with the reference's messages (src/kspace.rs:612-619) before touching a GPU."""
import ctypes as ct
import os
import re

import numpy as np
import pytest

from mrsim_amd import abi, gpu
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def L():
    if not os.path.exists(gpu.LIB_PATH):
        from mrsim_amd import build
        build.build()
    return gpu.lib()


def _declared():
    txt = open(os.path.join(ROOT, "include", "mrsim_gpu.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(mrs_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol(L):
    names = _declared()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_struct_layouts_match(L):
    for name, cls in abi.STRUCT_SIZES.items():
        assert L.mrs_sizeof(name.encode()) == ct.sizeof(cls), name


def test_version(L):
    assert b"gfx950" in L.mrs_version()


@pytest.mark.parametrize("nx,ny", [(17, 70), (16, 40), (8, 9), (64, 64), (128, 128)])
def test_host_tables_match_oracle(L, nx, ny):
    for kw in [dict(), dict(reverse_phase=1), dict(partial_fourier=0.75), dict(partial_fourier=0.75, pf_mode=2),
               dict(partial_fourier=0.8, pf_mode=1, reverse_phase=1), dict(accel=3), dict(t_line=0.37, t_echo=12.0)]:
        acq = abi.Acquisition.default(**kw)
        a = gpu.line_timing(acq, nx, ny)
        b = O.line_timing(acq, nx, ny)
        for u, v in zip(a, b):
            assert np.array_equal(u, v), kw
        assert np.array_equal(gpu.sampling_mask(nx, ny, acq), O.sampling_mask(nx, ny, acq)), kw


def test_ring_coil_maps_match_coil_sensitivity(L):
    nc, snx, sny, nx, ny = 4, 34, 140, 17, 70
    m = gpu.ring_coil_maps(nc, snx, sny, nx, ny)
    for c in range(nc):
        for (x, y) in [(0, 0), (5, 17), (33, 139)]:
            xa, ya = (x - 0.5) / 2, (y - 0.5) / 2
            want = O.lib().orc_coil_sensitivity(c, nc, xa, ya, nx, ny)
            assert m[c, x + snx * y].real == np.float32(want)
            assert m[c, x + snx * y].imag == 0


def test_refusals_carry_the_reference_messages(L):
    comps, fmap, phase0, _, _ = O.golden_inputs(1)
    bad = abi.SliceInputBuilder([comps[0]], [70.0], fmap, (16, 40), (6, 40))
    with pytest.raises(gpu.MrsError, match="sim grid must be an integer multiple of the acquired matrix"):
        gpu.build_coil_kspace(bad, abi.Acquisition.default())
    ok = abi.SliceInputBuilder([comps[0]], [70.0], fmap, (16, 40), (16, 40))
    t = O.line_timing(abi.Acquisition.default(), 16, 39)
    with pytest.raises(gpu.MrsError, match="line-timing table must have one entry per phase-encode line"):
        gpu.build_coil_kspace(ok, abi.Acquisition.default(), timing=t)
    with pytest.raises(gpu.MrsError, match="out of range"):
        gpu.build_coil_kspace(ok, abi.Acquisition.default(), coil=4, ncoils=4)
