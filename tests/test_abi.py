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


# ------------------------------------------------------------------ 3D gradient-echo host tables

def test_grase_block_matches_oracle(L):
    from oracle import ge3d as G
    for ny, kseg, tl in [(20, 2, 1.0), (64, 4, 0.2048), (16, 2, 0.5), (96, 4, 0.4), (7, 1, 0.3), (1, 1, 0.0001)]:
        for rev in (False, True):
            t, pol = gpu.grase_block(ny, kseg, tl, rev)
            tw, pw, _ = G.grase_block(ny, kseg, tl, rev)
            assert np.array_equal(t, tw) and np.array_equal(pol, pw), (ny, kseg, tl, rev)
    with pytest.raises(gpu.MrsError, match="69 phase-encode lines do not divide into 4 ky segments"):
        gpu.grase_block(69, 4, 0.2)


@pytest.mark.parametrize("centric", [False, True])
@pytest.mark.parametrize("ny,nz,kseg,kzs", [(16, 8, 2, 1), (16, 8, 2, 2), (16, 8, 2, 8), (64, 24, 4, 1),
                                            (96, 24, 4, 3), (12, 5, 3, 5)])
def test_ge3d_lines_match_oracle(L, centric, ny, nz, kseg, kzs):
    from oracle import ge3d as G
    a = gpu.ge3d_lines(ny, nz, kseg, 0.4, False, kzs, centric)
    b = G.ge3d_lines(ny, nz, kseg, 0.4, False, kzs, centric)
    for k in ("shot", "excitation"):
        assert np.array_equal(a[k], b[k]), k
    for k in ("n_shots", "n_exc", "e_c"):
        assert a[k] == b[k], k


def test_ge3d_line_refusals(L):
    for args, msg in [((16, 8, 2, 0.5, False, 3, False), "8 partitions do not divide into 3 kz segments"),
                      ((16, 1, 2, 0.5, False, 1, False), "a 3D gradient-echo train on 1 partition is a 2D readout"),
                      ((16, 8, 3, 0.5, False, 1, False), "16 phase-encode lines do not divide into 3 ky segments")]:
        with pytest.raises(gpu.MrsError, match=msg):
            gpu.ge3d_lines(*args)


def test_check_ge3d_timing_messages_match_oracle(L):
    # the reference's timing cases (src/readout.rs:1043-1088) and the config-4 train: the library
    # and the oracle accept the same trains and refuse the others with the same text
    from oracle import ge3d as G
    cases = [([12.0, 22.0], 40.0, 2000.0, 4), ([2.5], 40.0, 2000.0, 4), ([3.0], 40.0, 2000.0, 4),
             ([12.0, 15.5], 40.0, 2000.0, 4), ([12.0, 16.0], 40.0, 2000.0, 4), ([12.0, 32.0], 34.9, 2000.0, 4),
             ([12.0, 32.0], 35.0, 2000.0, 4), ([12.0], 40.0, 1133.9, 4), ([12.0], 40.0, 1134.0, 4),
             ([12.0], 10.0, 2000.0, 1), ([12.0], 10.0, 1013.9, 1), ([], 40.0, 2000.0, 4)]
    for tes, spacing, tr, n_exc in cases:
        want = None
        try:
            G.check_ge3d_timing(tes, spacing, 2.0, 16, 2, 0.5, False, n_exc, 1000.0, tr)
        except G.TimingError as e:
            want = str(e)
        got = None
        try:
            gpu.check_ge3d_timing(tes, spacing, 2.0, 16, 2, 0.5, False, n_exc, 1000.0, tr)
        except gpu.MrsError as e:
            got = str(e).split(": ", 1)[1]
        assert got == want, (tes, spacing, tr)
    for ny in (64, 96):
        gpu.check_ge3d_timing([12.0, 24.0, 36.0, 48.0], 60.0, 2.0, ny, 4, 0.4, False, 24, 10.0, 1500.0)


def test_zstage_refusals_before_any_device_work(L):
    import torch
    z = abi.ZSetDesc()
    z.inp, z.out = 16, 32          # never dereferenced: validation refuses first
    z.nz, z.n_comp, z.n_coils, z.ny, z.nx, z.ky_segments = 65, 1, 1, 4, 4, 1
    err = ct.create_string_buffer(256)
    assert L.mrs_zstage(ct.byref(z), 1, 0, None, err, 256) == -4 and b"1..64" in err.value
    z.nz, z.ky_segments = 4, 3
    assert L.mrs_zstage(ct.byref(z), 1, 0, None, err, 256) == -1 and b"ky segments" in err.value
    z.ky_segments, z.out = 1, 16
    assert L.mrs_zstage(ct.byref(z), 1, 0, None, err, 256) == -1 and b"not in place" in err.value
    z.out, z.n_comp, z.comp_stride = 32, 2, 10
    assert L.mrs_zstage(ct.byref(z), 1, 0, None, err, 256) == -1 and b"overlap" in err.value


def test_recon_refusals_before_any_device_work(L):
    acq = abi.Acquisition.default(n_coils=4, accel=2)
    h = ct.c_void_p()
    err = ct.create_string_buffer(256)
    k = np.zeros(4 * 16 * 16, dtype=np.complex128)
    out = np.zeros(16 * 16, dtype=np.complex128)
    P = ct.POINTER(abi.C64)
    assert L.mrs_reconstruct_coils(k.ctypes.data_as(P), 4, 16, 16, ct.byref(acq), None, out.ctypes.data_as(P),
                                   err, 256) == -1
    assert b"an accelerated slice needs its GRAPPA kernel" in err.value
    assert L.mrs_grappa_fit(None, 4, 16, 16, 2, 8, None, None, None, err, 256) == -1
    acq = abi.Acquisition.default(n_coils=4)
    assert L.mrs_recon_create(16, 16, 3, ct.byref(acq), None, ct.byref(h), err, 256) == -1
    assert b"one k-space per coil" in err.value
    assert L.mrs_recon_create(200, 16, 4, ct.byref(acq), None, ct.byref(h), err, 256) == -4


def test_nufft_refusals_before_any_device_work(L):
    # the NUFFT comparator's domain (src/kspace.rs:775-792) and the NUFFT's band (src/nufft.rs:41-42)
    comps, fmap, phase0, t2m, tim = O.golden_inputs(1)
    nx, ny, o = O.GOLDEN_SHAPES[1]
    err = ct.create_string_buffer(256)
    out = np.zeros(4 * nx * ny, dtype=np.complex128)
    P = ct.POINTER(abi.C64)

    def run(b, acq):
        return L.mrs_build_coil_kspaces_nufft(ct.byref(b.inp), ct.byref(acq), out.ctypes.data_as(P), err, 256)

    acq = abi.Acquisition.default(n_coils=4)
    mapped = abi.SliceInputBuilder(list(comps), [t2m, 100.0, t2m], fmap, (nx * o, ny * o), (nx, ny))
    assert run(mapped, acq) == -1 and b"uniform relaxation" in err.value
    eddy = abi.SliceInputBuilder(list(comps), [70.0, 100.0, 2000.0], fmap, (nx * o, ny * o), (nx, ny),
                                 eddy_drive=(0.3, -0.8, 0.5))
    assert run(eddy, acq.copy(eddy_strength=3.0)) == -1 and b"eddy polynomial" in err.value
    thin = abi.SliceInputBuilder([np.ones(16 * 2, dtype=np.float32)], [70.0], np.zeros(32, dtype=np.float32), (16, 2),
                                 (16, 2))
    assert run(thin, abi.Acquisition.default()) == -1 and b"at least 3 phase-encode lines" in err.value
    x = np.zeros(64)
    assert L.mrs_nufft1(x.ctypes.data, 1, x.ctypes.data, 4, 1, 1, 16, -9, 16, 1e-12, x.ctypes.data, None, err,
                        256) == -1
    assert b"Nyquist" in err.value
