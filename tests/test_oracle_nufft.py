"""The oracle's restatement of the type-1 NUFFT (orc_nufft1, src/nufft.rs:25-113) and of the `kspace`
feature's NUFFT y-stage (src/kspace.rs:778-831, 930-935), pinned by:

- the reference's own NUFFT tests: `matches_the_direct_sum` (src/nufft.rs:325-347: four
  configurations from one Rng(42) stream, max error <= 10 tol x total weight) and
  `integer_positions_reduce_to_the_dft` (:424-437, 1e-10);
- the reference's `kspace`-build golden record (tests/golden/p4_baseline/kspace, 39 slice images):
  every image bit for bit with the feature build switched on, 27 of them through the NUFFT y-stage;
- `gradient_echo_nufft_and_rotor_paths_agree` (src/kspace.rs:3681-3698): the NUFFT path (uniform T2)
  and the rotor path (a constant map of the same T2) agree to 1e-12;
- a planted mistake (the kernel width of tol 1e-3 where 1e-12 is asked) fails the direct-sum bound.
"""
import ctypes as ct
import os

import numpy as np
import pytest

from mrsim_amd import abi
from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "p4_baseline")


def _direct_sum_cases():
    st = ct.c_uint64(42)
    L = O.lib()
    unit = lambda: L.orc_rng_unit(ct.byref(st))
    out = []
    for n, k_lo, n_out, tol in [(48, -12, 24, 1e-12), (140, -35, 70, 1e-12), (16, -8, 16, 1e-12), (64, -16, 32, 1e-8)]:
        nsrc = 3 * n
        pos = np.array([(unit() - 0.5) * 3.0 * n for _ in range(nsrc)])
        re = np.array([unit() - 0.5 for _ in range(nsrc)])
        im = np.array([unit() - 0.5 for _ in range(nsrc)])
        re2 = np.array([unit() for _ in range(nsrc)])
        out.append((n, k_lo, n_out, tol, pos, [re + 1j * im, re2 + 0j]))
    return out


def test_nufft_matches_the_direct_sum():
    for n, k_lo, n_out, tol, pos, ws in _direct_sum_cases():
        got = O.nufft1(n, k_lo, n_out, tol, pos, ws)
        for q, w in enumerate(ws):
            want = O.nufft_direct(n, k_lo, n_out, pos, w)
            worst = float(np.max(np.abs(want - got[q])))
            assert worst <= 10.0 * tol * float(np.sum(np.abs(w))), (n, q, worst)


def test_planted_kernel_width_mistake_fails_the_bound():
    n, k_lo, n_out, tol, pos, ws = _direct_sum_cases()[0]
    got = O.nufft1(n, k_lo, n_out, 1e-3, pos, ws)          # width 4 where 1e-12 asks for 13
    want = O.nufft_direct(n, k_lo, n_out, pos, ws[0])
    assert float(np.max(np.abs(want - got[0]))) > 10.0 * tol * float(np.sum(np.abs(ws[0])))


def test_integer_positions_reduce_to_the_dft():
    n = 32
    pos = np.arange(n, dtype=np.float64)
    w = np.sin(0.3 * pos) + 0j
    got = O.nufft1(n, -16, 32, 1e-12, pos, [w])[0]
    want = np.array([np.sum(w * np.exp(2j * np.pi * k * pos / n)) for k in range(-16, 16)])
    assert np.max(np.abs(got.real - want.real)) < 1e-10 and np.max(np.abs(got.imag - want.imag)) < 1e-10


def test_band_outside_the_coarse_nyquist_band_is_refused():
    with pytest.raises(ValueError, match="Nyquist"):
        O.nufft1(16, -9, 16, 1e-12, np.zeros(4), [np.ones(4) + 0j])


@pytest.mark.parametrize("shape", [0, 1, 2])
def test_kspace_feature_build_matches_the_kspace_golden_record_bit_for_bit(shape):
    nx, ny, o = O.GOLDEN_SHAPES[shape]
    n_nufft = 0
    with O.kspace_feature():
        for name, acq, ed, maps in O.golden_cases():
            b = O.golden_slice_builder(shape, ed, maps)
            _, used = O.coil_kspace_nufft(b, acq, 0, max(int(acq.n_coils), 1))
            # the NUFFT y-stage runs where its conditions hold: uniform relaxation, no eddy polynomial
            assert used == (not maps and acq.eddy_strength == 0.0), name
            n_nufft += used
            got = O.simulate_slice(b, acq)
            want = np.fromfile(os.path.join(GOLDEN, "kspace", f"slice-{nx}x{ny}o{o}-{name}.bin"), dtype=np.float32)
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), name
    assert n_nufft == 9


def test_nufft_and_rotor_paths_agree_under_the_gradient_echo():
    nx = ny = 16
    comp = np.zeros(nx * ny, dtype=np.float32)
    O.lib().orc_box_hires(nx, ny, 4.0, 12.0, 4.0, 12.0, abi.fptr(comp))
    fmap = (15.0 + 5.0 * ((np.arange(nx * ny) % 7) - 3.0)).astype(np.float32)
    acq = abi.Acquisition.default(echo=abi.ECHO_GRADIENT)
    t2map = np.full(nx * ny, 100.0, dtype=np.float32)
    bu = abi.SliceInputBuilder([comp], [100.0], fmap, (nx, ny), (nx, ny), z=3, nz=9)
    bm = abi.SliceInputBuilder([comp], [t2map], fmap, (nx, ny), (nx, ny), z=3, nz=9)
    u, used = O.coil_kspace_nufft(bu, acq)
    assert used
    m = O.coil_kspace(bm, acq)
    assert np.linalg.norm(u - m) / np.linalg.norm(u) < 1e-12
