"""Pins the CPU FP64 oracle (test infrastructure) before any GPU result is trusted against it.

- every per-slice golden fixture of the reference (tests/golden_baseline.rs:28-89), bit for bit;
- EXPECTED_BITS_STD (src/kspace.rs:3476-3503), bit for bit;
- the rotor-path forward against the literal per-line sum (src/kspace.rs:3362-3433, 3580-3626);
- closed forms the mathematics fixes (gradient echo vs spin echo, src/kspace.rs:3634-3676;
  constant map = uniform, :3554-3575; a single k-space sample inverts to a plane wave;
  linearity; unacquired rows exactly zero);
- a planted mistake (a dropped factor in the oracle's input) is caught by the golden comparison.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O
from mrsim_amd import abi

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "p4_baseline")


def _golden(build, name):
    return np.fromfile(os.path.join(GOLDEN, build, f"{name}.bin"), dtype=np.float32)


CASES = [(s, c[0]) for s in range(3) for c in O.golden_cases()]


@pytest.mark.parametrize("shape,name", CASES, ids=[f"{O.GOLDEN_SHAPES[s][0]}x{O.GOLDEN_SHAPES[s][1]}o"
                                                     f"{O.GOLDEN_SHAPES[s][2]}-{n}" for s, n in CASES])
def test_oracle_matches_golden_slice_bit_for_bit(shape, name):
    nx, ny, o = O.GOLDEN_SHAPES[shape]
    case = {c[0]: c for c in O.golden_cases()}[name]
    _, acq, ed, maps = case
    b = O.golden_slice_builder(shape, ed, maps)
    got = O.simulate_slice(b, acq)
    want = _golden("default", f"slice-{nx}x{ny}o{o}-{name}")
    assert got.shape == want.shape
    diff = np.nonzero(got.view(np.uint32) != want.view(np.uint32))[0]
    assert diff.size == 0, f"first difference at {diff[0]}: {got[diff[0]]} vs {want[diff[0]]}"


def test_kspace_build_fixtures_agree_with_default_to_round_off():
    # The reference's two builds (twiddle table vs rustfft x-stage) differ only in round-off, so one
    # restatement pins both (the fixtures' two differing elements differ by <= 2.9e-11).
    worst = 0.0
    for f in sorted(os.listdir(os.path.join(GOLDEN, "default"))):
        a = _golden("default", f[:-4])
        b = _golden("kspace", f[:-4])
        worst = max(worst, float(np.max(np.abs(a - b))))
    assert worst < 1e-6


EXPECTED_BITS_STD = [(0xbf85ec1cb1daa8d4, 0x3fa5ffff06e41c35), (0xbf3ac84a01a68680, 0x3f817f28896f61be),
                     (0xbf83c1e814334026, 0xbfb080ac7ad7ff65), (0xbf7548454640584a, 0xbf8d9e9dd3358a7c),
                     (0x3fb6e06347299105, 0x3fc30724fd363632), (0x3fc9ee837d7455c5, 0x3fcc89859a43e2be),
                     (0x3fc5f2c7a85a66e2, 0x3fc03bf44552eb93), (0x3f8f6e71db6b4c1c, 0x3f7da1ff8d601054)]


def _box(snx, sny, x0, x1, y0, y1):
    v = np.zeros(snx * sny, dtype=np.float32)
    O.lib().orc_box_hires(snx, sny, x0, x1, y0, y1, abi.fptr(v))
    return v


def test_oracle_raw_kspace_bits_match_expected_bits_std():
    box = _box(16, 16, 4.0, 12.0, 4.0, 12.0)
    b = abi.SliceInputBuilder([box], [100.0], np.zeros(256, np.float32), (16, 16), (16, 16), z=0, nz=1,
                              eddy_drive=(0.6 * 1000.0, 0.8 * 1000.0, 0.0), slice_seed=7)
    k = O.coil_kspace(b, abi.Acquisition.default(eddy_strength=0.05))
    bits = k.view(np.uint64).reshape(-1, 2)[:8]
    for i, (re, im) in enumerate(EXPECTED_BITS_STD):
        assert (int(bits[i, 0]), int(bits[i, 1])) == (re, im), f"coefficient {i} moved"


def _literal_cases():
    full = abi.Acquisition.default(do_distortions=1, do_relaxation=1, signal_scale=100.0)
    G = abi.ECHO_GRADIENT
    return [
        ("clean", full.copy(do_distortions=0, do_relaxation=0), None, 0, 1),
        ("distortion+relaxation", full.copy(), None, 0, 1),
        ("reverse", full.copy(reverse_phase=1), None, 0, 1),
        ("ghost", full.copy(ghost_offset=0.015), None, 0, 1),
        ("eddy-poly", full.copy(eddy_strength=3.0, eddy_quad=0.4), (0.3, -0.8, 0.5), 0, 1),
        ("eddy-phase", full.copy(eddy_phase=0.2), (0.6, 0.6, 0.5), 0, 1),
        ("pf-fiberfox", full.copy(partial_fourier=0.75), None, 0, 1),
        ("pf-scanner-eddy", full.copy(partial_fourier=0.75, pf_mode=abi.PF_SCANNER, eddy_strength=3.0,
                                      eddy_quad=0.4, eddy_phase=0.1), (0.3, -0.8, 0.5), 0, 1),
        ("pf-scanner-reverse-grappa", full.copy(partial_fourier=0.8, pf_mode=abi.PF_SCANNER, reverse_phase=1,
                                                eddy_strength=2.0, accel=2, acs_lines=8, n_coils=3),
         (0.5, 0.5, 0.7), 1, 3),
        ("pf-contiguous-reverse", full.copy(partial_fourier=0.75, pf_mode=abi.PF_CONTIGUOUS, reverse_phase=1),
         None, 0, 1),
        ("grappa-coils", full.copy(accel=2, acs_lines=6, n_coils=4), None, 2, 4),
        ("grappa3", full.copy(accel=3, acs_lines=8, n_coils=4), None, 1, 4),
        ("everything", full.copy(ghost_offset=0.02, eddy_strength=2.0, eddy_quad=0.3, eddy_phase=0.1,
                                 partial_fourier=0.8, accel=2, acs_lines=8, n_coils=3), (0.5, 0.5, 0.7), 0, 3),
        ("ge-clean", full.copy(echo=G, do_distortions=0, do_relaxation=0), None, 0, 1),
        ("ge-distortion+relaxation", full.copy(echo=G), None, 0, 1),
        ("ge-everything", full.copy(echo=G, ghost_offset=0.02, eddy_strength=2.0, eddy_quad=0.3, eddy_phase=0.1,
                                    partial_fourier=0.8, accel=2, acs_lines=8, n_coils=3, reverse_phase=1),
         (0.5, 0.5, 0.7), 0, 3),
    ]


@pytest.mark.parametrize("shape", [0, 1, 2])
def test_rotor_forward_matches_literal_sum(shape):
    # src/kspace.rs:3362-3433: every effect alone and combined within 1e-10 x peak
    for name, acq, ed, coil, nc in _literal_cases():
        b = O.golden_slice_builder(shape, ed, False, slice_seed=0)
        a = O.coil_kspace(b, acq, coil, nc)
        r = O.literal_coil_kspace(b, acq, coil, nc)
        peak = np.abs(r).max()
        assert peak > 0, name
        assert np.abs(a - r).max() <= 1e-10 * peak, name


@pytest.mark.parametrize("echo", [abi.ECHO_SPIN, abi.ECHO_GRADIENT])
def test_maps_match_literal_sum(echo):
    # src/kspace.rs:3580-3626
    n = 12
    comps = [_box(n, n, 3.0, 9.0, 3.0, 9.0)]
    i = np.arange(n * n)
    fmap = (3.0 * ((i % 5).astype(np.float32) - 2.0)).astype(np.float32)
    m = (60.0 + (i % 7) * 10.0).astype(np.float32)
    tm = (30.0 + (i % 4) * 15.0).astype(np.float32)
    b = abi.SliceInputBuilder(comps, [m], fmap, (n, n), (n, n), t_inhom=[tm], slice_seed=5)
    acq = abi.Acquisition.default(echo=echo)
    got, want = O.coil_kspace(b, acq), O.literal_coil_kspace(b, acq)
    assert np.abs(got - want).max() < 1e-10 * np.abs(want).max()


def test_gradient_echo_static_phase_and_decay_from_rf():
    # src/kspace.rs:3634-3676 restated: the gradient echo is the spin echo times e^{i2pi f TE} for a
    # uniform fieldmap, and moving TE scales by exp(-dTE (1/T2 + 1/T2')).
    n = 16
    comps = [_box(2 * n, 2 * n, 9.0, 23.0, 7.0, 21.0)]
    ii = np.arange(4 * n * n)
    phase0 = 0.3 * ((ii % 11) - 5.0)
    f = 20.0
    withf = abi.SliceInputBuilder(comps, [80.0], np.full(4 * n * n, f, np.float32), (2 * n, 2 * n), (n, n),
                                  phase0=phase0, slice_seed=1)
    without = abi.SliceInputBuilder(comps, [80.0], np.zeros(4 * n * n, np.float32), (2 * n, 2 * n), (n, n),
                                    phase0=phase0, slice_seed=1)
    base = abi.Acquisition.default(do_relaxation=0, t_echo=40.0, t_line=0.5)
    spin = O.coil_kspace(withf, base)
    grad = O.coil_kspace(withf, base.copy(echo=abi.ECHO_GRADIENT))
    rot = np.exp(1j * 2 * np.pi * f * 0.040)
    peak = np.abs(spin).max()
    assert np.abs(grad - spin * rot).max() <= 1e-12 * peak
    dec = base.copy(echo=abi.ECHO_GRADIENT, do_relaxation=1, t_inhom=50.0)
    k1 = O.coil_kspace(without, dec)
    k2 = O.coil_kspace(without, dec.copy(t_echo=70.0))
    fac = np.exp(-30.0 * (1 / 80.0 + 1 / 50.0))
    assert np.abs(k2 - fac * k1).max() <= 1e-12 * np.abs(k1).max()
    sd = dec.copy(echo=abi.ECHO_SPIN)
    s1 = O.coil_kspace(without, sd)
    s2 = O.coil_kspace(without, sd.copy(t_echo=70.0))
    assert np.abs(s2 - np.exp(-30.0 / 80.0) * s1).max() <= 1e-12 * np.abs(k1).max()


def test_constant_map_reproduces_uniform():
    comps = [_box(16, 16, 4.0, 12.0, 4.0, 12.0)]
    fm = np.zeros(256, np.float32)
    u = O.coil_kspace(abi.SliceInputBuilder(comps, [100.0], fm, (16, 16), (16, 16), slice_seed=3),
                      abi.Acquisition.default())
    m = O.coil_kspace(abi.SliceInputBuilder(comps, [np.full(256, 100.0, np.float32)], fm, (16, 16), (16, 16),
                                            slice_seed=3), abi.Acquisition.default())
    assert np.linalg.norm(u - m) / np.linalg.norm(u) < 1e-12


def test_linearity_and_unacquired_rows_zero():
    comps, fmap, phase0, t2m, tim = O.golden_inputs(1)
    acq = abi.Acquisition.default(partial_fourier=0.75, accel=2)
    mk = lambda c: abi.SliceInputBuilder([c], [70.0], fmap, (16, 40), (16, 40), phase0=phase0)
    a, b2 = O.coil_kspace(mk(comps[0]), acq), O.coil_kspace(mk(comps[1]), acq)
    # the f32 sum of the inputs rounds (~6e-8 relative); the forward itself is linear to round-off
    ab = O.coil_kspace(mk(comps[0] * 2.0 + comps[1]), acq)
    assert np.abs(ab - (2 * a + b2)).max() <= 3e-7 * np.abs(ab).max()
    mask = O.sampling_mask(16, 40, acq)
    assert np.all(a[~mask] == 0)
    assert np.any(a[mask] != 0)


def test_inverse_dft_of_one_sample_is_a_plane_wave():
    nx, ny = 6, 5
    for (kx, ky) in [(0, 0), (3, 2), (5, 4)]:
        k = np.zeros(nx * ny, complex)
        k[kx + nx * ky] = 1.0
        img = O.inverse_2d(k, nx, ny).reshape(ny, nx)
        y, x = np.mgrid[0:ny, 0:nx]
        want = np.exp(-2j * np.pi * ((kx - nx // 2) / nx * (x - nx // 2) + (ky - ny // 2) / ny * (y - ny // 2)))
        assert np.abs(img - want).max() < 1e-12


def test_line_timing_is_the_snake_epi_closed_form():
    # SingleShotEpi (src/readout.rs:45-69): the train starts at the top line (bottom if reversed);
    # each line's centre sample kx = nx/2 is read at tick = line*nx + pos, pos the centre's position
    # in that line's (possibly reversed) readout.
    for nx, ny in [(17, 70), (16, 40), (8, 9), (64, 64)]:
        for rev in (0, 1):
            acq = abi.Acquisition.default(t_line=0.5, t_echo=30.0, reverse_phase=rev)
            t, trf, tread, pol = O.line_timing(acq, nx, ny)
            dt = 0.5 / nx
            for ky in range(ny):
                line = ky if rev else ny - 1 - ky
                pos = nx // 2
                if ky % 2 == 1:
                    pos = nx - 1 - nx // 2
                tick = line * nx + pos
                assert t[ky] == pytest.approx(dt * (tick + 0.5) - nx * ny * dt / 2, abs=1e-12)
                assert trf[ky] == pytest.approx(30.0 + t[ky], abs=1e-12)
                assert tread[ky] == pytest.approx(tick * dt + dt / 2, abs=1e-12)
                assert pol[ky] == (-1 if ky % 2 else 1)


def test_fiberfox_partial_fourier_keeps_line_zero():
    # PartialFourierMode::FiberfoxCompatible doc (src/kspace.rs:85-90): at ny=32, 6/8 keeps 25 lines
    m = O.sampling_mask(1, 32, abi.Acquisition.default(partial_fourier=0.75))
    assert m.sum() == 25 and m[0]
    m = O.sampling_mask(1, 128, abi.Acquisition.default(partial_fourier=0.75))
    assert m.sum() == 97


def test_planted_mistake_is_caught_by_the_golden_record():
    # Drop the T2' term (a plausible mistake) by giving an infinite T2': the fixture must differ.
    nx, ny, o = O.GOLDEN_SHAPES[2]
    _, acq, ed, maps = O.golden_cases()[1]
    b = O.golden_slice_builder(2, ed, maps)
    bad = acq.copy(t_inhom=float("inf"))
    got = O.simulate_slice(b, bad)
    want = _golden("default", f"slice-{nx}x{ny}o{o}-distortion+relaxation")
    assert np.any(got.view(np.uint32) != want.view(np.uint32))


def _ring_maps(nc, snx, sny, nx, ny):
    ox, oy = snx // nx, sny // ny
    y, x = np.mgrid[0:sny, 0:snx].astype(float)
    xa, ya = (x - (ox - 1) / 2) / ox, (y - (oy - 1) / 2) / oy
    out = np.zeros((nc, snx * sny), np.complex64)
    for c in range(nc):
        v = np.vectorize(lambda a, b: O.lib().orc_coil_sensitivity(c, nc, a, b, nx, ny))(xa, ya)
        out[c] = v.reshape(-1).astype(np.float32)
    return out


@pytest.mark.parametrize("shape", [0, 1, 2])
def test_direct_forward_matches_restatement(shape):
    # The plan-input FP64 forward (closed-form phase, explicit coil maps) against the bit-exact
    # restatement, every effect; the f32 coil maps bound the agreement at ~1e-7.
    nx, ny, o = O.GOLDEN_SHAPES[shape]
    comps, fmap, phase0, t2m, tim = O.golden_inputs(shape)
    for name, acq, ed, coil, nc in _literal_cases():
        b = O.golden_slice_builder(shape, ed, False, slice_seed=0)
        want = O.coil_kspace(b, acq, coil, nc).reshape(ny, nx)
        maps = _ring_maps(nc, nx * o, ny * o, nx, ny)[coil:coil + 1]
        got = O.forward_direct(nx * o, ny * o, nx, ny, 1, 3, acq, comps, fmap, maps, phase0=phase0,
                               t2_uniform=[70.0, 100.0, 2000.0], eddy_drive=ed, z=3, nz=9)[0]
        assert np.linalg.norm(got - want) <= 3e-7 * np.linalg.norm(want), name


def test_direct_forward_maps_complex_images_and_coil_phase():
    # T2/T2' maps, complex images: (a) an image m e^{i phi0} with phase0 = 0 equals the real image
    # with phase0; (b) a coil map i*s gives i times the map s.
    nx, ny, o = O.GOLDEN_SHAPES[1]
    comps, fmap, phase0, t2m, tim = O.golden_inputs(1)
    acq = abi.Acquisition.default(ghost_offset=0.01)
    maps = _ring_maps(1, nx, ny, nx, ny)
    t2 = np.stack([t2m, t2m * 1.5, t2m * 0.7]).astype(np.float32)
    ti = np.stack([tim, tim, tim * 2]).astype(np.float32)
    real = O.forward_direct(nx, ny, nx, ny, 1, 3, acq, comps, fmap, maps, phase0=phase0, t2=t2, t_inhom=ti)
    cimg = (comps * np.exp(1j * phase0)[None, :]).astype(np.complex64)
    cplx = O.forward_direct(nx, ny, nx, ny, 1, 3, acq, cimg, fmap, maps, t2=t2, t_inhom=ti, complex_images=True)
    assert np.linalg.norm(cplx - real) <= 1e-6 * np.linalg.norm(real)
    rot = O.forward_direct(nx, ny, nx, ny, 1, 3, acq, comps, fmap, (1j * maps).astype(np.complex64),
                           phase0=phase0, t2=t2, t_inhom=ti)
    assert np.abs(rot - 1j * real).max() <= 1e-12 * np.abs(real).max()
    # T2 map path agrees with the restatement's map path (one compartment)
    b = abi.SliceInputBuilder([comps[0]], [t2m], fmap, (nx, ny), (nx, ny), t_inhom=[tim], phase0=phase0)
    want = O.coil_kspace(b, acq).reshape(ny, nx)
    got = O.forward_direct(nx, ny, nx, ny, 1, 1, acq, comps[:1], fmap, maps, phase0=phase0, t2=t2m[None],
                           t_inhom=tim[None])[0]
    assert np.linalg.norm(got - want) <= 3e-7 * np.linalg.norm(want)


def _maps_slice(shape, scale=None):
    """A golden-shape slice with T2 and T2' maps on every compartment (the per-voxel relaxation path,
    where the reference adds log_amp in the exponent, src/kspace.rs:959-964)."""
    comps, fmap, phase0, t2m, tim = O.golden_inputs(shape)
    nx, ny, o = O.GOLDEN_SHAPES[shape]
    if scale is not None:
        comps = [(c * scale).astype(np.float32) for c in comps]
    return abi.SliceInputBuilder(list(comps), [t2m] * len(comps), fmap, (nx * o, ny * o), (nx, ny), z=3, nz=9,
                                 t_inhom=[tim] * len(comps), phase0=phase0, slice_seed=7)


@pytest.mark.parametrize("echo", [abi.ECHO_SPIN, abi.ECHO_GRADIENT])
def test_log_amp_constant_scales_the_kspace(echo):
    # log_amp = c everywhere: every voxel weight exp(c - trf/T2 - tau/T2') = e^c times the plain one
    b = _maps_slice(1)
    acq = abi.Acquisition.default(echo=echo, n_coils=2)
    nx, ny, o = O.GOLDEN_SHAPES[1]
    c = 0.37
    k0 = O.coil_kspace(b, acq, 1, 2)
    kc = O.coil_kspace(b, acq, 1, 2, log_amp=np.full(nx * o * ny * o, c))
    assert np.abs(kc - np.exp(c) * k0).max() <= 1e-12 * np.abs(k0).max()


def test_log_amp_map_equals_scaled_compartments():
    # log_amp = ln s(x, y) scales every compartment of voxel (x, y) by s; s a power of two, so the scaled
    # f32 compartments are exact and only exp's rounding separates the two
    shape = 0
    nx, ny, o = O.GOLDEN_SHAPES[shape]
    P = nx * o * ny * o
    s = 2.0 ** np.random.default_rng(4).integers(-2, 3, size=P)
    acq = abi.Acquisition.default(n_coils=3)
    k_la = O.coil_kspace(_maps_slice(shape), acq, 2, 3, log_amp=np.log(s))
    k_sc = O.coil_kspace(_maps_slice(shape, scale=s), acq, 2, 3)
    assert np.abs(k_la - k_sc).max() <= 1e-12 * np.abs(k_sc).max()


@pytest.mark.parametrize("shape,ghost", [(0, 0.0), (1, 0.015), (2, 0.015), ("128", 0.01)])
def test_fft_x_stage_matches_twiddle_x_stage(shape, ghost):
    # The `kspace` feature's x-stage (ramp, unnormalised inverse FFT, centre phase; src/kspace.rs:1077-1124)
    # is the twiddle sum of the default build (:1091-1137) reordered (the reference's own pin of the two
    # builds: fft_inverse_matches_the_direct_dft, :3437-3451).  Power-of-two and other readout lengths.
    if shape == "128":
        rng = np.random.default_rng(9)
        P = 128 * 128
        b = abi.SliceInputBuilder([rng.uniform(0, 1, P).astype(np.float32)], [rng.uniform(40, 120, P).astype(np.float32)],
                                  (40 * (rng.uniform(size=P) - 0.5)).astype(np.float32), (128, 128), (64, 64),
                                  t_inhom=[rng.uniform(20, 80, P).astype(np.float32)], phase0=rng.uniform(-3, 3, P))
    else:
        b = _maps_slice(shape)
    acq = abi.Acquisition.default(ghost_offset=ghost, n_coils=2)
    k_tw = O.coil_kspace(b, acq, 0, 2)
    with O.fft_xstage():
        k_fft = O.coil_kspace(b, acq, 0, 2)
    assert np.abs(k_fft - k_tw).max() <= 1e-12 * np.abs(k_tw).max()
    assert not np.array_equal(k_fft, k_tw) or ghost == 0.0
