"""The oracle's series restatement (acquire_volumes, src/kspace.rs:1851-2125; phase_slice :1156-1180;
PhaseModel / PrepPhase, src/phase.rs) against the reference's two series golden records and the
closed forms of the reference's own phase-model tests (src/phase.rs:217-317).

- series-uniform and series-mapped (tests/golden_baseline.rs:91-124: 16x16x3, o = 2, 4 volumes,
  3 coils, R = 2 GRAPPA with per-slice kernels fitted on volume 0, k-space noise, hbcd_like phase):
  magnitude and phase bit for bit;
- planted mistakes (no through-plane background term, calibration on the wrong volume's seed
  stream, the object phase dropped) each move the bits;
- PrepPhase::shot: the guards (zero magnitude, zero direction), a non-unit direction, sign reversal
  with the gradient, sqrt(b) scaling, rotation linear in position, per-volume draws;
- image-space noise (noise_sigma): the residual's SD per component, independence across volumes,
  magnitude and phase sharing one realization.
"""
import numpy as np
import pytest

from oracle import oracle as O

D = O.SERIES_DIMS


@pytest.mark.parametrize("name", ["uniform", "mapped"])
def test_series_match_the_golden_record_bit_for_bit(name):
    mag, ph = O.series(name)
    gm, gp = O.series_golden(name)
    assert mag.size == D["nx"] * D["ny"] * D["nz"] * D["nv"]
    assert np.array_equal(mag.view(np.uint32), gm.view(np.uint32))
    assert np.array_equal(ph.view(np.uint32), gp.view(np.uint32))


def _series_with(model=None, receiver_seed=42, t2=None):
    images, fmap, _ = O.series_inputs()
    name_t2, acq = O.series_cases()[0][1:]
    sim = (D["nx"] * D["o"], D["ny"] * D["o"], D["nz"])
    re, im = O.acquire_volumes(sim, (D["nx"], D["ny"], D["nz"]), D["nv"], images, t2 or name_t2, fmap, None, acq,
                               model=model if model is not None else O.PhaseModel.hbcd_like(),
                               excitation_seed=42, receiver_seed=receiver_seed)
    return O.mag_phase(re, im)


def test_planted_mistakes_in_the_series_are_caught():
    gm, gp = O.series_golden("uniform")
    no_z = O.PhaseModel.hbcd_like()
    no_z.coeffs[3] = 0.0                       # the background's through-plane term dropped
    flat = O.PhaseModel.hbcd_like()
    for i in range(10):
        flat.coeffs[i] = 0.0                   # no object phase at all
    for what, (m, p) in [("no z term", _series_with(no_z)), ("no object phase", _series_with(flat)),
                         ("another receiver seed", _series_with(receiver_seed=46))]:
        moved = np.count_nonzero(m.view(np.uint32) != gm.view(np.uint32)) + np.count_nonzero(
            p.view(np.uint32) != gp.view(np.uint32))
        assert moved > 100, what
    # the noise streams are Rng((slice_seed ^ coil * 0x9E37_79B9) | 1) (src/kspace.rs:1031): bit 0 of
    # the receiver seed never reaches them, so seeds 42 and 43 give the same series (no spikes here)
    m, p = _series_with(receiver_seed=43)
    assert np.array_equal(m.view(np.uint32), gm.view(np.uint32))
    assert np.array_equal(p.view(np.uint32), gp.view(np.uint32))


def _shot_phase(shot, n=9):
    """ShotPhase::at on a grid: phase_slice with a zero background, r = ((x - n/2), (y - n/2), z - 1)."""
    m = O.PhaseModel()
    return O.phase_slice(m, shot, n, n, 1, 0, 2).reshape(n, n)


def test_prep_shot_guards_and_scalings():
    m = O.PhaseModel.hbcd_like()
    a = O.prep_shot(m, 1000.0, [0.0, 0.0, 3.0], 0, 0, 1)
    b = O.prep_shot(m, 1000.0, [0.0, 0.0, 1.0], 0, 0, 1)
    assert np.allclose(a[:3], b[:3], atol=1e-12, rtol=0)          # the direction is normalised inside
    assert np.array_equal(O.prep_shot(m, 1000.0, [0.0, 0.0, 0.0], 0, 0, 1), np.zeros(9))
    assert np.array_equal(O.prep_shot(m, 0.0, [0.0, 0.0, 1.0], 0, 0, 1), np.zeros(9))
    assert np.array_equal(_shot_phase(O.prep_shot(m, 0.0, [1.0, 0.0, 0.0], 3, 1, 42)), np.zeros((9, 9)))
    # sign reverses with the gradient; phase scales as sqrt(b)
    p1 = _shot_phase(O.prep_shot(m, 1000.0, [0.0, 1.0, 0.0], 3, 1, 42))
    p2 = _shot_phase(O.prep_shot(m, 1000.0, [0.0, -1.0, 0.0], 3, 1, 42))
    assert np.max(np.abs(p1 + p2)) < 1e-12
    p4 = _shot_phase(O.prep_shot(m, 4000.0, [0.0, 1.0, 0.0], 3, 1, 42))
    nz = np.abs(p1) > 1e-6
    assert nz.any() and np.allclose(p4[nz] / p1[nz], 2.0, rtol=1e-9, atol=0)
    # the draws are keyed on the volume
    assert not np.array_equal(O.prep_shot(m, 1000.0, [1.0, 0.0, 0.0], 3, 1, 42)[3:],
                              O.prep_shot(m, 1000.0, [1.0, 0.0, 0.0], 4, 1, 42)[3:])


def test_shot_phase_is_linear_in_position_and_background_gradient():
    m = O.PhaseModel()
    m.has_prep, m.c_q, m.sigma_dx, m.sigma_rot = 1, 1e-3, 0.0, 1e-3
    s = O.prep_shot(m, 2000.0, [1.0, 0.0, 0.0], 1, 0, 7)
    p = _shot_phase(s)
    # linear field: every interior point is the mean of its neighbours on both axes
    assert np.max(np.abs(p[1:-1, 1:-1] - 0.5 * (p[:-2, 1:-1] + p[2:, 1:-1]))) < 1e-12
    assert np.max(np.abs(p[1:-1, 1:-1] - 0.5 * (p[1:-1, :-2] + p[1:-1, 2:]))) < 1e-12
    # hbcd_like's in-plane background gradient is the measured 0.164 rad/voxel (phase.rs:303-307)
    h = O.PhaseModel.hbcd_like()
    bg = O.phase_slice(h, None, 8, 8, 1, 1, 2).reshape(8, 8)
    g = np.hypot(bg[0, 1] - bg[0, 0], bg[1, 0] - bg[0, 0])
    assert abs(g - 0.164) < 0.005
    # oversampling puts the sim voxels at the acquired voxel's sub-positions: o = 2 halves the step
    bg2 = O.phase_slice(h, None, 16, 16, 2, 1, 2).reshape(16, 16)
    assert abs((bg2[0, 1] - bg2[0, 0]) - 0.5 * (bg[0, 1] - bg[0, 0])) < 1e-12


def test_image_noise_statistics_and_shared_realisation():
    images, fmap, _ = O.series_inputs()
    _, acq = O.series_cases()[0][1:]
    acq = acq.copy(noise_variance=0.0, accel=1)
    sim = (D["nx"] * D["o"], D["ny"] * D["o"], D["nz"])
    nvox = D["nx"] * D["ny"] * D["nz"]
    sigma = np.full(nvox, 0.25, dtype=np.float32)
    args = (sim, (D["nx"], D["ny"], D["nz"]), D["nv"], images, [80.0, 160.0], fmap, None, acq)
    kw = dict(model=O.PhaseModel.hbcd_like(), excitation_seed=9, receiver_seed=9)
    re0, im0 = O.acquire_volumes(*args, **kw)
    re1, im1 = O.acquire_volumes(*args, noise_sigma=sigma, **kw)
    res = np.stack([re1.astype(np.float64) - re0, im1.astype(np.float64) - im0])      # (2, nvox*nv)
    # SD per component within 5 standard errors of sigma
    n = res.size
    assert abs(res.std() - 0.25) < 5 * 0.25 / np.sqrt(2 * n)
    # volumes draw independent noise: residual correlation between volumes ~ 0
    r = res[0].reshape(nvox, D["nv"])
    c = np.corrcoef(r[:, 0], r[:, 1])[0, 1]
    assert abs(c) < 5 / np.sqrt(nvox)
    # the same call is bit-reproducible, and magnitude/phase come from that one complex sample
    re2, im2 = O.acquire_volumes(*args, noise_sigma=sigma, **kw)
    assert np.array_equal(re1, re2) and np.array_equal(im1, im2)
    mag, ph = O.mag_phase(re1, im1)
    assert np.allclose(mag * np.cos(ph), re1, atol=1e-4 * np.abs(mag).max())
    assert np.allclose(mag * np.sin(ph), im1, atol=1e-4 * np.abs(mag).max())
