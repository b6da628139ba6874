"""A whole 2D series on the GPU (mrsim_amd.series.SeriesAcquisition: acquire_volumes,
src/kspace.rs:1851-2125) against the oracle's restatement (orc_acquire_volumes, pinned bit for bit to
the reference's series records) and against those records themselves.

- the two golden series (16x16x3, o = 2, 4 volumes, 3 coils, R = 2, k-space noise, hbcd_like phase)
  with the oracle's per-slice GRAPPA kernels fed to the GPU: the same linear map on both sides, so
  the complex images agree to rel-L2 <= 2e-5 (forward and reconstruction each carry split-BF16
  round-off, about 4e-6 and 6e-6);
- the same series entirely on the GPU, calibration included, against the golden records: bounded by
  3x the image error the oracle's own pipeline shows when its k-spaces carry random relative errors
  of the GPU's measured k-space size (the regularised fit amplifies them);
- a series with per-volume prep drives (shot phases), eddy drives, T2 maps, T2' and image-space noise
  (no acceleration) against the oracle, rel-L2 <= 2e-5, and the magnitude/phase epilogue against the
  oracle's f32 magnitude/phase of the same complex samples.
"""
import numpy as np
import pytest

from mrsim_amd import abi, series
from oracle import oracle as O

pytestmark = pytest.mark.gpu
D = O.SERIES_DIMS
SIM = (D["nx"] * D["o"], D["ny"] * D["o"], D["nz"])
ACQD = (D["nx"], D["ny"], D["nz"])


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _builder(images, fmap, t2, z, g, phase0, seed, eddy=None, t_inhom=None):
    """SliceInputBuilder of (volume g, slice z) of a series in the entry point's layout."""
    nv, (snx, sny, nz) = D["nv"], SIM
    P = snx * sny
    comps = [im.reshape(nz * P, nv)[z * P:(z + 1) * P, g].copy() for im in images]
    t2s = [v if isinstance(v, float) else v[z * P:(z + 1) * P] for v in t2]
    tis = None if t_inhom is None else [v if isinstance(v, float) else v[z * P:(z + 1) * P] for v in t_inhom]
    return abi.SliceInputBuilder(comps, t2s, fmap[z * P:(z + 1) * P].copy(), (snx, sny), ACQD[:2], z=z, nz=nz,
                                 t_inhom=tis, phase0=phase0, eddy_drive=eddy, slice_seed=seed)


def _oracle_kernels(images, fmap, t2, acq, seed, model, pert=None):
    """The per-slice kernels of acquire_volumes (volume 0, no shot, calibration seed), optionally with
    the calibration k-spaces perturbed."""
    out = []
    acq_cal = acq.copy(accel=1, n_spikes=0)
    lines = O.sampling_mask(1, D["ny"], acq_cal)
    for z in range(D["nz"]):
        phi = O.phase_slice(model, None, SIM[0], SIM[1], D["o"], z, D["nz"])
        b = _builder(images, fmap, t2, z, 0, phi, ((z * 0x9E37) ^ seed ^ 0x47525043414C) & ((1 << 64) - 1))
        k = O.coil_kspaces(b, acq_cal)
        if pert is not None:
            k = pert(k)
        out.append(O.grappa_fit(k, D["nx"], D["ny"], acq.accel, acq.acs_lines, lines))
    return out


def _complex(mag, ph):
    return mag.astype(np.float64) * np.exp(1j * ph.astype(np.float64))


@pytest.mark.parametrize("name", ["uniform", "mapped"])
def test_golden_series_with_the_oracles_kernels(name):
    import torch
    images, fmap, _ = O.series_inputs()
    t2, acq = {c[0]: c for c in O.series_cases()}[name][1:]
    model = O.PhaseModel.hbcd_like()
    re_o, im_o = O.acquire_volumes(SIM, ACQD, D["nv"], images, t2, fmap, None, acq, model=model,
                                   excitation_seed=42, receiver_seed=42)
    s = series.SeriesAcquisition(SIM, ACQD, D["nv"], images, t2, fmap, acq, phase=series.PhaseModel.hbcd_like(),
                                 seed=42)
    shape = (acq.accel - 1, acq.n_coils, 6 * acq.n_coils)
    s.weights = [torch.from_numpy(w.reshape(shape)).cuda() for w in _oracle_kernels(images, fmap, t2, acq, 42, model)]
    re, im = (t.cpu().numpy() for t in s.acquire(magnitude_phase=False))
    e = rel(re + 1j * im, re_o + 1j * im_o.astype(np.float64))
    print(f"series-{name}: GPU with the oracle's kernels vs oracle rel-L2 {e:.2e}", flush=True)
    assert e <= 2e-5
    s.close()


@pytest.mark.parametrize("name", ["uniform", "mapped"])
def test_golden_series_entirely_on_the_gpu(name):
    images, fmap, _ = O.series_inputs()
    t2, acq = {c[0]: c for c in O.series_cases()}[name][1:]
    gm, gp = O.series_golden(name)
    mag, ph = series.simulate_acquisition_oversampled(SIM, ACQD, D["nv"], images, t2, fmap, None, acq, None, None,
                                                      series.PhaseModel.hbcd_like(), 42)
    e = rel(_complex(mag, ph), _complex(gm, gp))
    # the oracle's pipeline with k-space errors of the GPU's size (measured on the first slice)
    model = O.PhaseModel.hbcd_like()
    phi = O.phase_slice(model, None, SIM[0], SIM[1], D["o"], 0, D["nz"])
    from mrsim_amd import gpu
    b = _builder(images, fmap, t2, 0, 1, phi, (1 * 0x1000001 * 0x9E37) ^ 42)
    ek = rel(gpu.build_coil_kspaces(b, acq.copy(noise_variance=0.0)), O.coil_kspaces(b, acq.copy(noise_variance=0.0)))
    rng = np.random.default_rng(0)
    pert = lambda k: k + ek * np.sqrt(np.mean(np.abs(k[k != 0]) ** 2)) * (k != 0) * (
        rng.standard_normal(k.shape) + 1j * rng.standard_normal(k.shape)) / np.sqrt(2)
    ker = _oracle_kernels(images, fmap, t2, acq, 42, model, pert)
    nv, nz = D["nv"], D["nz"]
    npx = D["nx"] * D["ny"]
    want = _complex(gm, gp).reshape(nz, npx, nv)
    err2, ref2 = 0.0, 0.0
    for g in range(nv):
        for z in range(nz):
            phi = O.phase_slice(model, None, SIM[0], SIM[1], D["o"], z, nz)
            b = _builder(images, fmap, t2, z, g, phi, (((g * 0x1000001) + z) * 0x9E37) ^ 42)
            img = O.reconstruct(pert(O.coil_kspaces(b, acq)), D["nx"], D["ny"], acq, ker[z])
            err2 += np.sum(np.abs(img.astype(np.complex64) - want[z, :, g]) ** 2)
            ref2 += np.sum(np.abs(want[z, :, g]) ** 2)
    e_pert = float(np.sqrt(err2 / ref2))
    print(f"series-{name}: whole GPU series vs golden {e:.2e}; oracle with k-space errors of {ek:.1e}: "
          f"{e_pert:.2e}", flush=True)
    assert e <= max(2e-5, 3.0 * e_pert)


def test_series_with_shots_eddy_maps_and_image_noise():
    images, fmap, t2map = O.series_inputs()
    acq = abi.Acquisition.default(n_coils=4, noise_variance=0.5, signal_scale=100.0, t_echo=30.0, t_line=0.5,
                                  eddy_strength=0.05, eddy_quad=0.2, n_spikes=2, spike_amplitude=3.0)
    nv = D["nv"]
    nvox = D["nx"] * D["ny"] * D["nz"]
    t2 = [t2map, 160.0]
    ti = [45.0, t2map * 0.5]
    eddy = [None, (0.6, 0.8, 0.0), None, (0.0, 0.5, 0.5)]
    prep = [(1000.0, (1.0, 0.3, 0.0)), None, (3000.0, (0.0, 0.0, 1.0)), (800.0, (1.0, 1.0, 1.0))]
    sigma = (0.02 * (np.arange(nvox) % 5)).astype(np.float32)
    om = O.PhaseModel.hbcd_like()
    re_o, im_o = O.acquire_volumes(SIM, ACQD, nv, images, t2, fmap, ti, acq, eddy_drive=eddy, prep_drive=prep,
                                   model=om, excitation_seed=77, receiver_seed=77, noise_sigma=sigma)
    s = series.SeriesAcquisition(SIM, ACQD, nv, images, t2, fmap, acq, t_inhom=ti, eddy_drive=eddy, prep_drive=prep,
                                 phase=series.PhaseModel.hbcd_like(), seed=77, noise_sigma=sigma)
    print(f"series with shots: {s.n_slices} distinct slices for {nv * D['nz']} units", flush=True)
    re, im = (t.cpu().numpy() for t in s.acquire(magnitude_phase=False))
    e = rel(re + 1j * im, re_o + 1j * im_o.astype(np.float64))
    print(f"series with shots, eddy, maps, spikes, image noise: GPU vs oracle rel-L2 {e:.2e}", flush=True)
    assert e <= 2e-5
    # magnitude and phase from the same complex samples
    mag, ph = (t.cpu().numpy() for t in s.acquire(magnitude_phase=True))
    om_, op_ = O.mag_phase(re, im)
    assert np.max(np.abs(mag - om_)) <= 2e-6 * np.abs(om_).max()
    big = om_ > 1e-3 * om_.max()
    assert np.max(np.abs(np.angle(np.exp(1j * (ph[big] - op_[big]))))) <= 1e-5
    s.close()
