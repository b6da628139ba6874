"""Pins the 3D gradient-echo oracle (oracle/ge3d.py) before GPU results are checked against it.

- the reference's readout tests: the worked block numbers (src/readout.rs:822-866), the line tables
  (:1010-1041) and every timing refusal and boundary (:1043-1088);
- the reference's 3D-GE tests (src/kspace3d.rs:2601-2660): the "class" (line-weight) and "voxel"
  relaxation modes agree, the constant-weight 3D k-space taken back through the inverse z-DFT is the
  weight times each partition's 2D gradient-echo forward, and weights that vary across partitions
  blur through-plane;
- the z kernel is numpy's centred inverse FFT times nz; closed forms: identical partitions fill only
  p = nz/2 (with nz times the 2D k-space), one nonzero partition z0 gives the plane wave
  e^{+i2pi(p-nz/2)(z0-nz/2)/nz} across p; compartments add with their own weights and an unused
  compartment changes nothing;
- a planted mistake (the conjugate kernel) is caught by the plane-wave closed form.
"""
import numpy as np
import pytest

from mrsim_amd import abi
from oracle import ge3d as G
from oracle import oracle as O


def _peak(a):
    return float(np.abs(a).max())


# ------------------------------------------------------------------ readout tables


def test_grase_block_worked_numbers():
    t, pol, epi = G.grase_block(20, 2, 1.0, False)
    assert abs(t[10] + 0.5) < 1e-12 and epi == 10
    t5, _, _ = G.grase_block(64, 4, 0.2048, False)
    assert abs(t5[32] + 0.1024) < 1e-12
    with pytest.raises(G.TimingError, match="69"):
        G.grase_block(69, 4, 0.2, False)
    # polarity alternates with the acquisition index, not with ky; reversal mirrors the times
    t, pol, epi = G.grase_block(16, 2, 0.5, False)
    tr, polr, _ = G.grase_block(16, 2, 0.5, True)
    for ky in range(16):
        a = epi - 1 - ky // 2
        assert pol[ky] == (-1 if a % 2 else 1)
        assert tr[ky] == -t[ky]
    assert np.allclose(sorted(set(np.round(t, 9))), (np.arange(epi) - (epi - 1) / 2) * 0.5)


@pytest.mark.parametrize("centric", [True, False])
@pytest.mark.parametrize("kz_segments", [1, 2, 8])
def test_ge3d_line_tables(centric, kz_segments):
    ny, nz = 16, 8
    t = G.ge3d_lines(ny, nz, 2, 0.5, False, kz_segments, centric)
    assert (t["n_shots"], t["n_exc"]) == (2 * kz_segments, nz // kz_segments)
    order = G.kz_order(nz, centric)
    assert sorted(order) == list(range(nz))
    for p in range(nz):
        pos = order.index(p)
        for ky in range(ny):
            sy, sz = ky % 2, pos % kz_segments
            assert (t["shot"][p * ny + ky], t["excitation"][p * ny + ky]) == (sy * kz_segments + sz, pos // kz_segments)
    assert t["e_c"] == order.index(nz // 2) // kz_segments


def test_ge3d_line_table_refusals_and_centric_first():
    assert G.ge3d_lines(16, 8, 2, 0.5, False, 1, True)["e_c"] == 0
    assert G.kz_order(5, True) == [2, 3, 1, 4, 0]
    with pytest.raises(G.TimingError, match="kz segments"):
        G.ge3d_lines(16, 8, 2, 0.5, False, 3, False)
    with pytest.raises(G.TimingError, match="2D"):
        G.ge3d_lines(16, 1, 2, 0.5, False, 1, False)
    with pytest.raises(G.TimingError, match="ky segments"):
        G.ge3d_lines(16, 8, 3, 0.5, False, 1, False)


def _check(tes, spacing, tr, n_exc=4):
    G.check_ge3d_timing(tes, spacing, 2.0, 16, 2, 0.5, False, n_exc, 1000.0, tr)


def test_ge3d_timing_refusals_and_boundaries():
    t, _, _ = G.grase_block(16, 2, 0.5, False)
    assert abs(np.abs(t).max() + 0.25 - 2.0) < 1e-12
    _check([12.0, 22.0], 40.0, 2000.0)
    with pytest.raises(G.TimingError, match="excitation pulse"):
        _check([2.5], 40.0, 2000.0)
    _check([3.0], 40.0, 2000.0)
    with pytest.raises(G.TimingError, match="overlap"):
        _check([12.0, 15.5], 40.0, 2000.0)
    _check([12.0, 16.0], 40.0, 2000.0)
    with pytest.raises(G.TimingError, match="next excitation pulse"):
        _check([12.0, 32.0], 34.9, 2000.0)
    _check([12.0, 32.0], 35.0, 2000.0)
    with pytest.raises(G.TimingError, match="repetition"):
        _check([12.0], 40.0, 1133.9)
    _check([12.0], 40.0, 1134.0)
    with pytest.raises(G.TimingError):
        _check([], 40.0, 2000.0)
    # one excitation per shot: no next pulse, only the repetition bounds it
    _check([12.0], 10.0, 2000.0, n_exc=1)
    with pytest.raises(G.TimingError, match="repetition"):
        _check([12.0], 10.0, 1013.9, n_exc=1)


# ------------------------------------------------------------------ z kernel


@pytest.mark.parametrize("nz", [2, 3, 4, 5, 6, 24, 33])
def test_zkernel_is_nz_times_the_centred_inverse_fft(nz):
    rng = np.random.default_rng(nz)
    x = rng.standard_normal(nz) + 1j * rng.standard_normal(nz)
    want = nz * np.fft.fftshift(np.fft.ifft(np.fft.ifftshift(x)))
    assert np.abs(G.zkernel(nz) @ x - want).max() <= 1e-12 * np.abs(want).max()
    zk = G.zkernel(nz)
    assert np.allclose(zk, zk.T)
    if nz % 2 == 0:   # the centred integer DFT is unitary up to nz for even nz
        assert np.allclose(zk.conj().T @ zk, nz * np.eye(nz))


# ------------------------------------------------------------------ 3D gradient-echo k-space

NX, NY, NZ, OS = 12, 12, 4, 2
SNX, SNY = NX * OS, NY * OS


def _acq(n_coils):
    return abi.Acquisition.default(echo=abi.ECHO_GRADIENT, n_coils=n_coils, signal_scale=100.0, t_inhom=45.0)


def _maps(nc, nz=NZ):
    m = G.ring_maps(nc, SNX, SNY, NX, NY)
    return np.broadcast_to(m[None], (nz, nc, SNX * SNY)).copy()


def _ge(images, t2, t_inhom, weights, acq, te, nz=NZ, kseg=1, fmap=None, maps=None, mode=None, phase0=None):
    fmap = G.varying_fmap(SNX, SNY, nz) if fmap is None else fmap
    maps = _maps(max(int(acq.n_coils), 1), nz) if maps is None else maps
    return G.ge3d_kspace(SNX, SNY, NX, NY, nz, images, t2, t_inhom, fmap, phase0, maps, acq, te, kseg, 0.4, weights,
                         mode=mode)


def _slice_2d(image, t2, ti, fmap, maps, acq, te, z, nz, kseg=1):
    """One partition's 2D gradient-echo forward with its own decay at trf = TE + t (the reference
    test's per-slice reference, src/kspace3d.rs:2540-2567)."""
    a = acq.copy(echo=abi.ECHO_GRADIENT, t_echo=te, noise_variance=0.0, n_spikes=0, do_relaxation=1)
    timing = G.within_echo_timing(NY, kseg, 0.4, False, te)
    return O.forward_direct(SNX, SNY, NX, NY, maps.shape[0], 1, a, image, fmap, maps, t2_uniform=[t2],
                            t_inhom_uniform=[ti], z=z, nz=nz, timing=timing)


@pytest.mark.parametrize("n_coils", [1, 3])
def test_class_and_voxel_modes_agree(n_coils):
    img = G.blobs(SNX, SNY, NZ)
    acq = _acq(n_coils)
    w = np.full((NZ, 1, 1), 0.7)
    t2map = np.full((NZ, SNX * SNY), 80.0, np.float32)
    for te in (12.0, 30.0):
        cls = _ge([img], [80.0], [45.0], w, acq, te)
        vox = _ge([img], [t2map], [45.0], w, acq, te)
        forced = _ge([img], [80.0], [45.0], w, acq, te, mode="voxel")
        assert np.abs(cls - vox).max() <= 1e-9 * _peak(vox)
        assert np.abs(cls - forced).max() <= 1e-9 * _peak(vox)


def test_constant_weights_invert_to_the_per_partition_2d_forward():
    img = G.blobs(SNX, SNY, NZ)
    acq = _acq(2)
    fmap = G.varying_fmap(SNX, SNY, NZ)
    maps = _maps(2)
    k3 = _ge([img], [80.0], [45.0], np.full((NZ, 1, 1), 0.7), acq, 12.0, fmap=fmap, maps=maps)
    back = np.einsum("pz,pcyx->zcyx", G.zkernel(NZ).conj(), k3) / NZ
    for z in range(NZ):
        want = 0.7 * _slice_2d(img[z], 80.0, 45.0, fmap[z], maps[z], acq, 12.0, z, NZ)
        assert np.abs(back[z] - want).max() <= 1e-9 * _peak(want)
    # the fieldmap's static phase is per voxel: its mean as one phase factor misses it
    flat = _ge([img], [80.0], [45.0], np.full((NZ, 1, 1), 0.7), acq, 12.0, fmap=np.zeros_like(fmap), maps=maps)
    rot = np.exp(1j * 2 * np.pi * fmap.astype(np.float64).mean() * 0.012)
    assert np.abs(flat * rot - k3).max() > 0.05 * _peak(k3)


def test_partition_weights_blur_through_plane():
    nz = 6
    img = G.blobs(SNX, SNY, nz)
    acq = _acq(2)
    wp = 1.0 + 0.5 * np.cos(np.arange(nz) * 1.3)
    k3 = _ge([img], [80.0], [45.0], wp[:, None, None], acq, 12.0, nz=nz)
    flat = _ge([img], [80.0], [45.0], np.ones((nz, 1, 1)), acq, 12.0, nz=nz)
    assert np.abs(k3 - wp[:, None, None, None] * flat).max() <= 1e-12 * _peak(flat)
    back_w = np.einsum("pz,pcyx->zcyx", G.zkernel(nz).conj(), k3) / nz
    back_f = np.einsum("pz,pcyx->zcyx", G.zkernel(nz).conj(), flat) / nz
    assert np.abs(back_w - back_f).max() > 0.01 * _peak(back_f)


def test_identical_partitions_fill_only_the_centre():
    nz = 4
    one = G.blobs(SNX, SNY, 1)[0]
    img = np.stack([one] * nz)
    f1 = G.varying_fmap(SNX, SNY, 1)[0]
    fmap = np.stack([f1] * nz)
    acq = _acq(2)
    maps = _maps(2, nz)
    k3 = _ge([img], [80.0], [45.0], np.full((nz, 1, 1), 0.5), acq, 24.0, nz=nz, fmap=fmap, maps=maps)
    k2 = _slice_2d(one, 80.0, 45.0, f1, maps[0], acq, 24.0, 0, nz)
    for p in range(nz):
        if p == nz // 2:
            assert np.abs(k3[p] - nz * 0.5 * k2).max() <= 1e-9 * _peak(k2) * nz
        else:
            assert np.abs(k3[p]).max() <= 1e-9 * _peak(k2) * nz


def _plane_wave_error(zkernel=None):
    nz, z0 = 5, 1
    img = np.zeros((nz, SNX * SNY), np.float32)
    img[z0] = G.blobs(SNX, SNY, 1)[0]
    fmap = G.varying_fmap(SNX, SNY, nz)
    acq = _acq(1)
    maps = _maps(1, nz)
    saved = G.zkernel
    if zkernel is not None:
        G.zkernel = zkernel
    try:
        k3 = _ge([img], [80.0], [45.0], np.ones((nz, 1, 1)), acq, 12.0, nz=nz, fmap=fmap, maps=maps)
    finally:
        G.zkernel = saved
    k2 = _slice_2d(img[z0], 80.0, 45.0, fmap[z0], maps[z0], acq, 12.0, z0, nz)
    h = nz // 2
    err = 0.0
    for p in range(nz):
        want = np.exp(2j * np.pi * (p - h) * (z0 - h) / nz) * k2
        err = max(err, np.abs(k3[p] - want).max() / _peak(k2))
    return err


def test_one_partition_is_a_plane_wave_across_p():
    assert _plane_wave_error() <= 1e-9


def test_planted_conjugate_kernel_is_caught():
    orig = G.zkernel
    assert _plane_wave_error(lambda nz: np.conj(orig(nz))) > 0.1


def test_compartments_add_with_their_weights_and_unused_ones_change_nothing():
    nz = 4
    a = G.blobs(SNX, SNY, nz, seed=2)
    b = G.blobs(SNX, SNY, nz, seed=3)
    acq = _acq(2)
    kseg = 2
    rng = np.random.default_rng(5)
    w = rng.uniform(0.2, 1.0, size=(nz, kseg, 2))
    both = _ge([a, b], [80.0, 120.0], None, w, acq, 12.0, nz=nz, kseg=kseg)
    ka = _ge([a], [80.0], None, w[:, :, :1], acq, 12.0, nz=nz, kseg=kseg)
    kb = _ge([b], [120.0], None, w[:, :, 1:], acq, 12.0, nz=nz, kseg=kseg)
    assert np.abs(both - ka - kb).max() <= 1e-12 * _peak(both)
    w3 = np.concatenate([w, np.zeros((nz, kseg, 1))], axis=2)
    three = _ge([a, b, None], [80.0, 120.0, 60.0], None, w3, acq, 12.0, nz=nz, kseg=kseg)
    assert np.array_equal(three, both)


# ------------------------------------------------------------------ after the forward


def _k3(nz=NZ, nc=2, te=12.0):
    img = G.blobs(SNX, SNY, nz)
    acq = _acq(nc)
    return acq, _ge([img], [80.0], [45.0], np.ones((nz, 1, 1)), acq, te, nz=nz)


def test_partitions_to_slices_without_noise_is_the_inverse_z_dft():
    acq, k3 = _k3()
    mask = O.sampling_mask(1, NY, acq)
    sl = G.partitions_to_slices(k3.copy(), mask, 0, 123, acq)
    back = np.einsum("pz,zcyx->pcyx", G.zkernel(NZ), sl)
    assert np.abs(back - k3).max() <= 1e-12 * np.abs(k3).max()


def test_partition_noise_statistics_and_independence():
    # noise only on acquired rows, zero mean, variance sigma^2 per component; the two echoes' streams
    # (seed ^ echo_salt(e)) are uncorrelated, echo 0's is the series seed itself (src/kspace3d.rs:2662-2705)
    nz, nc = 4, 2
    acq0, k3 = _k3(nz, nc)
    acq = acq0.copy(noise_variance=0.5, accel=2, acs_lines=6)
    mask = O.sampling_mask(1, NY, acq)
    sigma = np.sqrt(0.5 / (NX * NY))
    r = []
    for e in (0, 1):
        k = k3.copy()
        G.partitions_to_slices(k, mask, 3, 77 ^ G.echo_salt(e), acq)
        r.append((k - k3))
    assert G.echo_salt(0) == 0
    for d in r:
        assert np.all(d[:, :, ~mask, :] == 0)
        v = d[:, :, mask, :].reshape(-1)
        assert abs(np.mean(v.real)) < 4 * sigma / np.sqrt(v.size)
        tol = 5 * np.sqrt(2.0 / v.size)   # five standard errors of a variance estimate
        assert abs(np.var(v.real) / sigma ** 2 - 1) < tol and abs(np.var(v.imag) / sigma ** 2 - 1) < tol
    a, b = r[0][:, :, mask, :].reshape(-1), r[1][:, :, mask, :].reshape(-1)
    corr = np.real(np.vdot(a, b)) / np.sqrt(np.vdot(a, a).real * np.vdot(b, b).real)
    assert abs(corr) < 3.0 / np.sqrt(2 * a.size)


def test_partition_spikes_replace_samples_with_the_scaled_peak():
    nz, nc = 4, 2
    acq0, k3 = _k3(nz, nc)
    acq = acq0.copy(n_spikes=3, spike_amplitude=2.5)
    mask = O.sampling_mask(1, NY, acq)
    k = k3.copy()
    G.partitions_to_slices(k, mask, 1, 5, acq)
    for p in range(nz):
        for q in range(nc):
            changed = np.nonzero(k[p, q] != k3[p, q])
            peak = k3[p, q].reshape(-1)[np.argmax(np.abs(k3[p, q]))]
            assert 1 <= len(changed[0]) <= 3
            assert np.all(k[p, q][changed] == peak * 2.5)
