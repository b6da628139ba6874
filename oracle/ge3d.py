"""TEST INFRASTRUCTURE: FP64 oracle of the 3D gradient-echo train's forward (numpy over the C oracle).

Only tests/, `__graft_entry__.smoke()` and bench.py's CPU-baseline leg may import this module, and only
as the checker.  Each function restates the reference function it cites:

  grase_block          src/readout.rs:206-221
  kz_order             src/readout.rs:170-189
  ge3d_lines           src/readout.rs:428-462
  check_ge3d_timing    src/readout.rs:465-501
  zkernel              src/kspace3d.rs:90-100
  ge_line_weight       src/kspace3d.rs:607-611
  ge3d_kspace          GePlan::{compartments,kspace}, src/kspace3d.rs:643-726 (no shot images)

The per-partition 2D forward is the C oracle's (oracle.forward_direct), itself pinned against the
reference's golden fixtures.  Pins of this module (tests/test_oracle_ge3d.py): the reference's own
3D-GE and readout tests (worked timing numbers, line tables, the timing refusals and boundaries, the
"class" and "voxel" relaxation modes agreeing, the constant-weight z round trip of its test
reference, partition-varying weights blurring), the z kernel against numpy's FFT, and closed forms
(identical partitions fill only p = nz/2; one nonzero partition gives a plane wave in p).
"""
from __future__ import annotations

import numpy as np

from . import oracle as O


class TimingError(ValueError):
    pass


def grase_block(ny, ky_segments, t_line_ms, reverse_phase=False):
    """(t_ms[ny], polarity[ny], epi), src/readout.rs:206-221."""
    if ky_segments == 0 or ny % ky_segments:
        raise TimingError(f"{ny} phase-encode lines do not divide into {ky_segments} ky segments")
    epi = ny // ky_segments
    t = np.zeros(ny)
    pol = np.zeros(ny, dtype=np.int8)
    for ky in range(ny):
        j = ky // ky_segments
        a = j if reverse_phase else epi - 1 - j
        t[ky] = (a - (epi - 1.0) / 2.0) * t_line_ms
        pol[ky] = -1 if a % 2 == 1 else 1
    return t, pol, epi


def kz_order(nz, centric):
    """src/readout.rs:170-189."""
    if not centric:
        return list(range(nz))
    c = nz // 2
    out = [c]
    d = 1
    while len(out) < nz:
        if c + d < nz:
            out.append(c + d)
        if d <= c and len(out) < nz:
            out.append(c - d)
        d += 1
    return out


def ge3d_lines(ny, nz, ky_segments, t_line_ms, reverse_phase, kz_segments, centric):
    """dict(shot, excitation [nz*ny], n_shots, n_exc, e_c), src/readout.rs:428-462."""
    if nz < 2:
        raise TimingError(f"a 3D gradient-echo train on {nz} partition is a 2D readout")
    _, _, epi = grase_block(ny, ky_segments, t_line_ms, reverse_phase)
    if kz_segments == 0 or nz % kz_segments:
        raise TimingError(f"{nz} partitions do not divide into {kz_segments} kz segments")
    n_exc = nz // kz_segments
    order = kz_order(nz, centric)
    shot = np.full(ny * nz, -1, dtype=np.int32)
    exc = np.full(ny * nz, -1, dtype=np.int32)
    for sy in range(ky_segments):
        for sz in range(kz_segments):
            for j in range(n_exc):
                p = order[sz + kz_segments * j]
                for i in range(epi):
                    at = p * ny + sy + ky_segments * i
                    assert shot[at] < 0, "line read twice"
                    shot[at] = sy * kz_segments + sz
                    exc[at] = j
    assert (shot >= 0).all(), "a line was never read"
    e_c = order.index(nz // 2) // kz_segments
    return {"shot": shot, "excitation": exc, "n_shots": ky_segments * kz_segments, "n_exc": n_exc, "e_c": e_c}


def check_ge3d_timing(tes, exc_spacing_ms, excitation_time_ms, ny, ky_segments, t_line_ms, reverse_phase, n_exc,
                      t_exc_ms, tr_ms):
    """Raises TimingError with the reference's text, src/readout.rs:465-501."""
    if len(tes) == 0:
        raise TimingError("a 3D gradient-echo train needs at least one echo time")
    t, _, epi = grase_block(ny, ky_segments, t_line_ms, reverse_phase)
    half = float(np.max(np.abs(t))) + t_line_ms / 2.0
    pulse = excitation_time_ms / 2.0
    if tes[0] - half < pulse - 1e-9:
        raise TimingError(f"echo 1 at {tes[0]:.4f} ms: its {epi}-line block starts {tes[0] - half:.4f} ms after the "
                          f"excitation, inside the excitation pulse (half of {excitation_time_ms:.3f} ms); a later "
                          "echo time or a shorter block would fit")
    for e in range(1, len(tes)):
        if tes[e] - half < tes[e - 1] + half - 1e-9:
            raise TimingError(f"echo {e} at {tes[e - 1]:.4f} ms and echo {e + 1} at {tes[e]:.4f} ms: their blocks of "
                              f"{half:.4f} ms each side overlap")
    end = tes[-1] + half
    if n_exc > 1 and end > exc_spacing_ms - pulse + 1e-9:
        raise TimingError(f"the last echo's block ends {end:.4f} ms after its excitation, after the next excitation "
                          f"pulse starts at {exc_spacing_ms - pulse:.4f} ms (excitation spacing {exc_spacing_ms:.4f} "
                          f"ms, pulse {excitation_time_ms:.3f} ms): the spacing must be at least {end + pulse:.4f} ms")
    last = t_exc_ms + (n_exc - 1.0) * exc_spacing_ms + end
    if last > tr_ms + 1e-9:
        raise TimingError(f"the train ends at {last:.4f} ms ({n_exc} excitations {exc_spacing_ms:.4f} ms apart from "
                          f"{t_exc_ms:.4f} ms), after the {tr_ms:.4f} ms repetition")


def zkernel(nz):
    """zk[p, z] = exp(+i 2 pi (p - nz/2)(z - nz/2) / nz), src/kspace3d.rs:90-100."""
    h = nz // 2
    p = np.arange(nz, dtype=np.float64)[:, None]
    z = np.arange(nz, dtype=np.float64)[None, :]
    th = 2.0 * np.pi * (p - h) * (z - h) / nz
    return np.cos(th) + 1j * np.sin(th)


def ge_line_weight(trf_ms, t2_ms, t2p_ms):
    """exp(-trf (1/T2 + 1/T2')), infinite times meaning no decay, src/kspace3d.rs:607-611."""
    d2 = 0.0 if np.isinf(t2_ms) else trf_ms / t2_ms
    dp = 0.0 if np.isinf(t2p_ms) else trf_ms / t2p_ms
    return np.exp(-d2 - dp)


def within_echo_timing(ny, ky_segments, t_line_ms, reverse_phase, te):
    """GePlan::compartments' LineTiming: t, trf = TE + t, tread = t, polarity (src/kspace3d.rs:654-660)."""
    t, pol, _ = grase_block(ny, ky_segments, t_line_ms, reverse_phase)
    return t, te + t, t.copy(), pol


def _ge_acquisition(acq, te, relax):
    a = acq.copy()
    a.echo = 1
    a.t_echo = te
    a.noise_variance = 0.0
    a.n_spikes = 0
    a.do_relaxation = 1 if relax else 0
    return a


def compartment_kspace(snx, sny, nx, ny, nz, image, t2, t_inhom, fmap, phase0, coil_maps, acq, te, ky_segments,
                       t_line_ms, mode=None):
    """GePlan::compartments for one compartment (src/kspace3d.rs:650-697): per partition z the 2D
    gradient-echo forward at TE with the within-echo timing, then the z-DFT, then the class line weight.

    image (nz, P) real; t2, t_inhom: float (uniform) or (nz, P) maps; fmap (nz, P); phase0 (nz, P) or
    None; coil_maps (nz, C, P) complex.  mode: None = the reference's choice ("class" when T2 and T2'
    are uniform, else "voxel"); forcing "voxel" on uniform inputs evaluates the decay per voxel.
    Returns (nz, C, ny, nx) complex128 indexed [p, coil, ky, kx].
    """
    C = coil_maps.shape[1]
    P = snx * sny
    ti = acq.t_inhom if t_inhom is None else t_inhom
    uniform = np.isscalar(t2) and np.isscalar(ti)
    if mode is None:
        mode = "class" if uniform else "voxel"
    if mode == "class" and not uniform:
        raise ValueError("class mode needs uniform T2 and T2'")
    voxel = mode == "voxel" and acq.do_relaxation
    a = _ge_acquisition(acq, te, voxel)
    timing = within_echo_timing(ny, ky_segments, t_line_ms, acq.reverse_phase, te)
    k2 = np.zeros((nz, C, ny, nx), dtype=np.complex128)
    for z in range(nz):
        if not np.any(image[z]):
            continue    # the 2D forward is linear in the image: a zero partition contributes zero
        t2z = None if np.isscalar(t2) else np.asarray(t2[z], dtype=np.float32)
        tiz = None if np.isscalar(ti) else np.asarray(ti[z], dtype=np.float32)
        k2[z] = O.forward_direct(snx, sny, nx, ny, C, 1, a, np.asarray(image[z], dtype=np.float32).reshape(P),
                                 np.asarray(fmap[z], dtype=np.float32), np.asarray(coil_maps[z]),
                                 phase0=None if phase0 is None else phase0[z], t2=t2z,
                                 t2_uniform=[float(t2)] if np.isscalar(t2) else None,
                                 t_inhom=tiz, t_inhom_uniform=[float(ti)] if np.isscalar(ti) else None,
                                 z=z, nz=nz, timing=timing)
    k3 = np.einsum("pz,zcyx->pcyx", zkernel(nz), k2)
    if mode == "class" and acq.do_relaxation:
        w = np.array([ge_line_weight(te + timing[0][ky], float(t2), float(ti)) for ky in range(ny)])
        k3 = k3 * w[None, None, :, None]
    return k3


def ge3d_kspace(snx, sny, nx, ny, nz, images, t2, t_inhom, fmap, phase0, coil_maps, acq, te, ky_segments, t_line_ms,
                weights, mode=None):
    """GePlan::kspace (src/kspace3d.rs:702-726) without shot images: per acquired line, the
    compartment sum weighted by w[(p*ky_segments + sy)*ncomp + c].  images: list (None = unused);
    weights (nz, ky_segments, ncomp).  Returns (nz, C, ny, nx) complex128."""
    ncomp = len(images)
    weights = np.asarray(weights, dtype=np.float64).reshape(nz, ky_segments, ncomp)
    C = coil_maps.shape[1]
    mask = O.sampling_mask(nx, ny, acq).reshape(ny, nx)[:, 0]
    sy = np.arange(ny) % ky_segments
    k = np.zeros((nz, C, ny, nx), dtype=np.complex128)
    for c in range(ncomp):
        if images[c] is None:
            continue
        ti = None if t_inhom is None else t_inhom[c]
        kc = compartment_kspace(snx, sny, nx, ny, nz, images[c], t2[c], ti, fmap, phase0, coil_maps, acq, te,
                                ky_segments, t_line_ms, mode)
        w = weights[:, sy, c]                                   # (nz, ny)
        k += kc * (w * mask[None, :])[:, None, :, None]
    return k


def ring_maps(nc, snx, sny, nx, ny):
    """coil_sensitivity (src/kspace.rs:226-237) on the simulation grid with the forward's half-cell
    registration, (nc, snx*sny) complex64: the maps the reference's 2D call applies to coil q of nc."""
    ox, oy = snx // nx, sny // ny
    y, x = np.mgrid[0:sny, 0:snx].astype(float)
    xa, ya = (x - (ox - 1) / 2) / ox, (y - (oy - 1) / 2) / oy
    out = np.zeros((nc, snx * sny), np.complex64)
    f = O.lib().orc_coil_sensitivity
    for c in range(nc):
        v = np.vectorize(lambda a, b: f(c, nc, a, b, nx, ny))(xa, ya)
        out[c] = v.reshape(-1).astype(np.float32)
    return out


def blobs(snx, sny, nz, seed=1):
    """A smooth positive test object on (nz, snx*sny): Gaussian blobs that drift with z."""
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:sny, 0:snx].astype(float)
    out = np.zeros((nz, snx * sny))
    cs = rng.uniform(0.3, 0.7, size=(4, 2))
    amp = rng.uniform(0.5, 1.5, size=4)
    for z in range(nz):
        img = np.zeros_like(x)
        for k in range(4):
            cx = (cs[k, 0] + 0.03 * z) * snx
            cy = (cs[k, 1] - 0.02 * z) * sny
            img += amp[k] * np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2 * (0.12 * snx) ** 2))
        out[z] = img.reshape(-1)
    return out.astype(np.float32)


def varying_fmap(snx, sny, nz):
    """A smooth fieldmap (Hz) that varies in-plane and with z, (nz, snx*sny)."""
    y, x = np.mgrid[0:sny, 0:snx].astype(float)
    fx, fy = x / snx - 0.5, y / sny - 0.5
    base = 25.0 * fx - 10.0 * fy + 15.0 * np.exp(-(fx * fx + fy * fy) / 0.02)
    return np.stack([(base + 2.0 * z).reshape(-1) for z in range(nz)]).astype(np.float32)


# ---------------------------------------------------------------------------- after the forward

SEED_SALT_3D = 0x334452454144            # src/kspace3d.rs:41
SEED_SALT_GRAPPA_CAL = 0x47525043414C    # src/kspace.rs (the GRAPPA calibration stream salt)
_M64 = (1 << 64) - 1


def echo_salt(e):
    """src/kspace.rs:2130-2132."""
    return (e * 0xD1B54A32D192ED03) & _M64


def _rng(seed):
    import ctypes as ct
    return ct.c_uint64(seed & _M64)


def partitions_to_slices(k3, mask_rows, g, seed, acq):
    """src/kspace3d.rs:235-279 for every coil of one volume: k3 (nz, C, ny, nx) complex128 is changed
    in place by the spikes and the noise; returns the slices (nz, C, ny, nx), the inverse z-DFT / nz."""
    import ctypes as ct
    L = O.lib()
    nz, C, ny, nx = k3.shape
    sigma = np.sqrt(acq.noise_variance / (nx * ny))
    mask = np.repeat(np.asarray(mask_rows, dtype=bool), nx)
    acquired = np.nonzero(mask)[0]
    for p in range(nz):
        pseed = (((g * 0x1000001 + p) & _M64) * 0x9E37 & _M64) ^ seed ^ SEED_SALT_3D
        for q in range(C):
            part = k3[p, q].reshape(-1)
            if acq.n_spikes > 0:
                mags = np.abs(part)
                peak = part[int(np.argmax(mags))] if mags.max() > 0 else 0.0
                st = _rng(pseed ^ 0xA5A5123456789ABC ^ ((q * 0xC2B2AE3D27D4EB4F) & _M64))
                for _ in range(int(acq.n_spikes)):
                    v = L.orc_rng_next(ct.byref(st))
                    part[acquired[v % len(acquired)]] = peak * acq.spike_amplitude
            if acq.noise_variance > 0.0:
                st = _rng((pseed ^ ((q * 0x9E3779B9) & _M64)) | 1)
                for i in acquired:
                    re = L.orc_rng_gauss(ct.byref(st)) * sigma
                    im = L.orc_rng_gauss(ct.byref(st)) * sigma
                    part[i] += re + 1j * im
    zk = zkernel(nz)
    return np.einsum("pz,pcyx->zcyx", zk.conj(), k3) / nz


def reconstruct_slices(slices, acq, kernels=None):
    """src/kspace3d.rs:283-296: every slice through the 2D coil reconstruction -> (nz, ny*nx)."""
    nz, C, ny, nx = slices.shape
    out = np.zeros((nz, ny * nx), dtype=np.complex128)
    for z in range(nz):
        w = None if kernels is None else kernels[z]
        out[z] = O.reconstruct(slices[z].reshape(C, -1).copy(), nx, ny, acq, w)
    return out
