"""Synthetic inputs of the benchmark workloads, seeded from the reference's SplitMix64 generator
(src/kspace.rs:60-77), evaluated counter-wise and vectorised: draw n of a stream seeded s is
mix(s + n * 0x9E3779B97F4A7C15).

Workloads (sizes are the benchmark definitions):
  slice_128      one slice, 128^2 -> 64^2, 1 coil, 1 compartment, T2/T2' maps, spin echo
  batch_128      30 slices x 20 volumes x 8 coils x 3 compartments, 128^2 -> 64^2
  pf_256         24 slices x 16 coils, 256^2 -> 128^2, partial Fourier 6/8 (97 lines), Hann recon window
  grad_128       one slice, 128^2 -> 64^2, 8 coils, 3 compartments, complex64 images (adjoint checks)
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import abi

GAMMA = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)


def splitmix(seed: int, n: int, start: int = 1) -> np.ndarray:
    """Draws start..start+n-1 of Rng(seed) as uint64."""
    with np.errstate(over="ignore"):
        k = np.arange(start, start + n, dtype=np.uint64)
        z = np.uint64(seed) + k * GAMMA
        z = (z ^ (z >> np.uint64(30))) * M1
        z = (z ^ (z >> np.uint64(27))) * M2
        return z ^ (z >> np.uint64(31))


class Stream:
    """A SplitMix64 stream (Rng) with vectorised draws."""

    def __init__(self, seed: int):
        self.seed = seed & 0xFFFFFFFFFFFFFFFF
        self.pos = 1

    def u64(self, n):
        v = splitmix(self.seed, n, self.pos)
        self.pos += n
        return v

    def unit(self, n=None):
        m = 1 if n is None else n
        v = ((self.u64(m) >> np.uint64(11)).astype(np.float64) + 0.5) / float(1 << 53)
        return v[0] if n is None else v

    def uniform(self, lo, hi, n=None):
        return lo + (hi - lo) * self.unit(n)

    def gauss(self, n):
        u = self.unit(2 * n).reshape(n, 2)
        return np.sqrt(-2.0 * np.log(u[:, 0])) * np.cos(2 * np.pi * u[:, 1])


def blob_image(st: Stream, snx: int, sny: int, nblobs: int = 8) -> np.ndarray:
    """Sum of Gaussian blobs, amplitude U(0,1), sigma U(4,20) px, centre U(16, n-16), phase U(-pi,pi)."""
    y, x = np.mgrid[0:sny, 0:snx].astype(np.float64)
    m = np.zeros((sny, snx), dtype=np.complex128)
    for _ in range(nblobs):
        a, sg = st.uniform(0, 1), st.uniform(4, 20)
        cx, cy = st.uniform(16, snx - 16), st.uniform(16, sny - 16)
        ph = st.uniform(-np.pi, np.pi)
        m += a * np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2 * sg * sg)) * np.exp(1j * ph)
    return m.reshape(-1)


def fieldmap(st: Stream, snx: int, sny: int, n: int = 4) -> np.ndarray:
    """Sum of Gaussians of amplitude +-U(20,120) Hz, width U(10,40) px."""
    y, x = np.mgrid[0:sny, 0:snx].astype(np.float64)
    f = np.zeros((sny, snx))
    for _ in range(n):
        amp = st.uniform(20, 120) * (1 if st.unit() < 0.5 else -1)
        s = st.uniform(10, 40)
        cx, cy = st.uniform(16, snx - 16), st.uniform(16, sny - 16)
        f += amp * np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2 * s * s))
    return f.reshape(-1).astype(np.float32)


def gaussian_ring_coils(n: int, snx: int, sny: int, sigma: float, radius: float, centre=None, z_gain=None):
    """Complex Gaussian sensitivities on a ring, phase 2 pi c / n: (n, P) complex64."""
    cx0, cy0 = centre if centre is not None else (snx / 2, sny / 2)
    y, x = np.mgrid[0:sny, 0:snx].astype(np.float64)
    out = np.zeros((n, snx * sny), dtype=np.complex64)
    for c in range(n):
        ang = 2 * np.pi * c / n
        px, py = cx0 + radius * np.cos(ang), cy0 + radius * np.sin(ang)
        g = np.exp(-((x - px) ** 2 + (y - py) ** 2) / (2 * sigma * sigma)) * np.exp(1j * 2 * np.pi * c / n)
        if z_gain is not None:
            g = g * z_gain[c]
        out[c] = g.reshape(-1)
    return out


@dataclass
class SliceData:
    fmap: np.ndarray
    phase0: np.ndarray
    t2: np.ndarray          # (Q, P)
    t_inhom: np.ndarray     # (Q, P)
    coil_maps: np.ndarray   # (C, P)
    mag: np.ndarray         # (P,)
    weights: np.ndarray     # (Q, P) Dirichlet


@dataclass
class Workload:
    name: str
    snx: int
    sny: int
    nx: int
    ny: int
    n_coils: int
    n_comp: int
    acq: abi.Acquisition
    slices: list = field(default_factory=list)   # SliceData
    units: list = field(default_factory=list)    # (slice index, images (Q,P) float32 or complex64)
    complex_images: bool = False
    recon_window: int = abi.WINDOW_NONE
    timing: object = None   # a caller's line-timing table (t_ms, trf_ms, tread_ms, polarity), or None


def _slice(st: Stream, snx, sny, Q, C, coil_sigma, coil_radius):
    P = snx * sny
    m = blob_image(st, snx, sny)
    fm = fieldmap(st, snx, sny)
    t2 = st.uniform(40, 120, Q * P).reshape(Q, P).astype(np.float32)
    ti = st.uniform(20, 80, Q * P).reshape(Q, P).astype(np.float32)
    e = -np.log(st.unit(Q * P).reshape(Q, P))
    w = e / e.sum(axis=0, keepdims=True)
    coils = gaussian_ring_coils(C, snx, sny, coil_sigma, coil_radius) if C > 1 else np.ones((1, P), np.complex64)
    return SliceData(fmap=fm, phase0=np.angle(m), t2=t2, t_inhom=ti, coil_maps=coils, mag=np.abs(m), weights=w)


def _volume_images(st: Stream, sd: SliceData, g: int):
    Q, P = sd.weights.shape
    pert = 1.0 + 0.02 * (2.0 * st.unit(Q * P).reshape(Q, P) - 1.0)
    return ((1.0 + 0.01 * g) * sd.weights * sd.mag[None, :] * pert).astype(np.float32)


def make(name: str, n_slices=None, n_volumes=None, seed_base=0x5EED0000) -> Workload:
    if name == "slice_128":
        acq = abi.Acquisition.default(t_line=0.5, t_echo=30.0, signal_scale=1.0, n_coils=1)
        w = Workload(name, 128, 128, 64, 64, 1, 1, acq)
        st = Stream(seed_base + 0)
        sd = _slice(st, 128, 128, 1, 1, 60.0, 90.0)
        sd.weights[:] = 1.0
        w.slices.append(sd)
        w.units.append((0, (sd.mag[None, :]).astype(np.float32)))
        return w
    if name == "batch_128":
        ns = 30 if n_slices is None else n_slices
        nv = 20 if n_volumes is None else n_volumes
        acq = abi.Acquisition.default(t_line=0.5, t_echo=30.0, signal_scale=1.0, n_coils=8)
        w = Workload(name, 128, 128, 64, 64, 8, 3, acq)
        for s in range(ns):
            st = Stream(seed_base + 1 + 7919 * s)
            sd = _slice(st, 128, 128, 3, 8, 60.0, 90.0)
            w.slices.append(sd)
            for g in range(nv):
                w.units.append((s, _volume_images(st, sd, g)))
        return w
    if name == "pf_256":
        ns = 24 if n_slices is None else n_slices
        acq = abi.Acquisition.default(t_line=0.5, t_echo=50.0, signal_scale=1.0, n_coils=16,
                                      partial_fourier=0.75, window=abi.WINDOW_HANN)
        w = Workload(name, 256, 256, 128, 128, 16, 1, acq, recon_window=abi.WINDOW_HANN)
        for s in range(ns):
            st = Stream(seed_base + 2 + 7919 * s)
            sd = _slice(st, 256, 256, 1, 16, 120.0, 180.0)
            sd.weights[:] = 1.0
            w.slices.append(sd)
            w.units.append((s, sd.mag[None, :].astype(np.float32)))
        return w
    if name in ("tile_96", "tile_192"):
        # x-stage blocks of 48 and 96 kx: 48 kx in one block of three 16-row kx tiles (96^2 -> 48^2,
        # 8 coils, 2 compartments) and 96 kx in one block of six tiles and eight waves (192^2 -> 96^2,
        # 4 coils)
        sn, n, C, Q = (96, 48, 8, 2) if name == "tile_96" else (192, 96, 4, 1)
        acq = abi.Acquisition.default(t_line=0.5, t_echo=30.0, signal_scale=1.0, n_coils=C)
        w = Workload(name, sn, sn, n, n, C, Q, acq)
        st = Stream(seed_base + (5 if name == "tile_96" else 6))
        sd = _slice(st, sn, sn, Q, C, 0.45 * sn, 0.7 * sn)
        w.slices.append(sd)
        w.units.append((0, _volume_images(st, sd, 0)))
        return w
    if name == "grad_128":
        acq = abi.Acquisition.default(t_line=0.5, t_echo=30.0, signal_scale=1.0, n_coils=8)
        w = Workload(name, 128, 128, 64, 64, 8, 3, acq, complex_images=True)
        st = Stream(seed_base + 3)
        sd = _slice(st, 128, 128, 3, 8, 60.0, 90.0)
        w.slices.append(sd)
        img = (sd.weights * (sd.mag * np.exp(1j * sd.phase0))[None, :]).astype(np.complex64)
        w.units.append((0, img))
        return w
    raise KeyError(name)


def to_plan(w: Workload, precision=abi.PREC_SPLIT_BF16, device="cuda"):
    """Upload a workload and build its Plan.  Returns (plan, outputs, keep-alive tensors)."""
    import torch

    from .gpu import Group, Plan, Slice, Unit
    g = Group(w.snx, w.sny, w.nx, w.ny, w.n_coils, w.n_comp, w.acq, complex_images=w.complex_images,
              timing=w.timing)
    keep, slices, units, outs = [], [], [], []
    for sd in w.slices:
        t = dict(fmap=torch.from_numpy(sd.fmap).to(device),
                 coil_maps=torch.from_numpy(np.ascontiguousarray(sd.coil_maps)).to(device),
                 phase0=None if w.complex_images else torch.from_numpy(sd.phase0).to(device),
                 t2=torch.from_numpy(np.ascontiguousarray(sd.t2)).to(device),
                 t_inhom=torch.from_numpy(np.ascontiguousarray(sd.t_inhom)).to(device))
        keep.append(t)
        slices.append(Slice(0, t["fmap"], t["coil_maps"], phase0=t["phase0"], t2=t["t2"], t_inhom=t["t_inhom"]))
    for si, img in w.units:
        ti = torch.from_numpy(np.ascontiguousarray(img)).to(device)
        out = torch.zeros((w.n_coils, w.ny, w.nx), dtype=torch.complex64, device=device)
        keep.append(ti)
        outs.append(out)
        units.append(Unit(si, ti, out))
    plan = Plan([g], slices, units, precision)
    return plan, outs, keep


# ---------------------------------------------------------------------------- 3D gradient-echo train

@dataclass
class Ge3dVolumeData:
    snx: int
    sny: int
    nx: int
    ny: int
    nz: int
    images: list            # per compartment (nz, P) float32
    t2: list                # per compartment (nz, P) float32 map
    t_inhom: list           # per compartment float (uniform, ms)
    fmap: np.ndarray        # (nz, P) float32
    coil_maps: np.ndarray   # (nz, C, P) complex64
    weights: np.ndarray     # (nz, ky_segments, Q) float64


@dataclass
class Ge3dWorkload:
    name: str
    acq: abi.Acquisition
    echo_times_ms: tuple
    ky_segments: int
    t_line_ms: float
    kz_segments: int
    centric: bool
    exc_spacing_ms: float
    excitation_time_ms: float
    t_exc_ms: float
    tr_ms: float
    n_coils: int
    volumes: list = field(default_factory=list)


def _blobs3d(st: Stream, snx, sny, nz, nblobs=8):
    """Real positive object: Gaussian blobs in (x, y, z), sigma U(0.06, 0.18) of the field of view."""
    y, x = np.mgrid[0:sny, 0:snx].astype(np.float64)
    out = np.zeros((nz, sny * snx))
    for _ in range(nblobs):
        a, s = st.uniform(0.3, 1.0), st.uniform(0.06, 0.18)
        cx, cy, cz = st.uniform(0.25, 0.75) * snx, st.uniform(0.25, 0.75) * sny, st.uniform(0.2, 0.8) * nz
        sz = st.uniform(0.15, 0.5) * nz
        plane = np.exp(-((x - cx) ** 2 / (2 * (s * snx) ** 2) + (y - cy) ** 2 / (2 * (s * sny) ** 2))).reshape(-1)
        zg = np.exp(-((np.arange(nz) - cz) ** 2) / (2 * sz * sz))
        out += a * zg[:, None] * plane[None, :]
    return out.astype(np.float32)


def _ge3d_volume(st: Stream, snx, sny, nx, ny, nz, C, Q, kseg, lines, exc_spacing, seed_gain):
    P = snx * sny
    images = [_blobs3d(st, snx, sny, nz) for _ in range(Q)]
    t2 = [st.uniform(40, 120, nz * P).reshape(nz, P).astype(np.float32) for _ in range(Q)]
    fm2 = fieldmap(st, snx, sny).astype(np.float64)
    fmap = np.stack([fm2 * (1.0 - 0.3 * z / nz) + 2.0 * z for z in range(nz)]).astype(np.float32)
    # two rings of C/2 coils stacked along z, Gaussian in z around nz/4 and 3nz/4
    per_ring = max(C // 2, 1)
    rings = max(C // per_ring, 1)
    plane = np.concatenate([gaussian_ring_coils(per_ring, snx, sny, 0.45 * snx, 0.7 * snx) for _ in range(rings)])
    zc = np.array([(r + 0.5) * nz / rings - 0.5 for r in range(rings)])
    zz = np.arange(nz, dtype=np.float64)
    gain = np.exp(-((zz[:, None] - zc[None, :]) ** 2) / (2 * (nz / 3.0) ** 2))      # (nz, rings)
    zgain = np.repeat(gain, per_ring, axis=1)[:, :C]
    coil_maps = (zgain[:, :, None] * plane[None, :C, :]).astype(np.complex64)
    # per-excitation weights: approach to the steady state over the shot's excitations,
    # w = M_ss + (1 - M_ss) exp(-n TR_exc / T1*), times a per-shot gain
    m_ss = st.uniform(0.2, 0.4)
    t1s = 1100.0
    gains = 1.0 + 0.02 * Stream(seed_gain).gauss(int(lines["n_shots"]))
    w = np.zeros((nz, kseg, Q))
    for p in range(nz):
        for sy in range(kseg):
            n = lines["excitation"][p * ny + sy]
            sh = lines["shot"][p * ny + sy]
            for q in range(Q):
                w[p, sy, q] = (m_ss + (1.0 - m_ss) * np.exp(-n * exc_spacing / t1s)) * gains[sh] * (1.0 - 0.1 * q)
    return Ge3dVolumeData(snx, sny, nx, ny, nz, images, t2, [45.0] * Q, fmap, coil_maps, w)


def make_ge3d(name: str = "ge3d", seed_base=0x5EED3D00) -> Ge3dWorkload:
    """3D gradient-echo trains.

    ge3d        two volumes, 128^2 -> 64^2 and 192^2 -> 96^2, 24 partitions each, 32 coils (two
                rings of 16 along z), 1 compartment, echoes at 12/24/36/48 ms, 4 ky segments of
                0.4 ms lines, 2 kz segments in centric order, 60 ms excitation spacing
    ge3d_small  the same train on 24^2 -> 12^2 (4 partitions) and 32^2 -> 16^2 (6 partitions),
                4 coils, 2 compartments, echoes at 12/24 ms, 2 ky segments
    """
    from .gpu import check_ge3d_timing, ge3d_lines
    if name == "ge3d":
        geoms, C, Q, tes, kseg, kzs = [(128, 64, 24), (192, 96, 24)], 32, 1, (12.0, 24.0, 36.0, 48.0), 4, 2
    elif name == "ge3d_small":
        geoms, C, Q, tes, kseg, kzs = [(24, 12, 4), (32, 16, 6)], 4, 2, (12.0, 24.0), 2, 2
    else:
        raise KeyError(name)
    acq = abi.Acquisition.default(echo=abi.ECHO_GRADIENT, t_line=0.4, signal_scale=1.0, n_coils=C, t_inhom=45.0)
    w = Ge3dWorkload(name, acq, tes, kseg, 0.4, kzs, True, 60.0, 2.0, 10.0, 1000.0, C)
    for i, (sn, n, nz) in enumerate(geoms):
        lines = ge3d_lines(n, nz, kseg, w.t_line_ms, False, kzs, w.centric)
        check_ge3d_timing(list(tes), w.exc_spacing_ms, w.excitation_time_ms, n, kseg, w.t_line_ms, False,
                          lines["n_exc"], w.t_exc_ms, w.tr_ms)
        st = Stream(seed_base + 101 * i)
        w.volumes.append(_ge3d_volume(st, sn, sn, n, n, nz, C, Q, kseg, lines, w.exc_spacing_ms,
                                      seed_base + 101 * i + 1))
    return w


def to_ge3d(w: Ge3dWorkload, precision=abi.PREC_SPLIT_BF16, device="cuda"):
    """Upload a 3D gradient-echo workload and build its batched forward."""
    import torch

    from .ge3d import Ge3dForward, Ge3dVolume

    def dev(a, dt=None):
        t = torch.from_numpy(np.ascontiguousarray(a))
        return (t if dt is None else t.to(dt)).to(device)

    vols = []
    for v in w.volumes:
        vols.append(Ge3dVolume(v.snx, v.sny, v.nx, v.ny, v.nz,
                               images=[dev(im.reshape(-1)) for im in v.images],
                               t2=[dev(t.reshape(-1)) for t in v.t2], fmap=dev(v.fmap.reshape(-1)),
                               coil_maps=dev(v.coil_maps), weights=dev(v.weights.reshape(-1), torch.float32),
                               t_inhom=[dev(t.reshape(-1)) if isinstance(t, np.ndarray) else t for t in v.t_inhom]))
    fw = Ge3dForward(vols, w.acq, w.echo_times_ms, w.ky_segments, w.t_line_ms, precision)
    return fw, vols
