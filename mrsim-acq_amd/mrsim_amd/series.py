"""A whole 2D series on the GPU: acquire_volumes (src/kspace.rs:1851-2125) with default options, as
simulate_acquisition_oversampled (:1687-1718) and simulate_acquisition_complex (:1803-1807) call it.

Every (volume, slice) unit of the series is one unit of ONE batched plan (the forward with spikes
and k-space noise on the reference's per-unit streams); under acceleration a second plan reads the
GRAPPA reference scan of every slice, one kernel is fitted per slice on the GPU and applied to all
of the slice's volumes; one reconstruction launch covers every unit; the epilogue adds the
image-space noise and writes the entry point's layout.  The calling contract is the reference's:

- images per compartment in the layout (x + snx*(y + sny*z))*n_volumes + g, gathered per slice with
  stride n_volumes (:1988-1996);
- slice_seed = ((g*0x100_0001 + z)*0x9E37) ^ receiver_seed (:2005-2009);
- the object phase phase_slice(model, shot, ...) (:1156-1180) with the shot of PrepPhase::shot keyed
  on (g, z, excitation_seed) when both the model and the volume carry a prep drive (:1997-2003);
- the GRAPPA reference scan: volume 0's images (or the given calibration), no shot phase, no eddy
  or prep drive, no acceleration or spikes, seed (z*0x9E37) ^ receiver_seed ^ 0x4752_5043_414C
  (:1924-1950, 1655-1657).

The object phase is built on the host in FP64 (a closed form per sim voxel, like the reference's);
everything after it runs on the GPU.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import abi
from .gpu import (Group, MrsError, Plan, Recon, Slice, Unit, grappa_apply, grappa_fit, ring_coil_maps, sampling_mask,
                  series_epilogue)

SEED_SALT_GRAPPA_CAL = 0x4752_5043_414C
_M64 = (1 << 64) - 1


# ----------------------------------------------------------------------------- object phase (src/phase.rs)

@dataclass
class PrepPhase:
    """PrepPhase (src/phase.rs:93-99): effective q scale and the per-shot motion SDs."""
    c_q: float
    sigma_dx: float
    sigma_rot: float


@dataclass
class PhaseModel:
    """PhaseModel (src/phase.rs:129-137): global + BackgroundPhase (coefficients of 1, x, y, z, x^2,
    y^2, z^2, xy, xz, yz in voxel units from the FOV centre) + an optional PrepPhase."""
    global_phase: float = 0.0
    coeffs: Tuple[float, ...] = (0.0,) * 10
    prep: Optional[PrepPhase] = None

    @classmethod
    def hbcd_like(cls) -> "PhaseModel":
        """PhaseModel::hbcd_like (src/phase.rs:188-198)."""
        return cls(0.0, (0.0, 0.13, 0.10, 0.02, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0), PrepPhase(0.0425, 1.0, 2.0e-3))

    @classmethod
    def none(cls) -> "PhaseModel":
        """PhaseModel::none (src/phase.rs:201-208)."""
        return cls(0.0, (0.0,) * 10, PrepPhase(0.0, 0.0, 0.0))


def _splitmix(state: int):
    state = (state + 0x9E3779B97F4A7C15) & _M64
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
    return state, z ^ (z >> 31)


def _gauss(state: int):
    state, a = _splitmix(state)
    state, b = _splitmix(state)
    u1 = ((a >> 11) + 0.5) / float(1 << 53)
    u2 = ((b >> 11) + 0.5) / float(1 << 53)
    return state, math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)


ZERO_SHOT = np.zeros(9)


def prep_shot(prep: PrepPhase, magnitude: float, direction, volume: int, slice_group: int, seed: int) -> np.ndarray:
    """PrepPhase::shot (src/phase.rs:108-125): q_eff[3], dx[3], rot[3]."""
    d = [float(v) for v in direction]
    n = math.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])
    if magnitude <= 0.0 or n < 1e-12:
        return ZERO_SHOT.copy()
    s = prep.c_q * math.sqrt(magnitude) / n
    out = np.zeros(9)
    out[:3] = [d[0] * s, d[1] * s, d[2] * s]
    st = (seed ^ ((volume * 0x9E3779B97F4A7C15) & _M64) ^ ((slice_group * 0xBF58476D1CE4E5B9) & _M64)) & _M64
    for i in range(3):
        st, g = _gauss(st)
        out[3 + i] = g * prep.sigma_dx
    for i in range(3):
        st, g = _gauss(st)
        out[6 + i] = g * prep.sigma_rot
    return out


def phase_slice(model: PhaseModel, shot, snx: int, sny: int, o: int, z: int, nz: int) -> np.ndarray:
    """phase_slice (src/kspace.rs:1156-1180): the model's phase at every sim voxel of slice z, (P,) f64,
    in the reference's operation order (BackgroundPhase::at, ShotPhase::at)."""
    sxs, sys_ = snx / 2.0, sny / 2.0
    off = (o - 1.0) / 2.0
    y, x = np.mgrid[0:sny, 0:snx].astype(np.float64)
    r0 = (x - sxs - off) / o
    r1 = (y - sys_ - off) / o
    r2 = np.full_like(r0, z - nz / 2.0)
    c = model.coeffs
    bg = (c[0] + c[1] * r0 + c[2] * r1 + c[3] * r2 + c[4] * r0 * r0 + c[5] * r1 * r1 + c[6] * r2 * r2
          + c[7] * r0 * r1 + c[8] * r0 * r2 + c[9] * r1 * r2)
    q, dx, rot = shot[:3], shot[3:6], shot[6:9]
    cr0 = rot[1] * r2 - rot[2] * r1
    cr1 = rot[2] * r0 - rot[0] * r2
    cr2 = rot[0] * r1 - rot[1] * r0
    sh = q[0] * (dx[0] + cr0) + q[1] * (dx[1] + cr1) + q[2] * (dx[2] + cr2)
    return (model.global_phase + bg + sh).reshape(-1)


# ----------------------------------------------------------------------------- the series

def _refuse(msg):
    raise ValueError(msg)


class SeriesAcquisition:
    """acquire_volumes on the GPU for one series.

    sim_dims (snx, sny, nz), acq_dims (nx, ny, nz); images: per compartment, (snx*sny*nz*n_volumes,)
    float32 in the layout (x + snx*(y + sny*z))*n_volumes + g (numpy or a CUDA tensor); t2, t_inhom:
    per compartment a float (ms) or a (snx*sny*nz,) map; fmap: (snx*sny*nz,) Hz; eddy_drive,
    prep_drive: per volume None or the drive; phase: PhaseModel; noise_sigma: None or (nx*ny*nz,) SD
    per component; grappa_calibration: None (volume 0 and the receiver seed) or (images, seed).
    `acquire()` runs the whole series and returns device tensors in the entry point's layout.
    """

    def __init__(self, sim_dims, acq_dims, n_volumes: int, images, t2, fmap, acq: abi.Acquisition, t_inhom=None,
                 eddy_drive=None, prep_drive=None, phase: Optional[PhaseModel] = None, seed: int = 0,
                 excitation_seed: Optional[int] = None, noise_sigma=None, grappa_calibration=None,
                 precision: int = abi.PREC_SPLIT_BF16):
        import torch
        snx, sny, nz = (int(v) for v in sim_dims)
        nx, ny, nzo = (int(v) for v in acq_dims)
        # the reference's refusals (src/kspace.rs:1863-1898)
        if nz != nzo:
            _refuse("slice count must match; z is never oversampled")
        if snx % nx or sny % ny:
            _refuse("sim grid must be an integer multiple of the acquired matrix")
        o = snx // nx
        if o != sny // ny:
            _refuse("oversampling must match on both axes")
        nv = int(n_volumes)
        P, nsim = snx * sny, snx * sny * nz
        Q = len(images)
        if not 1 <= Q <= abi.MRS_MAX_COMP:
            _refuse(f"1..{abi.MRS_MAX_COMP} compartments are supported, got {Q}")
        if len(t2) != Q:
            _refuse(f"t2 has {len(t2)} entries for {Q} compartments")
        if t_inhom is not None and len(t_inhom) != Q:
            _refuse(f"t_inhom has {len(t_inhom)} entries for {Q} compartments")
        eddy_drive = list(eddy_drive) if eddy_drive is not None else [None] * nv
        prep_drive = list(prep_drive) if prep_drive is not None else [None] * nv
        if len(eddy_drive) != nv:
            _refuse(f"eddy_drive has {len(eddy_drive)} entries for {nv} volumes")
        if len(prep_drive) != nv:
            _refuse(f"prep_drive has {len(prep_drive)} entries for {nv} volumes")
        self.dev = torch.device("cuda")
        dev = self.dev

        def up(a, dt=torch.float32):
            t = a if isinstance(a, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(a))
            return t.to(device=dev, dtype=dt).contiguous()

        imgs = [up(im) for im in images]
        for im in imgs:
            if im.numel() != nsim * nv:
                _refuse("compartment image is not on the simulation grid")
        fm = up(fmap)
        if fm.numel() != nsim:
            _refuse("fieldmap is not on the simulation grid")

        def maps(vals, name):
            """Per-slice (Q, P) maps, or None when every compartment is uniform (then the values)."""
            if vals is None:
                return None, None
            if all(isinstance(v, (int, float)) for v in vals):
                return None, [float(v) for v in vals]
            out = torch.empty((nz, Q, P), dtype=torch.float32, device=dev)
            for q, v in enumerate(vals):
                if isinstance(v, (int, float)):
                    out[:, q].fill_(float(v))
                else:
                    m = up(v)
                    if m.numel() != nsim:
                        _refuse(f"{name} map {q} is not on the simulation grid")
                    out[:, q] = m.view(nz, P)
            return out, None

        t2m, t2u = maps(t2, "T2")
        tim, tiu = maps(t_inhom, "T2'")
        self.nx, self.ny, self.nz, self.nv, self.Q, self.o = nx, ny, nz, nv, Q, o
        self.snx, self.sny = snx, sny
        self.acq = acq.copy()
        self.seed = seed & _M64
        xseed = (seed if excitation_seed is None else excitation_seed) & _M64
        phase = phase if phase is not None else PhaseModel.none()
        C = max(int(acq.n_coils), 1)
        self.C = C
        # sim-grid coil sensitivities (coil_sensitivity at the acquired voxel of each sim voxel)
        self.coil = up(ring_coil_maps(C, snx, sny, nx, ny), torch.complex64)
        # images gathered per (g, z): (nv, nz, Q, P)
        self.images = torch.stack([im.view(nz, P, nv).permute(2, 0, 1) for im in imgs], dim=2).contiguous()
        self.t2m, self.tim = t2m, tim
        self.fm = fm.view(nz, P)
        self._keep = []

        def slice_desc(z, phi, eddy):
            ph = torch.from_numpy(phi).to(dev)
            self._keep.append(ph)
            kw = dict(phase0=ph, z=z, nz=nz, eddy_drive=None if eddy is None else tuple(float(e) for e in eddy))
            if t2m is not None:
                kw["t2"] = t2m[z]
            else:
                kw["t2_uniform"] = tuple(t2u) + (100.0,) * (abi.MRS_MAX_COMP - Q)
            if tim is not None:
                kw["t_inhom"] = tim[z]
            elif tiu is not None:
                kw["t_inhom_uniform"] = tuple(tiu) + (50.0,) * (abi.MRS_MAX_COMP - Q)
            return Slice(0, self.fm[z], self.coil, **kw)

        # slices: one per distinct (z, eddy drive, shot); the plan shares coil planes between them
        slices, key_of, units = [], {}, []
        self.kspace = torch.zeros((nv, nz, C, ny, nx), dtype=torch.complex64, device=dev)
        for g in range(nv):
            for z in range(nz):
                shot = ZERO_SHOT
                if phase.prep is not None and prep_drive[g] is not None:
                    mag, direction = prep_drive[g]
                    shot = prep_shot(phase.prep, float(mag), direction, g, z, xseed)
                eddy = eddy_drive[g]
                key = (z, None if eddy is None else tuple(eddy), tuple(shot))
                if key not in key_of:
                    key_of[key] = len(slices)
                    slices.append(slice_desc(z, phase_slice(phase, shot, snx, sny, o, z, nz), eddy))
                sseed = ((((g * 0x1000001) & _M64) + z) * 0x9E37 & _M64) ^ self.seed
                units.append(Unit(key_of[key], self.images[g, z], self.kspace[g, z], slice_seed=sseed))
        group = Group(snx, sny, nx, ny, C, Q, self.acq)
        self.plan = Plan([group], slices, units, precision)
        self.n_slices = len(slices)

        # GRAPPA reference scan: one unit per slice
        self.cal_plan = None
        if int(acq.accel) > 1 and nv > 0:
            if grappa_calibration is not None:
                cal_images, cal_seed = grappa_calibration
                cal = [up(im) for im in cal_images]
                if len(cal) != Q or any(c.numel() != nsim for c in cal):
                    _refuse("the GRAPPA calibration images are not one simulation-grid volume per compartment")
                cal = torch.stack([c.view(nz, P) for c in cal], dim=1).contiguous()
            else:
                cal, cal_seed = self.images[0], self.seed
            self.cal_images = cal
            self.acq_cal = acq.copy(accel=1, n_spikes=0)
            cal_slices = [slice_desc(z, phase_slice(phase, ZERO_SHOT, snx, sny, o, z, nz), None) for z in range(nz)]
            self.cal_k = torch.zeros((nz, C, ny, nx), dtype=torch.complex64, device=dev)
            cal_units = [Unit(z, cal[z], self.cal_k[z], slice_seed=((z * 0x9E37) & _M64) ^ (cal_seed & _M64)
                              ^ SEED_SALT_GRAPPA_CAL) for z in range(nz)]
            self.cal_plan = Plan([Group(snx, sny, nx, ny, C, Q, self.acq_cal)], cal_slices, cal_units, precision)
            self.cal_lines = sampling_mask(1, ny, self.acq_cal)
            self.acquired = sampling_mask(1, ny, self.acq)
        self.recon = Recon(nx, ny, C, self.acq)
        self.combined = torch.empty((nv, nz, ny, nx), dtype=torch.complex64, device=dev)
        self.noise_sigma = None
        if noise_sigma is not None:
            self.noise_sigma = up(noise_sigma)
            if self.noise_sigma.numel() != nx * ny * nz:
                _refuse("noise_sigma is not on the acquired grid")
        self.weights = None

    def calibrate(self):
        """The per-slice GRAPPA kernels (device tensors), fitted on the reference scan."""
        if self.cal_plan is None:
            return None
        self.cal_plan.forward()
        self.weights = [grappa_fit(self.cal_k[z], int(self.acq.accel), int(self.acq.acs_lines), self.cal_lines)
                        for z in range(self.nz)]
        return self.weights

    def acquire(self, magnitude_phase: bool = True):
        """Run the series: (magnitude, phase) or (re, im), (nx*ny*nz*n_volumes,) float32 device tensors
        in the layout (x + nx*(y + ny*z))*n_volumes + g."""
        if self.cal_plan is not None and self.weights is None:
            self.calibrate()
        self.plan.forward()
        if self.weights is not None:
            for z in range(self.nz):
                grappa_apply([self.kspace[g, z] for g in range(self.nv)], int(self.acq.accel), self.acquired,
                             self.weights[z])
        ks = [self.kspace[g, z] for g in range(self.nv) for z in range(self.nz)]
        ims = [self.combined[g, z] for g in range(self.nv) for z in range(self.nz)]
        self.recon.run(ks, ims)
        return series_epilogue(self.combined.view(self.nv, -1), self.noise_sigma, self.seed, magnitude_phase)

    def close(self):
        for p in (self.plan, self.cal_plan):
            if p is not None:
                p.close()
        self.recon.close()


def simulate_acquisition_oversampled(sim_dims, acq_dims, n_volumes, images, t2, fmap, t_inhom, acq, eddy_drive,
                                     prep_drive, phase, seed, noise_sigma=None):
    """simulate_acquisition_oversampled (src/kspace.rs:1687-1718) on the GPU: (magnitude, phase) as
    numpy float32 arrays in the reference's layout."""
    s = SeriesAcquisition(sim_dims, acq_dims, n_volumes, images, t2, fmap, acq, t_inhom=t_inhom, eddy_drive=eddy_drive,
                          prep_drive=prep_drive, phase=phase, seed=seed, noise_sigma=noise_sigma)
    try:
        mag, ph = s.acquire(True)
        return mag.cpu().numpy(), ph.cpu().numpy()
    finally:
        s.close()


__all__ = ["PhaseModel", "PrepPhase", "prep_shot", "phase_slice", "SeriesAcquisition",
           "simulate_acquisition_oversampled", "MrsError"]
