"""The 3D gradient-echo train's forward on the GPU: per-partition 2D forward, then the partition
transform with per-excitation weights, for several volumes (geometries) and echoes in one plan.

Mirrors the reference's `simulate_acquisition_3d_ge_complex` up to the acquired 3D k-space
(src/kspace3d.rs:732-760 for the refusals, GePlan::{compartments,kspace} at src/kspace3d.rs:643-726
for the arithmetic): for echo e at TE_e, compartment c and partition z, the 2D forward runs with
gradient-echo formation at TE_e and the within-echo timing of the GRASE block
(t, trf = TE_e + t, tread = t, polarity; src/readout.rs:206-221), its decay evaluated per voxel;
then K3[p] = sum_c w(p, ky % ky_segments, c) sum_z zk[p][z] K2_c[z] (src/kspace3d.rs:90-100, 716-722).
Where every relaxation input is uniform the reference applies the decay as a line weight after the
z-DFT instead; the two agree exactly in real arithmetic (exp(-(TE+t)(1/T2 + 1/T2'))).

Output layout per (volume, echo): [p][coil][ky][kx] complex64 (the reference's per-coil
[p][ky][kx] with the coil axis moved outside ky).  `acquire` continues as the reference does after
the forward: per (volume, echo) the spikes and noise of partitions_to_slices on the echo's stream
(src/kspace3d.rs:235-264, seed ^ echo_salt(e)), the inverse z-DFT over 1/nz (:265-278), GRAPPA per
slice under acceleration (kernels fitted by `grappa_calibration`, src/kspace3d.rs:303-323), and the
reconstruction of every slice (reconstruct_slices, :283-296).
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Optional, Sequence

from . import abi
from .gpu import (Group, MrsError, Plan, Recon, Slice, Unit, ZPlan, ZSet, grappa_apply, grappa_fit, grase_block,
                  partition_noise, sampling_mask, zstage)

SEED_SALT_GRAPPA_CAL = 0x47525043414C


def echo_salt(e: int) -> int:
    """The receiver-noise salt of echo e, echo 0 unsalted (src/kspace.rs:2130-2132)."""
    return (e * 0xD1B54A32D192ED03) & 0xFFFFFFFFFFFFFFFF


@dataclass
class Ge3dVolume:
    """One volume of a 3D gradient-echo train (torch device tensors, planes of P = snx*sny).

    images[c]: (nz*P,) float32 or None (a compartment this volume does not use; its weights must be 0)
    t2[c], t_inhom[c]: float (uniform, ms) or (nz*P,) float32 maps; t_inhom None = acq.t_inhom
    fmap: (nz*P,) float32 Hz; phase0: (nz*P,) float64 rad or None
    coil_maps: (nz, C, P) complex64 (sensitivities may vary with the partition)
    weights: (nz*ky_segments*ncomp,) float32, w[(p*ky_segments + sy)*ncomp + c]
    """
    snx: int
    sny: int
    nx: int
    ny: int
    nz: int
    images: Sequence[object]
    t2: Sequence[object]
    fmap: object
    coil_maps: object
    weights: object
    t_inhom: Optional[Sequence[object]] = None
    phase0: object = None


def _refuse(msg):
    raise ValueError(msg)


def check_volume(v: Ge3dVolume, acq: abi.Acquisition, ky_segments: int):
    """The reference's assertions on the 3D gradient-echo inputs (src/kspace3d.rs:741-760, 787-797)."""
    import torch
    if v.snx % v.nx or v.sny % v.ny:
        _refuse("sim grid must be an integer multiple of the acquired matrix")
    if v.snx // v.nx != v.sny // v.ny:
        _refuse("oversampling must match on both axes")
    ncomp = len(v.t2)
    P = v.snx * v.sny
    nvox = P * v.nz
    if v.fmap.numel() != nvox:
        _refuse("fieldmap is not on the simulation grid")
    if v.t_inhom is not None and len(v.t_inhom) != ncomp:
        _refuse(f"t_inhom has {len(v.t_inhom)} entries for {ncomp} compartments")
    if acq.echo != abi.ECHO_GRADIENT:
        _refuse("the 3D gradient-echo train is a gradient-echo acquisition")
    if acq.eddy_strength != 0.0 or acq.eddy_quad != 0.0 or acq.eddy_phase != 0.0:
        _refuse("the eddy model is not available in 3D (an excitation-dependent eddy evolution breaks the z "
                "factorization)")
    if acq.partial_fourier < 1.0:
        _refuse("partial Fourier is not available on the 3D gradient-echo train")
    if v.nz < 2:
        _refuse(f"a 3D gradient-echo train on {v.nz} partition is a 2D readout")
    if len(v.images) != ncomp:
        _refuse(f"{len(v.images)} compartment images for {ncomp} compartments")
    for c, im in enumerate(v.images):
        if im is not None and im.numel() != nvox:
            _refuse(f"compartment {c} is not on the simulation grid")
    if v.weights.numel() != v.nz * ky_segments * ncomp:
        _refuse("weights are not (partition, ky segment, compartment)")
    w = v.weights.reshape(v.nz * ky_segments, ncomp)
    for c, im in enumerate(v.images):
        if im is None and bool(torch.any(w[:, c] != 0)):
            _refuse(f"compartment {c} has no image but a nonzero weight")
    if tuple(v.coil_maps.shape[:1]) != (v.nz,) or v.coil_maps.shape[2] != P:
        _refuse("coil maps are not (partition, coil, sim-grid plane)")


class Ge3dForward:
    """The batched forward of several 3D gradient-echo volumes at several echo times.

    One plan holds a group per (volume, echo), a slice and a unit per (volume, echo, used
    compartment, partition), and one partition-transform set per (volume, echo) that runs after the
    2D stages inside the same launch sequence (and the same HIP graph).
    """

    def __init__(self, volumes: Sequence[Ge3dVolume], acq: abi.Acquisition, echo_times_ms: Sequence[float],
                 ky_segments: int, t_line_ms: float, precision: int = abi.PREC_SPLIT_BF16):
        import torch
        if len(echo_times_ms) == 0:
            _refuse("a 3D gradient-echo train needs at least one echo time")
        self.volumes = list(volumes)
        self.tes = [float(t) for t in echo_times_ms]
        self.ky_segments = ky_segments
        self.acq, self.t_line_ms, self.precision = acq, t_line_ms, precision
        groups: List[Group] = []
        slices: List[Slice] = []
        units: List[Unit] = []
        self.k2: List[List[object]] = []
        self.out: List[List[object]] = []
        zsets: List[ZSet] = []
        self._keep = []
        self._umeta = []      # per unit: (volume, echo, index among the used compartments, partition)
        self._used = []       # per volume: the compartments with an image
        self._zadj = {}       # (volume, echo) -> (n_used, weights) of its partition transform
        for vi, v in enumerate(self.volumes):
            check_volume(v, acq, ky_segments)
            dev = v.fmap.device
            P = v.snx * v.sny
            C = v.coil_maps.shape[1]
            ncomp = len(v.t2)
            used = [c for c in range(ncomp) if v.images[c] is not None]
            self._used.append(used)
            t, pol = grase_block(v.ny, ky_segments, t_line_ms, bool(acq.reverse_phase))
            w_all = v.weights.reshape(v.nz * ky_segments, ncomp)
            w_used = w_all[:, used].contiguous().reshape(-1) if used else None
            self._keep.append(w_used)
            fm = v.fmap.reshape(v.nz, P)
            ph = None if v.phase0 is None else v.phase0.reshape(v.nz, P)
            k2_v, out_v = [], []
            for e, te in enumerate(self.tes):
                a = acq.copy(echo=abi.ECHO_GRADIENT, t_echo=te, noise_variance=0.0, n_spikes=0)
                gi = len(groups)
                groups.append(Group(v.snx, v.sny, v.nx, v.ny, C, 1, a, timing=(t, te + t, t.copy(), pol),
                                    complex_images=False))
                k2 = torch.zeros((max(len(used), 1), v.nz, C, v.ny, v.nx), dtype=torch.complex64, device=dev)
                out = torch.zeros((v.nz, C, v.ny, v.nx), dtype=torch.complex64, device=dev)
                for ci, c in enumerate(used):
                    img = v.images[c].reshape(v.nz, P)
                    t2c, tic = v.t2[c], None if v.t_inhom is None else v.t_inhom[c]
                    t2m = None if isinstance(t2c, (int, float)) else t2c.reshape(v.nz, P)
                    tim = None if (tic is None or isinstance(tic, (int, float))) else tic.reshape(v.nz, P)
                    for z in range(v.nz):
                        si = len(slices)
                        slices.append(Slice(gi, fm[z], v.coil_maps[z], phase0=None if ph is None else ph[z],
                                            t2=None if t2m is None else t2m[z],
                                            t2_uniform=(float(t2c) if t2m is None else 100.0,) * abi.MRS_MAX_COMP,
                                            t_inhom=None if tim is None else tim[z],
                                            t_inhom_uniform=None if (tic is None or tim is not None)
                                            else (float(tic),) * abi.MRS_MAX_COMP,
                                            z=z, nz=v.nz))
                        units.append(Unit(si, img[z], k2[ci, z]))
                        self._umeta.append((vi, e, ci, z))
                if used:
                    self._zadj[(vi, e)] = (len(used), w_used)
                    zsets.append(ZSet(inp=k2, out=out, nz=v.nz, n_coils=C, ny=v.ny, nx=v.nx, n_comp=len(used),
                                      ky_segments=ky_segments, weights=w_used))
                k2_v.append(k2)
                out_v.append(out)
            self.k2.append(k2_v)
            self.out.append(out_v)
        if not units:
            _refuse("no compartment image in any volume")
        self.plan = Plan(groups, slices, units, precision)
        self.plan.set_zstage(zsets)
        self.zsets = zsets

    def forward(self, stream=None):
        self.plan.forward(stream)
        return self.out

    def vjp(self, grad_out, stream=None, after_forward=False):
        """Vector-Jacobian product of the train for a real loss L: grad_out[v][e] = dL/d out[v][e]
        ((nz, C, ny, nx) complex64, torch's convention dL/dRe + i dL/dIm).

        The partition transform's adjoint (conjugate transpose with the per-excitation weights,
        src/kspace3d.rs:90-100, 716-722) takes each (volume, echo) cotangent to its per-partition 2D
        cotangents; mrs_plan_vjp runs the adjoint of every partition unit (gradient-echo decay at the
        echo's TE); the units' gradients are summed over the echoes, and for the fieldmap over the
        compartments, that share an input.  The forward's inputs must be bound as at construction.
        Returns per volume a dict of device tensors:
          images:  per compartment (nz*P,) float32, None for a compartment without an image
          fmap:    (nz*P,) float32, per Hz
          t_inhom: per compartment (nz*P,) float32, per ms of T2' (at the uniform value where T2' is
                   uniform: sum it over the voxels for the scalar's gradient)
        after_forward: the inputs are unchanged since this train's last forward on `stream`; the
        per-unit adjoints then reuse that forward's slice planes.
        """
        import torch
        if not hasattr(self, "_vjp_buf"):
            gk2 = {k: torch.empty_like(self.k2[k[0]][k[1]]) for k in self._zadj}
            per_vol = []
            for vi, v in enumerate(self.volumes):
                shape = (len(self.tes), max(len(self._used[vi]), 1), v.nz, v.snx * v.sny)
                dev = v.fmap.device
                per_vol.append([torch.empty(shape, dtype=torch.float32, device=dev) for _ in range(3)])
            # the per-unit views stay fixed with the buffers: build them once
            views = ([], [], [], [])
            for vi, e, ci, z in self._umeta:
                b = per_vol[vi]
                views[0].append(gk2[(vi, e)][ci, z])
                views[1].append(b[0][e, ci, z])
                views[2].append(b[1][e, ci, z])
                views[3].append(b[2][e, ci, z])
            # the views never move either: their pointer arrays are built once too
            self._vjp_buf = (gk2, per_vol, self.plan.pointer_arrays(*views))
        gk2, per_vol, ptrs = self._vjp_buf
        # The adjoint partition transform runs from an mrs_zplan whose sets stay on the device.  The
        # first caller's cotangent tensors get a plan of their own, reused while they are passed
        # again; cotangents in other tensors are copied (asynchronously, on `stream`) into cotangent
        # buffers this train owns, whose plan is built once: no allocation and no wait per call.
        sets = self._zadj_sets(grad_out, gk2)
        key = tuple(z.inp.data_ptr() for z in sets)
        zp = getattr(self, "_zadj_plan", None)
        if zp is None:
            self._zadj_plan = zp = (key, ZPlan(sets, adjoint=True))
        if zp[0] == key:
            zp[1].run(stream)
        else:
            own = getattr(self, "_zadj_own", None)
            if own is None:
                bufs = [torch.empty_like(z.inp) for z in sets]
                own_sets = self._zadj_sets(self._nest(bufs), gk2)
                self._zadj_own = own = (bufs, ZPlan(own_sets, adjoint=True))
            cur = torch.cuda.current_stream() if stream is None else stream
            with torch.cuda.stream(cur):
                for b, z in zip(own[0], sets):
                    b.copy_(z.inp, non_blocking=True)
            own[1].run(stream)
        self.plan.vjp_arrays(ptrs, stream, after_forward=after_forward)
        out = []
        for vi, v in enumerate(self.volumes):
            b = per_vol[vi]
            used = self._used[vi]
            imgs, tis = [None] * len(v.images), [None] * len(v.images)
            for ci, c in enumerate(used):
                imgs[c] = b[0][:, ci].sum(dim=0).reshape(-1)
                tis[c] = b[2][:, ci].sum(dim=0).reshape(-1)
            fm = b[1].sum(dim=(0, 1)).reshape(-1) if used else torch.zeros_like(v.fmap)
            out.append({"images": imgs, "fmap": fm, "t_inhom": tis})
        return out

    def _nest(self, flat):
        """A flat list in the order of the adjoint sets as grad_out[volume][echo]."""
        nested = [[None] * len(self.tes) for _ in self.volumes]
        for (vi, e), b in zip(self._zadj.keys(), flat):
            nested[vi][e] = b
        return nested

    def _zadj_sets(self, grad_out, gk2):
        import torch
        sets = []
        for (vi, e), (n_used, w) in self._zadj.items():
            v = self.volumes[vi]
            g = grad_out[vi][e]
            if g.dtype != torch.complex64 or not g.is_contiguous() or g.shape != self.out[vi][e].shape:
                raise ValueError(f"grad_out[{vi}][{e}] must be a contiguous complex64 tensor of shape "
                                 f"{tuple(self.out[vi][e].shape)}")
            sets.append(ZSet(inp=g, out=gk2[(vi, e)], nz=v.nz, n_coils=v.coil_maps.shape[1], ny=v.ny, nx=v.nx,
                             n_comp=n_used, ky_segments=self.ky_segments, weights=w))
        return sets

    def _slices(self, vi, e, g, seed, acq):
        """Spikes and noise on a copy of the (volume, echo) 3D k-space, then the inverse z-DFT / nz:
        (nz, C, ny, nx) complex64 2D coil k-spaces per slice."""
        import torch
        v = self.volumes[vi]
        k3 = self.out[vi][e].clone()
        partition_noise(k3, acq, g, seed)
        nz, C = v.nz, v.coil_maps.shape[1]
        slices = torch.empty_like(k3)
        w = torch.full((nz,), 1.0 / nz, dtype=torch.float32, device=k3.device)
        zstage([ZSet(k3, slices, nz, C, v.ny, v.nx, weights=w)], adjoint=True)
        return slices

    def grappa_calibration(self, seed: int):
        """GRAPPA kernels per (volume, echo, slice), fitted on a reference scan of this forward's
        volumes (grappa_kernels_3d, src/kspace3d.rs:303-323): the acquisition without acceleration and
        spikes, volume index 0, noise on seed ^ echo_salt(e) ^ SEED_SALT_GRAPPA_CAL.  Returns
        weights[v][e][z] device tensors; None without acceleration."""
        if self.acq.accel <= 1:
            return None
        acq_cal = self.acq.copy(accel=1, n_spikes=0)
        cal = Ge3dForward(self.volumes, acq_cal, self.tes, self.ky_segments, self.t_line_ms, self.precision)
        cal.forward()
        out = []
        for vi, v in enumerate(self.volumes):
            lines = sampling_mask(1, v.ny, acq_cal).reshape(-1)
            row = []
            for e in range(len(self.tes)):
                sl = cal._slices(vi, e, 0, seed ^ echo_salt(e) ^ SEED_SALT_GRAPPA_CAL, acq_cal)
                row.append([grappa_fit(sl[z].contiguous(), self.acq.accel, self.acq.acs_lines, lines)
                            for z in range(v.nz)])
            out.append(row)
        cal.close()
        return out

    def acquire(self, seed: int, volume_index: Optional[Sequence[int]] = None, grappa=None):
        """The acquired, reconstructed complex volumes after `forward`: images[v][e] of (nz, ny, nx)
        complex64 (vox = x + nx*(y + ny*z)).  `volume_index[v]` is volume v's index g in its series
        (default v); `grappa` the kernels of `grappa_calibration` (required under acceleration)."""
        import torch
        if self.acq.accel > 1 and grappa is None:
            _refuse("an accelerated slice needs its GRAPPA kernel (fitted on a reference scan)")
        images = []
        for vi, v in enumerate(self.volumes):
            g = vi if volume_index is None else int(volume_index[vi])
            C = v.coil_maps.shape[1]
            rec = Recon(v.nx, v.ny, C, self.acq)
            acquired = sampling_mask(1, v.ny, self.acq).reshape(-1)
            row = []
            for e in range(len(self.tes)):
                sl = self._slices(vi, e, g, seed ^ echo_salt(e), self.acq)
                if grappa is not None:
                    for z in range(v.nz):
                        grappa_apply([sl[z]], self.acq.accel, acquired, grappa[vi][e][z])
                img = torch.empty((v.nz, v.ny, v.nx), dtype=torch.complex64, device=sl.device)
                rec.run([sl[z] for z in range(v.nz)], [img[z] for z in range(v.nz)])
                row.append(img)
            rec.close()
            images.append(row)
        return images

    def forward_graph(self, stream=None):
        self.plan.forward_graph(stream)
        return self.out

    def tune(self, trials=3):
        return self.plan.tune(trials)

    def describe(self):
        return self.plan.describe()

    def stage_times(self, iters=5):
        return self.plan.stage_times(iters)

    def close(self):
        for attr in ("_zadj_plan", "_zadj_own"):
            zp = getattr(self, attr, None)
            if zp is not None:
                zp[1].close()
                setattr(self, attr, None)
        self.plan.close()


__all__ = ["Ge3dVolume", "Ge3dForward", "check_volume", "MrsError"]
