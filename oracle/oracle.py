"""TEST INFRASTRUCTURE: Python access to the CPU FP64 oracle (oracle/mrsim_oracle.c).

Only tests/, `__graft_entry__.smoke()` and bench.py's CPU-baseline leg may import this module, and
only as the checker.  The product path never imports it.  Struct layouts come from the product's
ctypes mirror of include/mrsim_gpu.h (both sides compile against that header).
"""
from __future__ import annotations

import ctypes as ct
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(ROOT, "mrsim-acq_amd"))

from mrsim_amd import abi  # noqa: E402
from mrsim_amd.abi import C64, Acquisition, LineTiming, SliceInput  # noqa: E402

LIB_PATH = os.path.join(HERE, "build", "libmrsim_oracle.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ct.CDLL(LIB_PATH)
        P = ct.POINTER
        L.orc_build_coil_kspace_timed.argtypes = [P(SliceInput), P(Acquisition), ct.c_size_t, ct.c_size_t,
                                                  P(LineTiming), P(ct.c_double), P(C64)]
        L.orc_build_coil_kspace.argtypes = [P(SliceInput), P(Acquisition), ct.c_size_t, ct.c_size_t, P(C64)]
        L.orc_literal_coil_kspace.argtypes = [P(SliceInput), P(Acquisition), ct.c_size_t, ct.c_size_t, P(C64)]
        L.orc_simulate_slice.argtypes = [P(SliceInput), P(Acquisition), P(ct.c_float)]
        L.orc_line_timing.argtypes = [P(Acquisition), ct.c_size_t, ct.c_size_t, P(ct.c_double), P(ct.c_double),
                                      P(ct.c_double), P(ct.c_int8)]
        L.orc_sampling_mask.argtypes = [ct.c_size_t, ct.c_size_t, P(Acquisition), P(ct.c_uint8)]
        L.orc_coil_sensitivity.argtypes = [ct.c_size_t, ct.c_size_t, ct.c_double, ct.c_double, ct.c_size_t,
                                           ct.c_size_t]
        L.orc_coil_sensitivity.restype = ct.c_double
        L.orc_golden_slice_inputs.argtypes = [ct.c_int, P(ct.c_float), P(ct.c_float), P(ct.c_double),
                                              P(ct.c_float), P(ct.c_float)]
        L.orc_inverse_2d.argtypes = [P(C64), ct.c_size_t, ct.c_size_t, P(C64)]
        L.orc_reconstruct.argtypes = [P(C64), ct.c_size_t, ct.c_size_t, ct.c_size_t, P(Acquisition), P(C64),
                                      P(C64)]
        L.orc_grappa_fit.argtypes = [P(P(C64)), ct.c_size_t, ct.c_size_t, ct.c_size_t, ct.c_size_t, ct.c_size_t,
                                     P(ct.c_uint8), P(C64)]
        L.orc_box_hires.argtypes = [ct.c_size_t, ct.c_size_t, ct.c_double, ct.c_double, ct.c_double, ct.c_double,
                                    P(ct.c_float)]
        L.orc_rng_next.argtypes = [P(ct.c_uint64)]
        L.orc_rng_next.restype = ct.c_uint64
        L.orc_rng_unit.argtypes = [P(ct.c_uint64)]
        L.orc_rng_unit.restype = ct.c_double
        L.orc_rng_gauss.argtypes = [P(ct.c_uint64)]
        L.orc_rng_gauss.restype = ct.c_double
        L.orc_window_at.argtypes = [ct.c_int, ct.c_double, ct.c_double, ct.c_double, ct.c_double]
        L.orc_window_at.restype = ct.c_double
        L.orc_prep_shot.argtypes = [P(PhaseModel), ct.c_double, P(ct.c_double), ct.c_uint64, ct.c_uint64,
                                    ct.c_uint64, P(ct.c_double)]
        L.orc_phase_slice.argtypes = [P(PhaseModel), P(ct.c_double), ct.c_size_t, ct.c_size_t, ct.c_size_t,
                                      ct.c_size_t, ct.c_size_t, P(ct.c_double)]
        L.orc_acquire_volumes.argtypes = [P(ct.c_size_t), P(ct.c_size_t), ct.c_size_t, P(P(ct.c_float)),
                                          ct.c_size_t, P(abi.T2Slice), P(abi.T2Slice), P(ct.c_float),
                                          P(Acquisition), P(ct.c_int32), P(ct.c_double), P(ct.c_int32),
                                          P(ct.c_double), P(ct.c_double), P(PhaseModel), ct.c_uint64,
                                          ct.c_uint64, P(ct.c_float), P(ct.c_float), P(ct.c_float)]
        L.orc_mag_phase.argtypes = [P(ct.c_float), P(ct.c_float), ct.c_size_t, P(ct.c_float), P(ct.c_float)]
        L.orc_series_inputs.argtypes = [P(ct.c_float), P(ct.c_float), P(ct.c_float)]
        L.orc_nufft1.argtypes = [ct.c_size_t, ct.c_long, ct.c_size_t, ct.c_double, P(ct.c_double), ct.c_size_t,
                                 P(ct.c_double), P(ct.c_double), ct.c_size_t, P(ct.c_double), P(ct.c_double)]
        L.orc_nufft_direct.argtypes = [ct.c_size_t, ct.c_long, ct.c_size_t, P(ct.c_double), ct.c_size_t,
                                       P(ct.c_double), P(ct.c_double), P(ct.c_double), P(ct.c_double)]
        L.orc_build_coil_kspace_nufft.argtypes = [P(SliceInput), P(Acquisition), ct.c_size_t, ct.c_size_t,
                                                  P(ct.c_int), P(C64)]
        L.orc_set_kspace_feature.argtypes = [ct.c_int]
        L.orc_set_fft_xstage.argtypes = [ct.c_int]
        _lib = L
    return _lib


class PhaseModel(ct.Structure):
    """PhaseModel (src/phase.rs:129-137) as the oracle takes it."""
    _fields_ = [("global_", ct.c_double), ("coeffs", ct.c_double * 10), ("has_prep", ct.c_int32),
                ("c_q", ct.c_double), ("sigma_dx", ct.c_double), ("sigma_rot", ct.c_double)]

    @classmethod
    def hbcd_like(cls):
        """PhaseModel::hbcd_like (src/phase.rs:188-198)."""
        m = cls()
        for i, v in enumerate((0.0, 0.13, 0.10, 0.02)):
            m.coeffs[i] = v
        m.has_prep, m.c_q, m.sigma_dx, m.sigma_rot = 1, 0.0425, 1.0, 2.0e-3
        return m


def prep_shot(model, magnitude, direction, volume, slice_group, seed):
    d = np.ascontiguousarray(direction, dtype=np.float64)
    out = np.zeros(9)
    lib().orc_prep_shot(ct.byref(model), magnitude, abi.dptr(d), volume, slice_group, seed & _M64, abi.dptr(out))
    return out


def phase_slice(model, shot, snx, sny, o, z, nz):
    sh = np.ascontiguousarray(np.zeros(9) if shot is None else shot, dtype=np.float64)
    out = np.zeros(snx * sny)
    lib().orc_phase_slice(ct.byref(model), abi.dptr(sh), snx, sny, o, z, nz, abi.dptr(out))
    return out


_M64 = (1 << 64) - 1


def _t2_volumes(vals):
    keep, arr = [], (abi.T2Slice * max(len(vals), 1))()
    for q, v in enumerate(vals):
        if isinstance(v, (int, float)):
            arr[q] = abi.T2Slice(None, float(v))
        else:
            a = np.ascontiguousarray(v, dtype=np.float32)
            keep.append(a)
            arr[q] = abi.T2Slice(abi.fptr(a), 0.0)
    return arr, keep


def acquire_volumes(sim_dims, acq_dims, nv, images, t2, fmap, t_inhom, acq, eddy_drive=None, prep_drive=None,
                    model=None, excitation_seed=0, receiver_seed=0, noise_sigma=None):
    """acquire_volumes_complex (src/kspace.rs:1851-2125) with default options: (re, im) f32 arrays of
    nx*ny*nz*nv in the layout (x + nx*(y + ny*z))*nv + g.  images: per compartment, (vox*nv + g);
    t2/t_inhom: per compartment a float or a simulation-volume map; eddy_drive/prep_drive: per volume
    None or a value ((magnitude, direction) for prep)."""
    nx, ny, nz = acq_dims
    n = nx * ny * nz * nv
    ims = [np.ascontiguousarray(im, dtype=np.float32) for im in images]
    ptrs = (ct.POINTER(ct.c_float) * len(ims))(*[abi.fptr(a) for a in ims])
    t2a, k1 = _t2_volumes(t2)
    tia, k2 = _t2_volumes(t_inhom) if t_inhom is not None else (None, [])
    fm = np.ascontiguousarray(fmap, dtype=np.float32)
    he = np.array([e is not None for e in (eddy_drive or [None] * nv)], dtype=np.int32)
    ed = np.array([e if e is not None else (0.0, 0.0, 0.0) for e in (eddy_drive or [None] * nv)], dtype=np.float64)
    hp = np.array([p is not None for p in (prep_drive or [None] * nv)], dtype=np.int32)
    pm = np.array([p[0] if p is not None else 0.0 for p in (prep_drive or [None] * nv)], dtype=np.float64)
    pd = np.array([p[1] if p is not None else (0.0, 0.0, 0.0) for p in (prep_drive or [None] * nv)],
                  dtype=np.float64)
    model = model if model is not None else PhaseModel()
    ns = None if noise_sigma is None else np.ascontiguousarray(noise_sigma, dtype=np.float32)
    re, im = np.zeros(n, dtype=np.float32), np.zeros(n, dtype=np.float32)
    sd = (ct.c_size_t * 3)(*sim_dims)
    ad = (ct.c_size_t * 3)(*acq_dims)
    rc = lib().orc_acquire_volumes(sd, ad, nv, ptrs, len(ims), t2a, tia, abi.fptr(fm), ct.byref(acq),
                                   he.ctypes.data_as(ct.POINTER(ct.c_int32)), abi.dptr(ed.reshape(-1)),
                                   hp.ctypes.data_as(ct.POINTER(ct.c_int32)), abi.dptr(pm), abi.dptr(pd.reshape(-1)),
                                   ct.byref(model), excitation_seed & _M64, receiver_seed & _M64,
                                   abi.fptr(ns) if ns is not None else ct.POINTER(ct.c_float)(),
                                   abi.fptr(re), abi.fptr(im))
    if rc != 0:
        raise ValueError("oracle acquire_volumes refused the input")
    return re, im


def mag_phase(re, im):
    """acquire_volumes' f32 magnitude and atan2 phase (src/kspace.rs:1840-1846)."""
    re, im = np.ascontiguousarray(re, dtype=np.float32), np.ascontiguousarray(im, dtype=np.float32)
    mag, ph = np.zeros_like(re), np.zeros_like(re)
    lib().orc_mag_phase(abi.fptr(re), abi.fptr(im), len(re), abi.fptr(mag), abi.fptr(ph))
    return mag, ph


def nufft1(n, k_lo, n_out, tol, pos, weights):
    """Nufft1::new(n, k_lo, n_out, tol).run(pos, weights) (src/nufft.rs:25-113): weights is a list of
    complex vectors over the sources; returns a list of complex n_out vectors."""
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    w = np.ascontiguousarray(np.stack([np.asarray(x) for x in weights]), dtype=np.complex128)
    wre, wim = np.ascontiguousarray(w.real), np.ascontiguousarray(w.imag)
    ore, oim = np.zeros((len(weights), n_out)), np.zeros((len(weights), n_out))
    rc = lib().orc_nufft1(n, k_lo, n_out, tol, abi.dptr(pos), len(pos), abi.dptr(wre), abi.dptr(wim), len(weights),
                          abi.dptr(ore), abi.dptr(oim))
    if rc:
        raise ValueError("output band must sit inside the coarse Nyquist band")
    return list(ore + 1j * oim)


def nufft_direct(n, k_lo, n_out, pos, weight):
    """The literal sum a type-1 NUFFT approximates (src/nufft.rs:301-315)."""
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    w = np.asarray(weight, dtype=np.complex128)
    re, im = np.ascontiguousarray(w.real), np.ascontiguousarray(w.imag)
    ore, oim = np.zeros(n_out), np.zeros(n_out)
    lib().orc_nufft_direct(n, k_lo, n_out, abi.dptr(pos), len(pos), abi.dptr(re), abi.dptr(im), abi.dptr(ore),
                           abi.dptr(oim))
    return ore + 1j * oim


def coil_kspace_nufft(b, acq, coil=0, ncoils=1):
    """build_coil_kspace as the `kspace` feature builds it: (k-space, whether the NUFFT y-stage ran)."""
    nx, ny = b.acq_matrix
    out = np.zeros(nx * ny, dtype=np.complex128)
    used = ct.c_int(0)
    rc = lib().orc_build_coil_kspace_nufft(ct.byref(b.inp), ct.byref(acq), coil, ncoils, ct.byref(used), _c64(out))
    if rc != 0:
        raise ValueError("oracle refused the input")
    return out, bool(used.value)


class kspace_feature:
    """Context manager: the per-slice entry points restate the `kspace` feature build inside it."""

    def __enter__(self):
        lib().orc_set_kspace_feature(1)
        return self

    def __exit__(self, *exc):
        lib().orc_set_kspace_feature(0)
        return False


class fft_xstage:
    """Context manager: inside it the per-slice forward runs the `kspace` feature's FFT x-stage
    (ramp, unnormalised inverse FFT, centre phase; src/kspace.rs:1077-1088, 1110-1124) in place of
    the twiddle tables (:1091-1101, 1126-1137)."""

    def __enter__(self):
        lib().orc_set_fft_xstage(1)
        return self

    def __exit__(self, *exc):
        lib().orc_set_fft_xstage(0)
        return False


SERIES_DIMS = dict(nx=16, ny=16, nz=3, o=2, nv=4)


def series_inputs():
    """The series inputs of tests/golden_baseline.rs:91-116: (images per compartment, fmap, t2map)."""
    d = SERIES_DIMS
    nsim = d["nx"] * d["o"] * d["ny"] * d["o"] * d["nz"]
    images = np.zeros(2 * nsim * d["nv"], dtype=np.float32)
    fmap, t2map = np.zeros(nsim, dtype=np.float32), np.zeros(nsim, dtype=np.float32)
    lib().orc_series_inputs(abi.fptr(images), abi.fptr(fmap), abi.fptr(t2map))
    return list(images.reshape(2, -1)), fmap, t2map


def series_cases():
    """The two series configurations (tests/golden_baseline.rs:113-124): (name, t2, acq)."""
    _, _, t2map = series_inputs()
    acq = Acquisition.default(n_coils=3, accel=2, acs_lines=6, noise_variance=2.0, signal_scale=100.0,
                              t_echo=30.0, t_line=0.5)
    return [("uniform", [80.0, 160.0], acq), ("mapped", [t2map, 160.0], acq)]


def series_golden(name):
    """The reference's record of series-<name>: (magnitude, phase) f32, each nx*ny*nz*nv."""
    raw = np.fromfile(os.path.join(ROOT, "tests", "golden", "p4_baseline", "default", f"series-{name}.bin"),
                      dtype=np.float32)
    return raw[: raw.size // 2], raw[raw.size // 2:]


def series(name, model=None):
    """One golden series through the oracle: (magnitude, phase) f32."""
    d = SERIES_DIMS
    images, fmap, _ = series_inputs()
    t2, acq = {c[0]: c for c in series_cases()}[name][1:]
    sim = (d["nx"] * d["o"], d["ny"] * d["o"], d["nz"])
    re, im = acquire_volumes(sim, (d["nx"], d["ny"], d["nz"]), d["nv"], images, t2, fmap, None, acq,
                             model=model if model is not None else PhaseModel.hbcd_like(),
                             excitation_seed=42, receiver_seed=42)
    return mag_phase(re, im)


def _c64(a):
    return a.ctypes.data_as(ct.POINTER(C64))


def line_timing(acq, nx, ny):
    t = np.zeros(ny)
    trf = np.zeros(ny)
    tread = np.zeros(ny)
    pol = np.zeros(ny, dtype=np.int8)
    lib().orc_line_timing(ct.byref(acq), nx, ny, abi.dptr(t), abi.dptr(trf), abi.dptr(tread),
                          pol.ctypes.data_as(ct.POINTER(ct.c_int8)))
    return t, trf, tread, pol


def sampling_mask(nx, ny, acq):
    m = np.zeros(nx * ny, dtype=np.uint8)
    lib().orc_sampling_mask(nx, ny, ct.byref(acq), m.ctypes.data_as(ct.POINTER(ct.c_uint8)))
    return m.astype(bool)


def coil_kspace(b: abi.SliceInputBuilder, acq, coil=0, ncoils=1, timing=None, log_amp=None):
    """build_coil_kspace(_timed): nx*ny complex128, layout kx + nx*ky."""
    nx, ny = b.acq_matrix
    out = np.zeros(nx * ny, dtype=np.complex128)
    if timing is None and log_amp is None:
        rc = lib().orc_build_coil_kspace(ct.byref(b.inp), ct.byref(acq), coil, ncoils, _c64(out))
    else:
        t, trf, tread, pol = timing if timing is not None else line_timing(acq, nx, ny)
        keep = [np.ascontiguousarray(v, dtype=np.float64) for v in (t, trf, tread)]
        pol = np.ascontiguousarray(pol, dtype=np.int8)
        lt = LineTiming(abi.dptr(keep[0]), abi.dptr(keep[1]), abi.dptr(keep[2]),
                        pol.ctypes.data_as(ct.POINTER(ct.c_int8)), ny)
        la = None
        if log_amp is not None:
            la = np.ascontiguousarray(log_amp, dtype=np.float64)
        rc = lib().orc_build_coil_kspace_timed(ct.byref(b.inp), ct.byref(acq), coil, ncoils, ct.byref(lt),
                                               abi.dptr(la) if la is not None else ct.POINTER(ct.c_double)(),
                                               _c64(out))
    if rc != 0:
        raise ValueError(f"oracle forward refused the input (code {rc})")
    return out


def coil_kspaces(b, acq):
    nc = max(int(acq.n_coils), 1)
    return np.stack([coil_kspace(b, acq, c, nc) for c in range(nc)])


def literal_coil_kspace(b, acq, coil=0, ncoils=1):
    nx, ny = b.acq_matrix
    out = np.zeros(nx * ny, dtype=np.complex128)
    rc = lib().orc_literal_coil_kspace(ct.byref(b.inp), ct.byref(acq), coil, ncoils, _c64(out))
    if rc != 0:
        raise ValueError("oracle literal sum refused the input")
    return out


def simulate_slice(b, acq):
    """simulate_slice: nx*ny complex64 (bit patterns of the reference's f32 pairs)."""
    nx, ny = b.acq_matrix
    out = np.zeros(2 * nx * ny, dtype=np.float32)
    rc = lib().orc_simulate_slice(ct.byref(b.inp), ct.byref(acq), abi.fptr(out))
    if rc != 0:
        raise ValueError("oracle simulate_slice refused the input")
    return out


def grappa_fit(coils, nx, ny, accel, acs_lines, lines):
    coils = np.ascontiguousarray(coils, dtype=np.complex128)
    nc = coils.shape[0]
    ptrs = (ct.POINTER(C64) * nc)(*[_c64(coils[c]) for c in range(nc)])
    w = np.zeros((accel - 1) * nc * 6 * nc, dtype=np.complex128)
    lines = np.ascontiguousarray(lines, dtype=np.uint8)
    lib().orc_grappa_fit(ptrs, nc, nx, ny, accel, acs_lines, lines.ctypes.data_as(ct.POINTER(ct.c_uint8)), _c64(w))
    return w


def reconstruct(coil_kspace, nx, ny, acq, weights=None):
    k = np.ascontiguousarray(coil_kspace, dtype=np.complex128).copy()
    nc = k.shape[0]
    out = np.zeros(nx * ny, dtype=np.complex128)
    lib().orc_reconstruct(_c64(k), nc, nx, ny, ct.byref(acq),
                          _c64(weights) if weights is not None else ct.POINTER(C64)(), _c64(out))
    return out


def inverse_2d(k, nx, ny):
    k = np.ascontiguousarray(k, dtype=np.complex128)
    out = np.zeros(nx * ny, dtype=np.complex128)
    lib().orc_inverse_2d(_c64(k), nx, ny, _c64(out))
    return out


GOLDEN_SHAPES = [(17, 70, 2), (16, 40, 1), (8, 9, 3)]


def golden_inputs(shape_idx):
    """The inputs of tests/golden_baseline.rs:32-55 for shape `shape_idx`."""
    nx, ny, o = GOLDEN_SHAPES[shape_idx]
    n = nx * o * ny * o
    comps = np.zeros(3 * n, dtype=np.float32)
    fmap = np.zeros(n, dtype=np.float32)
    phase0 = np.zeros(n, dtype=np.float64)
    t2m = np.zeros(n, dtype=np.float32)
    tim = np.zeros(n, dtype=np.float32)
    lib().orc_golden_slice_inputs(shape_idx, abi.fptr(comps), abi.fptr(fmap), abi.dptr(phase0), abi.fptr(t2m),
                                  abi.fptr(tim))
    return comps.reshape(3, n), fmap, phase0, t2m, tim


def golden_cases():
    """The 13 configurations of tests/golden_baseline.rs:56-83: (name, acq, eddy_drive, maps)."""
    full = Acquisition.default(do_distortions=1, do_relaxation=1, signal_scale=100.0)
    return [
        ("clean", full.copy(do_distortions=0, do_relaxation=0), None, False),
        ("distortion+relaxation", full.copy(), None, False),
        ("reverse", full.copy(reverse_phase=1), None, False),
        ("ghost", full.copy(ghost_offset=0.015), None, False),
        ("eddy-poly", full.copy(eddy_strength=3.0, eddy_quad=0.4, eddy_tau=70.0), (0.3, -0.8, 0.5), False),
        ("eddy-phase", full.copy(eddy_phase=0.2), (0.6, 0.6, 0.5), False),
        ("pf-fiberfox", full.copy(partial_fourier=0.75), None, False),
        ("pf-contiguous-reverse", full.copy(partial_fourier=0.75, pf_mode=abi.PF_CONTIGUOUS, reverse_phase=1),
         None, False),
        ("grappa-coils", full.copy(accel=2, acs_lines=6, n_coils=4), None, False),
        ("grappa3", full.copy(accel=3, acs_lines=8, n_coils=4), None, False),
        ("everything", full.copy(ghost_offset=0.02, eddy_strength=2.0, eddy_quad=0.3, eddy_phase=0.1,
                                 partial_fourier=0.8, accel=2, acs_lines=8, n_coils=3), (0.5, 0.5, 0.7), False),
        ("maps", full.copy(), None, True),
        ("maps-coils-noise", full.copy(n_coils=3, noise_variance=4.0), None, True),
    ]


def golden_slice_builder(shape_idx, eddy_drive, maps, slice_seed=7):
    nx, ny, o = GOLDEN_SHAPES[shape_idx]
    comps, fmap, phase0, t2m, tim = golden_inputs(shape_idx)
    t2 = [t2m, 100.0, t2m] if maps else [70.0, 100.0, 2000.0]
    ti = [tim, 45.0, tim] if maps else None
    return abi.SliceInputBuilder(list(comps), t2, fmap, (nx * o, ny * o), (nx, ny), z=3, nz=9, t_inhom=ti,
                                 phase0=phase0, eddy_drive=eddy_drive, slice_seed=slice_seed)


def forward_direct(snx, sny, nx, ny, n_coils, n_comp, acq, images, fmap, coil_maps, phase0=None, t2=None,
                   t2_uniform=None, t_inhom=None, t_inhom_uniform=None, log_amp=None, z=0, nz=1, eddy_drive=None,
                   timing=None, complex_images=False):
    """FP64 forward on the batched plan's inputs (orc_forward_direct): (C, ny, nx) complex128."""
    L = lib()
    if not hasattr(L, "_fd"):
        L.orc_forward_direct.argtypes = [ct.POINTER(abi.GroupDesc), ct.POINTER(abi.SliceDesc), ct.c_void_p,
                                         ct.POINTER(C64)]
        L._fd = True
    keep = []

    def arr(a, dt):
        if a is None:
            return None
        a = np.ascontiguousarray(a, dtype=dt)
        keep.append(a)
        return a.ctypes.data

    P = snx * sny
    for name, a, n in (("coil_maps", coil_maps, n_coils * P), ("fmap", fmap, P), ("images", images, n_comp * P),
                       ("phase0", phase0, P), ("t2", t2, n_comp * P), ("t_inhom", t_inhom, n_comp * P),
                       ("log_amp", log_amp, P)):
        if a is not None and np.size(a) != n:
            raise ValueError(f"forward_direct: {name} has {np.size(a)} values, expected {n}")
    g = abi.GroupDesc()
    g.snx, g.sny, g.nx, g.ny, g.n_coils, g.n_compartments = snx, sny, nx, ny, n_coils, n_comp
    g.complex_images = 1 if complex_images else 0
    g.acq = acq
    if timing is not None:
        tk = [np.ascontiguousarray(v, dtype=np.float64) for v in timing[:3]]
        pol = np.ascontiguousarray(timing[3], dtype=np.int8)
        keep += tk + [pol]
        lt = LineTiming(abi.dptr(tk[0]), abi.dptr(tk[1]), abi.dptr(tk[2]),
                        pol.ctypes.data_as(ct.POINTER(ct.c_int8)), ny)
        keep.append(lt)
        g.timing = ct.pointer(lt)
    s = abi.SliceDesc()
    s.group = 0
    s.fmap = arr(fmap, np.float32)
    s.coil_maps = arr(coil_maps, np.complex64)
    s.phase0 = arr(phase0, np.float64)
    s.t2 = arr(t2, np.float32)
    s.t_inhom = arr(t_inhom, np.float32)
    s.log_amp = arr(log_amp, np.float64)
    if t2_uniform is not None:
        for q, v in enumerate(t2_uniform):
            s.t2_uniform[q] = v
    if t_inhom_uniform is not None:
        s.has_t_inhom = 1
        for q, v in enumerate(t_inhom_uniform):
            s.t_inhom_uniform[q] = v
    if t_inhom is not None:
        s.has_t_inhom = 1
    s.z, s.nz = z, nz
    if eddy_drive is not None:
        s.has_eddy_drive = 1
        for k in range(3):
            s.eddy_drive[k] = eddy_drive[k]
    img = arr(images, np.complex64 if complex_images else np.float32)
    out = np.zeros(n_coils * ny * nx, dtype=np.complex128)
    rc = L.orc_forward_direct(ct.byref(g), ct.byref(s), img, _c64(out))
    if rc != 0:
        raise ValueError("oracle direct forward refused the input")
    return out.reshape(n_coils, ny, nx)


def forward_workload_unit(w, unit_index):
    """FP64 oracle of one unit of a workload (mrsim_amd.workloads.Workload)."""
    si, img = w.units[unit_index]
    sd = w.slices[si]
    return forward_direct(w.snx, w.sny, w.nx, w.ny, w.n_coils, w.n_comp, w.acq, img, sd.fmap, sd.coil_maps,
                          phase0=None if w.complex_images else sd.phase0, t2=sd.t2, t_inhom=sd.t_inhom,
                          complex_images=w.complex_images)
