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
        _lib = L
    return _lib


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
