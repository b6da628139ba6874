"""Python binding of libmrsim_gpu.so (the HIP kernels behind include/mrsim_gpu.h).

The library must be built (``python -m mrsim_amd.build`` or ``__graft_entry__.build()``); loading
fails loudly otherwise -- there is no CPU fallback.  torch supplies device memory and streams only.

Entry points mirror the reference's per-slice forward:
  build_coil_kspace   <-> build_coil_kspace_timed / simulate_slice_kspace (src/kspace.rs:597-609, 1144)
  build_coil_kspaces  <-> build_coil_kspaces (src/kspace.rs:1248-1251)
and add the batched ``Plan`` (many slice x volume x echo units in one launch) with its adjoint.
"""
from __future__ import annotations

import ctypes as ct
import os
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

from . import abi

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(PKG), "build", "libmrsim_gpu.so")

_lib = None


class MrsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"mrsim_gpu error {code}: {msg}")
        self.code = code


def lib():
    """The loaded C-ABI library (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                          "(the HIP extension is required; there is no CPU fallback)")
    L = ct.CDLL(LIB_PATH)
    P = ct.POINTER
    E = [ct.c_char_p, ct.c_size_t]
    L.mrs_acquisition_default.argtypes = [P(abi.Acquisition)]
    L.mrs_line_timing_for_acquisition.argtypes = [P(abi.Acquisition), ct.c_size_t, ct.c_size_t, P(ct.c_double),
                                                  P(ct.c_double), P(ct.c_double), P(ct.c_int8)]
    L.mrs_sampling_mask.argtypes = [ct.c_size_t, ct.c_size_t, P(abi.Acquisition), P(ct.c_uint8)]
    L.mrs_ring_coil_maps.argtypes = [ct.c_size_t] * 7 + [ct.c_double, P(abi.C32)]
    L.mrs_build_coil_kspace.argtypes = [P(abi.SliceInput), P(abi.Acquisition), ct.c_size_t, ct.c_size_t,
                                        P(abi.LineTiming), P(ct.c_double), P(abi.C64)] + E
    L.mrs_build_coil_kspaces.argtypes = [P(abi.SliceInput), P(abi.Acquisition), P(abi.LineTiming), P(ct.c_double),
                                         P(abi.C64)] + E
    L.mrs_plan_create.argtypes = [P(abi.GroupDesc), ct.c_size_t, P(abi.SliceDesc), ct.c_size_t, P(abi.UnitDesc),
                                  ct.c_size_t, ct.c_int32, P(ct.c_void_p)] + E
    L.mrs_plan_set_pointers.argtypes = [ct.c_void_p, P(abi.SliceDesc), P(abi.UnitDesc)] + E
    L.mrs_plan_forward.argtypes = [ct.c_void_p, ct.c_void_p] + E
    L.mrs_plan_forward_graph.argtypes = [ct.c_void_p, ct.c_void_p] + E
    L.mrs_plan_vjp.argtypes = [ct.c_void_p, P(ct.c_void_p), P(ct.c_void_p), P(ct.c_void_p), P(ct.c_void_p),
                               ct.c_void_p] + E
    L.mrs_plan_vjp_after_forward.argtypes = L.mrs_plan_vjp.argtypes
    L.mrs_plan_stage_times.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_int32, P(ct.c_double)] + E
    L.mrs_plan_stats.argtypes = [ct.c_void_p, P(ct.c_double), P(ct.c_double), P(ct.c_double)]
    L.mrs_plan_describe.argtypes = [ct.c_void_p, ct.c_char_p, ct.c_size_t]
    L.mrs_plan_destroy.argtypes = [ct.c_void_p]
    L.mrs_plan_tune.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_int32, ct.c_char_p, ct.c_size_t]
    L.mrs_plan_set_tiles.argtypes = [ct.c_void_p, ct.c_int32, ct.c_int32, ct.c_int32] + E
    L.mrs_plan_get_tiles.argtypes = [ct.c_void_p, P(ct.c_int32), P(ct.c_int32), P(ct.c_int32)]
    L.mrs_plan_set_basis.argtypes = [ct.c_void_p, ct.c_int32] + E
    L.mrs_plan_get_basis.argtypes = [ct.c_void_p, P(ct.c_int32), P(ct.c_int32)]
    L.mrs_plan_set_vjp_basis.argtypes = [ct.c_void_p, ct.c_int32] + E
    L.mrs_plan_get_vjp_basis.argtypes = [ct.c_void_p, P(ct.c_int32)]
    L.mrs_version.restype = ct.c_char_p
    L.mrs_sizeof.argtypes = [ct.c_char_p]
    L.mrs_sizeof.restype = ct.c_size_t
    L.mrs_probe_sincos.argtypes = [P(ct.c_double), P(ct.c_float), P(ct.c_float), ct.c_size_t, ct.c_int32] + E
    L.mrs_grase_block.argtypes = [ct.c_size_t, ct.c_size_t, ct.c_double, ct.c_int32, P(ct.c_double), P(ct.c_int8)] + E
    L.mrs_ge3d_lines.argtypes = [ct.c_size_t, ct.c_size_t, ct.c_size_t, ct.c_double, ct.c_int32, ct.c_size_t,
                                 ct.c_int32, P(ct.c_int32), P(ct.c_int32), P(ct.c_size_t), P(ct.c_size_t),
                                 P(ct.c_size_t)] + E
    L.mrs_check_ge3d_timing.argtypes = [P(ct.c_double), ct.c_size_t, ct.c_double, ct.c_double, ct.c_size_t,
                                        ct.c_size_t, ct.c_double, ct.c_int32, ct.c_size_t, ct.c_double,
                                        ct.c_double] + E
    L.mrs_zstage.argtypes = [P(abi.ZSetDesc), ct.c_size_t, ct.c_int32, ct.c_void_p] + E
    L.mrs_zplan_create.argtypes = [P(abi.ZSetDesc), ct.c_size_t, ct.c_int32, P(ct.c_void_p)] + E
    L.mrs_zplan_run.argtypes = [ct.c_void_p, ct.c_void_p] + E
    L.mrs_zplan_destroy.argtypes = [ct.c_void_p]
    L.mrs_zplan_destroy.restype = None
    L.mrs_recon_create.argtypes = [ct.c_size_t, ct.c_size_t, ct.c_size_t, P(abi.Acquisition), P(ct.c_float),
                                   P(ct.c_void_p)] + E
    L.mrs_recon_run.argtypes = [ct.c_void_p, P(ct.c_void_p), P(ct.c_void_p), ct.c_size_t, ct.c_void_p] + E
    L.mrs_recon_destroy.argtypes = [ct.c_void_p]
    L.mrs_recon_set_units.argtypes = [ct.c_void_p, P(ct.c_void_p), P(ct.c_void_p), ct.c_size_t] + E
    L.mrs_recon_launch.argtypes = [ct.c_void_p, ct.c_void_p] + E
    L.mrs_reconstruct_coils.argtypes = [P(abi.C64), ct.c_size_t, ct.c_size_t, ct.c_size_t, P(abi.Acquisition),
                                        P(abi.C64), P(abi.C64)] + E
    L.mrs_grappa_fit.argtypes = [ct.c_void_p, ct.c_size_t, ct.c_size_t, ct.c_size_t, ct.c_size_t, ct.c_size_t,
                                 P(ct.c_uint8), ct.c_void_p, ct.c_void_p] + E
    L.mrs_grappa_fit_noalloc.argtypes = [ct.c_void_p, ct.c_size_t, ct.c_size_t, ct.c_size_t, ct.c_size_t, ct.c_size_t,
                                 P(ct.c_uint8), ct.c_void_p, ct.c_void_p] + E
    L.mrs_grappa_apply.argtypes = [P(ct.c_void_p), ct.c_size_t, ct.c_size_t, ct.c_size_t, ct.c_size_t, ct.c_size_t,
                                   P(ct.c_uint8), ct.c_void_p, ct.c_void_p] + E
    L.mrs_simulate_slice.argtypes = [P(abi.SliceInput), P(abi.Acquisition), P(ct.c_float)] + E
    L.mrs_partition_noise.argtypes = [ct.c_void_p, ct.c_size_t, ct.c_size_t, ct.c_size_t, ct.c_size_t,
                                      P(abi.Acquisition), ct.c_uint64, ct.c_uint64, ct.c_void_p] + E
    L.mrs_plan_set_zstage.argtypes = [ct.c_void_p, P(abi.ZSetDesc), ct.c_size_t] + E
    L.mrs_series_epilogue.argtypes = [ct.c_void_p, ct.c_size_t, ct.c_size_t, ct.c_void_p, ct.c_uint64, ct.c_int32,
                                      ct.c_void_p, ct.c_void_p, ct.c_void_p] + E
    L.mrs_nufft1.argtypes = [ct.c_void_p, ct.c_size_t, ct.c_void_p, ct.c_size_t, ct.c_size_t, ct.c_size_t, ct.c_size_t,
                             ct.c_long, ct.c_size_t, ct.c_double, ct.c_void_p, ct.c_void_p] + E
    L.mrs_build_coil_kspaces_nufft.argtypes = [P(abi.SliceInput), P(abi.Acquisition), P(abi.C64)] + E
    _lib = L
    return L


def _check(rc, buf):
    if rc != 0:
        raise MrsError(rc, buf.value.decode(errors="replace"))


def _errbuf():
    return ct.create_string_buffer(512)


# ----------------------------------------------------------------------------- host tables

def line_timing(acq: abi.Acquisition, nx: int, ny: int):
    """LineTiming::for_acquisition (src/kspace.rs:310-326): (t_ms, trf_ms, tread_ms, polarity)."""
    t, trf, tread = np.zeros(ny), np.zeros(ny), np.zeros(ny)
    pol = np.zeros(ny, dtype=np.int8)
    rc = lib().mrs_line_timing_for_acquisition(ct.byref(acq), nx, ny, abi.dptr(t), abi.dptr(trf), abi.dptr(tread),
                                               pol.ctypes.data_as(ct.POINTER(ct.c_int8)))
    if rc:
        raise MrsError(rc, "invalid acquisition")
    return t, trf, tread, pol


def sampling_mask(nx: int, ny: int, acq: abi.Acquisition) -> np.ndarray:
    m = np.zeros(nx * ny, dtype=np.uint8)
    lib().mrs_sampling_mask(nx, ny, ct.byref(acq), m.ctypes.data_as(ct.POINTER(ct.c_uint8)))
    return m.astype(bool)


def ring_coil_maps(n_coils: int, snx: int, sny: int, nx: int, ny: int, coil0: int = 0, n: Optional[int] = None,
                   scale: float = 1.0) -> np.ndarray:
    """coil_sensitivity (src/kspace.rs:226-237) on the sim grid, (n, snx*sny) complex64."""
    n = n_coils if n is None else n
    out = np.zeros((n, snx * sny), dtype=np.complex64)
    rc = lib().mrs_ring_coil_maps(coil0, n, n_coils, snx, sny, nx, ny, scale,
                                  out.ctypes.data_as(ct.POINTER(abi.C32)))
    if rc:
        raise MrsError(rc, "invalid coil map request")
    return out


def _timing_struct(timing, ny):
    if timing is None:
        return None, []
    keep = [np.ascontiguousarray(v, dtype=np.float64) for v in timing[:3]]
    pol = np.ascontiguousarray(timing[3], dtype=np.int8)
    keep.append(pol)
    lt = abi.LineTiming(abi.dptr(keep[0]), abi.dptr(keep[1]), abi.dptr(keep[2]),
                        pol.ctypes.data_as(ct.POINTER(ct.c_int8)), len(keep[0]))
    return lt, keep


# ----------------------------------------------------------------------------- drop-in calls

def build_coil_kspace(b: abi.SliceInputBuilder, acq: abi.Acquisition, coil: int = 0, ncoils: int = 1,
                      timing=None, log_amp=None) -> np.ndarray:
    """One coil's acquired k-space (nx*ny complex128, layout kx + nx*ky) computed on the GPU."""
    nx, ny = b.acq_matrix
    out = np.zeros(nx * ny, dtype=np.complex128)
    lt, keep = _timing_struct(timing, ny)
    la = np.ascontiguousarray(log_amp, dtype=np.float64) if log_amp is not None else None
    err = _errbuf()
    rc = lib().mrs_build_coil_kspace(ct.byref(b.inp), ct.byref(acq), coil, ncoils,
                                     ct.byref(lt) if lt is not None else None,
                                     abi.dptr(la) if la is not None else None,
                                     out.ctypes.data_as(ct.POINTER(abi.C64)), err, len(err))
    _check(rc, err)
    return out


def build_coil_kspaces(b: abi.SliceInputBuilder, acq: abi.Acquisition, timing=None, log_amp=None) -> np.ndarray:
    """All coils of one slice in one launch: (n_coils, nx*ny) complex128."""
    nx, ny = b.acq_matrix
    nc = max(int(acq.n_coils), 1)
    out = np.zeros((nc, nx * ny), dtype=np.complex128)
    lt, keep = _timing_struct(timing, ny)
    la = np.ascontiguousarray(log_amp, dtype=np.float64) if log_amp is not None else None
    err = _errbuf()
    rc = lib().mrs_build_coil_kspaces(ct.byref(b.inp), ct.byref(acq), ct.byref(lt) if lt is not None else None,
                                      abi.dptr(la) if la is not None else None,
                                      out.ctypes.data_as(ct.POINTER(abi.C64)), err, len(err))
    _check(rc, err)
    return out


def probe_sincos(cyc: np.ndarray, mode: int = 0):
    cyc = np.ascontiguousarray(cyc, dtype=np.float64)
    s = np.zeros(cyc.size, dtype=np.float32)
    c = np.zeros(cyc.size, dtype=np.float32)
    err = _errbuf()
    rc = lib().mrs_probe_sincos(abi.dptr(cyc), abi.fptr(s), abi.fptr(c), cyc.size, mode, err, len(err))
    _check(rc, err)
    return s, c


# ----------------------------------------------------------------------------- batched plan

def grase_block(ny: int, ky_segments: int, t_line_ms: float, reverse_phase: bool = False):
    """mrs_grase_block: (t_ms[ny], polarity[ny]) of a GRASE / 3D-GE block."""
    t = np.zeros(ny)
    pol = np.zeros(ny, dtype=np.int8)
    err = _errbuf()
    _check(lib().mrs_grase_block(ny, ky_segments, t_line_ms, 1 if reverse_phase else 0, abi.dptr(t),
                                 pol.ctypes.data_as(ct.POINTER(ct.c_int8)), err, len(err)), err)
    return t, pol


def ge3d_lines(ny: int, nz: int, ky_segments: int, t_line_ms: float, reverse_phase: bool, kz_segments: int,
               centric: bool):
    """mrs_ge3d_lines: dict of shot[p*ny+ky], excitation[p*ny+ky], n_shots, n_exc, e_c."""
    shot = np.zeros(ny * nz, dtype=np.int32)
    exc = np.zeros(ny * nz, dtype=np.int32)
    ns, ne, ec = ct.c_size_t(), ct.c_size_t(), ct.c_size_t()
    err = _errbuf()
    P = ct.POINTER(ct.c_int32)
    _check(lib().mrs_ge3d_lines(ny, nz, ky_segments, t_line_ms, 1 if reverse_phase else 0, kz_segments,
                                1 if centric else 0, shot.ctypes.data_as(P), exc.ctypes.data_as(P), ct.byref(ns),
                                ct.byref(ne), ct.byref(ec), err, len(err)), err)
    return {"shot": shot, "excitation": exc, "n_shots": ns.value, "n_exc": ne.value, "e_c": ec.value}


def check_ge3d_timing(echo_times_ms, exc_spacing_ms, excitation_time_ms, ny, ky_segments, t_line_ms, reverse_phase,
                      n_exc, t_exc_ms, tr_ms):
    """mrs_check_ge3d_timing: raises MrsError with the reference's text when the train does not fit."""
    tes = np.ascontiguousarray(echo_times_ms, dtype=np.float64)
    err = _errbuf()
    _check(lib().mrs_check_ge3d_timing(abi.dptr(tes), len(tes), exc_spacing_ms, excitation_time_ms, ny, ky_segments,
                                       t_line_ms, 1 if reverse_phase else 0, n_exc, t_exc_ms, tr_ms, err, len(err)),
           err)


def reconstruct_coils(coil_kspace: np.ndarray, nx: int, ny: int, acq: abi.Acquisition,
                      grappa_weights: Optional[np.ndarray] = None) -> np.ndarray:
    """mrs_reconstruct_coils (reconstruct_coils, src/kspace.rs:1290-1298): (C, nx*ny) complex -> (ny*nx,)
    complex128 combined image, x + nx*y, computed on the GPU (GRAPPA first under acceleration)."""
    k = np.ascontiguousarray(coil_kspace, dtype=np.complex128)
    out = np.zeros(nx * ny, dtype=np.complex128)
    w = None if grappa_weights is None else np.ascontiguousarray(grappa_weights, dtype=np.complex128)
    err = _errbuf()
    _check(lib().mrs_reconstruct_coils(k.ctypes.data_as(ct.POINTER(abi.C64)), k.shape[0], nx, ny, ct.byref(acq),
                                       None if w is None else w.ctypes.data_as(ct.POINTER(abi.C64)),
                                       out.ctypes.data_as(ct.POINTER(abi.C64)), err, len(err)), err)
    return out


def grappa_fit(ref_kspace, accel: int, acs_lines: int, lines: np.ndarray, stream=None):
    """mrs_grappa_fit: ref_kspace (C, ny, nx) complex64 device tensor, lines (ny,) flags of the reference
    scan -> weights (accel-1, C, 6C) complex128 device tensor."""
    import torch
    C, ny, nx = ref_kspace.shape
    _check_dev("ref_kspace", ref_kspace, torch.complex64, C * ny * nx)
    w = torch.zeros((max(accel - 1, 1), C, 6 * C), dtype=torch.complex128, device=ref_kspace.device)
    fl = np.ascontiguousarray(lines, dtype=np.uint8)
    stream = stream or torch.cuda.current_stream()
    err = _errbuf()
    _check(lib().mrs_grappa_fit(ref_kspace.data_ptr(), C, nx, ny, accel, acs_lines,
                                fl.ctypes.data_as(ct.POINTER(ct.c_uint8)), w.data_ptr(),
                                ct.c_void_p(stream.cuda_stream), err, len(err)), err)
    return w


def grappa_fit_noalloc(ref_kspace, accel: int, acs_lines: int, lines: np.ndarray, weights, stream=None):
    """mrs_grappa_fit_noalloc: the fit into the caller's `weights` (accel-1, C, 6C) complex128 device
    tensor, with the library's fixed device scratch (no allocation by the library)."""
    import torch
    C, ny, nx = ref_kspace.shape
    _check_dev("ref_kspace", ref_kspace, torch.complex64, C * ny * nx)
    _check_dev("weights", weights, torch.complex128, max(accel - 1, 1) * C * 6 * C)
    fl = np.ascontiguousarray(lines, dtype=np.uint8)
    stream = stream or torch.cuda.current_stream()
    err = _errbuf()
    _check(lib().mrs_grappa_fit_noalloc(ref_kspace.data_ptr(), C, nx, ny, accel, acs_lines,
                                        fl.ctypes.data_as(ct.POINTER(ct.c_uint8)), weights.data_ptr(),
                                        ct.c_void_p(stream.cuda_stream), err, len(err)), err)


def grappa_apply(kspaces, accel: int, acquired: np.ndarray, weights, stream=None):
    """mrs_grappa_apply: synthesize the missing lines of each (C, ny, nx) complex64 device tensor in place."""
    import torch
    n = len(kspaces)
    C, ny, nx = kspaces[0].shape
    ptrs = (ct.c_void_p * max(n, 1))()
    for i, k in enumerate(kspaces):
        _check_dev(f"kspace {i}", k, torch.complex64, C * ny * nx)
        ptrs[i] = k.data_ptr()
    _check_dev("weights", weights, torch.complex128, max(accel - 1, 1) * C * 6 * C)
    fl = np.ascontiguousarray(acquired, dtype=np.uint8)
    stream = stream or torch.cuda.current_stream()
    err = _errbuf()
    _check(lib().mrs_grappa_apply(ptrs, n, C, nx, ny, accel, fl.ctypes.data_as(ct.POINTER(ct.c_uint8)),
                                  weights.data_ptr(), ct.c_void_p(stream.cuda_stream), err, len(err)), err)


def partition_noise(k3, acq: abi.Acquisition, volume: int, seed: int, stream=None):
    """mrs_partition_noise: spikes and noise of one volume's acquired 3D k-space, (nz, C, ny, nx)
    complex64 device tensor, in place."""
    import torch
    nz, C, ny, nx = k3.shape
    _check_dev("k3", k3, torch.complex64, nz * C * ny * nx)
    stream = stream or torch.cuda.current_stream()
    err = _errbuf()
    _check(lib().mrs_partition_noise(k3.data_ptr(), nz, C, nx, ny, ct.byref(acq), volume & 0xFFFFFFFFFFFFFFFF,
                                     seed & 0xFFFFFFFFFFFFFFFF, ct.c_void_p(stream.cuda_stream), err, len(err)), err)


def series_epilogue(images, noise_sigma, receiver_seed: int, magnitude_phase: bool, stream=None):
    """mrs_series_epilogue: images (nv, nvox) complex64 device tensor -> two (nvox*nv,) float32 device
    tensors in the layout vox*nv + g: (re, im), or (magnitude, phase) when `magnitude_phase`."""
    import torch
    nv, nvox = images.shape
    _check_dev("images", images, torch.complex64, nv * nvox)
    _check_dev("noise_sigma", noise_sigma, torch.float32, nvox)
    o0 = torch.empty(nv * nvox, dtype=torch.float32, device=images.device)
    o1 = torch.empty_like(o0)
    stream = stream or torch.cuda.current_stream()
    err = _errbuf()
    _check(lib().mrs_series_epilogue(images.data_ptr(), nv, nvox, _ptr(noise_sigma), receiver_seed & 0xFFFFFFFFFFFFFFFF,
                                     1 if magnitude_phase else 0, o0.data_ptr(), o1.data_ptr(),
                                     ct.c_void_p(stream.cuda_stream), err, len(err)), err)
    return o0, o1


def nufft1(pos, weights, n: int, k_lo: int, n_out: int, tol: float, stream=None):
    """mrs_nufft1 (Nufft1, src/nufft.rs:25-113) in FP64: pos (pos_cols, nsrc) float64 and weights
    (ncols, nw, nsrc) complex128 device tensors -> (ncols, nw, n_out) complex128."""
    import torch
    pos_cols, nsrc = pos.shape
    ncols, nw, nsrc_w = weights.shape
    if nsrc_w != nsrc:
        raise ValueError("weights and positions disagree on the number of sources")
    _check_dev("pos", pos, torch.float64, pos_cols * nsrc)
    _check_dev("weights", weights, torch.complex128, ncols * nw * nsrc)
    out = torch.empty((ncols, nw, n_out), dtype=torch.complex128, device=weights.device)
    stream = stream or torch.cuda.current_stream()
    err = _errbuf()
    _check(lib().mrs_nufft1(pos.data_ptr(), pos_cols, weights.data_ptr(), nsrc, nw, ncols, n, k_lo, n_out, tol,
                            out.data_ptr(), ct.c_void_p(stream.cuda_stream), err, len(err)), err)
    return out


def build_coil_kspaces_nufft(b: abi.SliceInputBuilder, acq: abi.Acquisition) -> np.ndarray:
    """mrs_build_coil_kspaces_nufft: every coil's k-space as the `kspace` feature build computes it
    (NUFFT y-stage, FP64), (n_coils, nx*ny) complex128."""
    nx, ny = b.acq_matrix
    nc = max(int(acq.n_coils), 1)
    out = np.zeros(nc * nx * ny, dtype=np.complex128)
    err = _errbuf()
    _check(lib().mrs_build_coil_kspaces_nufft(ct.byref(b.inp), ct.byref(acq), out.ctypes.data_as(ct.POINTER(abi.C64)),
                                              err, len(err)), err)
    return out.reshape(nc, nx * ny)


def simulate_slice(b: abi.SliceInputBuilder, acq: abi.Acquisition) -> np.ndarray:
    """mrs_simulate_slice (simulate_slice, src/kspace.rs:1229-1235) on the GPU: (nx*ny,) complex64 image."""
    nx, ny = b.acq_matrix
    out = np.zeros(2 * nx * ny, dtype=np.float32)
    err = _errbuf()
    _check(lib().mrs_simulate_slice(ct.byref(b.inp), ct.byref(acq), abi.fptr(out), err, len(err)), err)
    return out[0::2] + 1j * out[1::2]


class Recon:
    """Batched reconstruction of one geometry (mrs_recon_*): window, centred inverse 2D DFT per coil,
    Roemer combine.  `run` takes lists of device tensors: k-spaces (C, ny, nx) complex64 -> images
    (ny, nx) complex64."""

    def __init__(self, nx: int, ny: int, n_coils: int, acq: abi.Acquisition, sensitivities=None):
        self.nx, self.ny, self.C = nx, ny, n_coils
        sens = None
        if sensitivities is not None:
            sens = np.ascontiguousarray(sensitivities, dtype=np.float32)
        h = ct.c_void_p()
        err = _errbuf()
        _check(lib().mrs_recon_create(nx, ny, n_coils, ct.byref(acq),
                                      None if sens is None else sens.ctypes.data_as(ct.POINTER(ct.c_float)),
                                      ct.byref(h), err, len(err)), err)
        self._h = h

    def bind(self, kspaces, images):
        """Upload the unit list (k-spaces (C, ny, nx) and images (ny, nx), complex64 device tensors)."""
        import torch
        n = len(kspaces)
        if len(images) != n:
            raise ValueError("one image per k-space")
        kp, ip = (ct.c_void_p * max(n, 1))(), (ct.c_void_p * max(n, 1))()
        for i, (k, im) in enumerate(zip(kspaces, images)):
            _check_dev(f"kspace {i}", k, torch.complex64, self.C * self.ny * self.nx)
            _check_dev(f"image {i}", im, torch.complex64, self.ny * self.nx)
            kp[i], ip[i] = k.data_ptr(), im.data_ptr()
        self._bound = (list(kspaces), list(images))
        err = _errbuf()
        _check(lib().mrs_recon_set_units(self._h, kp, ip, n, err, len(err)), err)

    def launch(self, stream=None):
        """Run the bound units asynchronously on `stream` (torch's current stream by default)."""
        import torch
        if stream is None:
            stream = torch.cuda.current_stream()
        err = _errbuf()
        _check(lib().mrs_recon_launch(self._h, ct.c_void_p(stream.cuda_stream), err, len(err)), err)

    def run(self, kspaces, images, stream=None):
        """bind + launch, then wait for the stream."""
        import torch
        self.bind(kspaces, images)
        self.launch(stream)
        (stream or torch.cuda.current_stream()).synchronize()

    def close(self):
        if getattr(self, "_h", None):
            lib().mrs_recon_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class ZSet:
    """One set of the batched partition transform (torch complex64 / float32 device tensors).

    forward: out[p] (+)= sum_c w(p, sy, c) sum_z zk[p][z] inp[c][z];  adjoint: out[c][z] (+)= ... (see
    include/mrsim_gpu.h).  ``inp``/``out`` hold [n_comp][nz][n_coils][ny][nx] on the compartment side and
    [nz][n_coils][ny][nx] on the other; ``weights`` is [nz][ky_segments][n_comp] or None.
    """
    inp: object
    out: object
    nz: int
    n_coils: int
    ny: int
    nx: int
    n_comp: int = 1
    ky_segments: int = 1
    weights: object = None
    accumulate: bool = False
    comp_stride: Optional[int] = None


def _zset_array(sets):
    arr = (abi.ZSetDesc * max(len(sets), 1))()
    for i, z in enumerate(sets):
        d = arr[i]
        d.inp, d.out = _ptr(z.inp), _ptr(z.out)
        d.weights = _ptr(z.weights)
        d.comp_stride = z.comp_stride if z.comp_stride is not None else z.nz * z.n_coils * z.ny * z.nx
        d.nz, d.n_comp, d.n_coils, d.ny, d.nx, d.ky_segments = z.nz, z.n_comp, z.n_coils, z.ny, z.nx, z.ky_segments
        d.accumulate = 1 if z.accumulate else 0
    return arr


def zstage_check(sets: Sequence[ZSet], adjoint: bool = False):
    """Raise ValueError unless every set's tensors are device tensors of the set's shape."""
    import torch
    for i, z in enumerate(sets):
        ncols = z.n_coils * z.ny * z.nx
        comp_side = z.n_comp * (z.comp_stride or z.nz * ncols)
        _check_dev(f"zset {i} inp", z.inp, torch.complex64, None)
        _check_dev(f"zset {i} out", z.out, torch.complex64, None)
        need_in = comp_side if not adjoint else z.nz * ncols
        need_out = z.nz * ncols if not adjoint else comp_side
        if z.inp.numel() < need_in or z.out.numel() < need_out:
            raise ValueError(f"zset {i}: buffers smaller than the set's shape")
        if z.weights is not None:
            _check_dev(f"zset {i} weights", z.weights, torch.float32, z.nz * z.ky_segments * z.n_comp)


def zstage(sets: Sequence[ZSet], adjoint: bool = False, stream=None):
    """mrs_zstage: run the sets on `stream` (torch's current stream by default) and wait."""
    import torch
    zstage_check(sets, adjoint)
    arr = _zset_array(sets)
    if stream is None:
        stream = torch.cuda.current_stream()
    err = _errbuf()
    _check(lib().mrs_zstage(arr, len(sets), 1 if adjoint else 0, ct.c_void_p(stream.cuda_stream), err, len(err)), err)


class ZPlan:
    """mrs_zplan: a set list whose descriptors stay on the device; run() enqueues it without waiting.
    The sets' tensors are kept referenced for the z-plan's life."""

    def __init__(self, sets: Sequence[ZSet], adjoint: bool = False):
        zstage_check(sets, adjoint)
        self._keep = list(sets)
        arr = _zset_array(sets)
        h = ct.c_void_p()
        err = _errbuf()
        _check(lib().mrs_zplan_create(arr, len(sets), 1 if adjoint else 0, ct.byref(h), err, len(err)), err)
        self._h = h

    def run(self, stream=None):
        import torch
        if stream is None:
            stream = torch.cuda.current_stream()
        err = _errbuf()
        _check(lib().mrs_zplan_run(self._h, ct.c_void_p(stream.cuda_stream), err, len(err)), err)

    def close(self):
        if self._h:
            lib().mrs_zplan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class Group:
    """Geometry and protocol shared by a set of units."""
    snx: int
    sny: int
    nx: int
    ny: int
    n_coils: int
    n_comp: int
    acq: abi.Acquisition
    timing: Optional[tuple] = None
    complex_images: bool = False


@dataclass
class Slice:
    """Static per-slice inputs: torch device tensors (flattened sim-grid planes)."""
    group: int
    fmap: object                     # (P,) float32
    coil_maps: object                # (C, P) complex64
    phase0: object = None            # (P,) float64
    t2: object = None                # (Q, P) float32 maps
    t2_uniform: Sequence[float] = (100.0,) * abi.MRS_MAX_COMP
    t_inhom: object = None           # (Q, P) float32 maps
    t_inhom_uniform: Optional[Sequence[float]] = None
    log_amp: object = None           # (P,) float64
    z: int = 0
    nz: int = 1
    eddy_drive: Optional[Sequence[float]] = None


@dataclass
class Unit:
    """One (slice, volume, echo) forward: images (Q, P) float32 or complex64 -> out (C, ny, nx) complex64."""
    slice: int
    images: object
    out: object
    slice_seed: int = 0
    coil0: int = 0


def _ptr(t):
    return None if t is None else ct.c_void_p(t.data_ptr())


def _check_dev(name, t, dtype, numel=None):
    import torch
    if t is None:
        return
    if not isinstance(t, torch.Tensor) or t.device.type != "cuda":
        raise TypeError(f"{name} must be a CUDA (HIP) tensor")
    if t.dtype != dtype:
        raise TypeError(f"{name} must be {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if numel is not None and t.numel() != numel:
        raise ValueError(f"{name} has {t.numel()} elements, expected {numel}")


class Plan:
    """Batched forward (and adjoint) of many units in one launch per stage."""

    def __init__(self, groups: Sequence[Group], slices: Sequence[Slice], units: Sequence[Unit],
                 precision: int = abi.PREC_SPLIT_BF16):
        self.groups, self.slices, self.units = list(groups), list(slices), list(units)
        self.precision = precision
        self._keep = []
        gd = (abi.GroupDesc * len(groups))()
        for i, g in enumerate(groups):
            gd[i].snx, gd[i].sny, gd[i].nx, gd[i].ny = g.snx, g.sny, g.nx, g.ny
            gd[i].n_coils, gd[i].n_compartments = g.n_coils, g.n_comp
            gd[i].complex_images = 1 if g.complex_images else 0
            gd[i].acq = g.acq
            if g.timing is not None:
                lt, keep = _timing_struct(g.timing, g.ny)
                self._keep += [lt, keep]
                gd[i].timing = ct.pointer(lt)
        sd, ud = self._descs()
        h = ct.c_void_p()
        err = _errbuf()
        rc = lib().mrs_plan_create(gd, len(groups), sd, len(slices), ud, len(units), precision, ct.byref(h), err,
                                   len(err))
        _check(rc, err)
        self._h = h

    def _descs(self):
        import torch
        sd = (abi.SliceDesc * len(self.slices))()
        for i, s in enumerate(self.slices):
            g = self.groups[s.group]
            P = g.snx * g.sny
            _check_dev(f"slice {i} fmap", s.fmap, torch.float32, P)
            _check_dev(f"slice {i} coil_maps", s.coil_maps, torch.complex64, g.n_coils * P)
            _check_dev(f"slice {i} phase0", s.phase0, torch.float64, P)
            _check_dev(f"slice {i} t2", s.t2, torch.float32, g.n_comp * P)
            _check_dev(f"slice {i} t_inhom", s.t_inhom, torch.float32, g.n_comp * P)
            _check_dev(f"slice {i} log_amp", s.log_amp, torch.float64, P)
            d = sd[i]
            d.group = s.group
            d.fmap, d.coil_maps = _ptr(s.fmap), _ptr(s.coil_maps)
            d.phase0, d.t2, d.t_inhom, d.log_amp = _ptr(s.phase0), _ptr(s.t2), _ptr(s.t_inhom), _ptr(s.log_amp)
            for q in range(min(len(s.t2_uniform), abi.MRS_MAX_COMP)):
                d.t2_uniform[q] = s.t2_uniform[q]
            if s.t_inhom_uniform is not None:
                d.has_t_inhom = 1
                for q in range(min(len(s.t_inhom_uniform), abi.MRS_MAX_COMP)):
                    d.t_inhom_uniform[q] = s.t_inhom_uniform[q]
            if s.t_inhom is not None:
                d.has_t_inhom = 1
            d.z, d.nz = s.z, s.nz
            if s.eddy_drive is not None:
                d.has_eddy_drive = 1
                for k in range(3):
                    d.eddy_drive[k] = float(s.eddy_drive[k])
        ud = (abi.UnitDesc * len(self.units))()
        for i, u in enumerate(self.units):
            g = self.groups[self.slices[u.slice].group]
            P = g.snx * g.sny
            _check_dev(f"unit {i} images", u.images, torch.complex64 if g.complex_images else torch.float32,
                       g.n_comp * P)
            _check_dev(f"unit {i} out", u.out, torch.complex64, g.n_coils * g.ny * g.nx)
            ud[i].slice = u.slice
            ud[i].images, ud[i].out = _ptr(u.images), _ptr(u.out)
            ud[i].slice_seed, ud[i].coil0 = u.slice_seed, u.coil0
        self._sd, self._ud = sd, ud
        return sd, ud

    def set_pointers(self, slices=None, units=None):
        if slices is not None:
            self.slices = list(slices)
        if units is not None:
            self.units = list(units)
        sd, ud = self._descs()
        err = _errbuf()
        _check(lib().mrs_plan_set_pointers(self._h, sd, ud, err, len(err)), err)

    @staticmethod
    def _stream(stream):
        import torch
        if stream is None:
            stream = torch.cuda.current_stream()
        return ct.c_void_p(stream.cuda_stream)

    def forward(self, stream=None):
        err = _errbuf()
        _check(lib().mrs_plan_forward(self._h, self._stream(stream), err, len(err)), err)

    def forward_graph(self, stream=None):
        err = _errbuf()
        _check(lib().mrs_plan_forward_graph(self._h, self._stream(stream), err, len(err)), err)

    def vjp(self, grad_out, grad_images=None, grad_fmap=None, grad_t_inhom=None, stream=None, after_forward=False):
        """after_forward: the slice inputs are unchanged since this plan's last forward on `stream`;
        the adjoint reuses that forward's slice planes (mrs_plan_vjp_after_forward)."""
        self.vjp_arrays(self.pointer_arrays(grad_out, grad_images, grad_fmap, grad_t_inhom), stream, after_forward)

    def pointer_arrays(self, grad_out, grad_images=None, grad_fmap=None, grad_t_inhom=None):
        """The per-unit device pointers of vjp's tensors as ctypes arrays (for vjp_arrays)."""
        n = len(self.units)

        def arr(ts):
            a = (ct.c_void_p * n)()
            if ts is not None:
                for i, t in enumerate(ts):
                    a[i] = None if t is None else t.data_ptr()
            return a

        return arr(grad_out), arr(grad_images), arr(grad_fmap), arr(grad_t_inhom)

    def vjp_arrays(self, arrays, stream=None, after_forward=False):
        """vjp on pointer arrays built once by pointer_arrays, for a caller whose tensors stay put."""
        err = _errbuf()
        fn = lib().mrs_plan_vjp_after_forward if after_forward else lib().mrs_plan_vjp
        rc = fn(self._h, *arrays, self._stream(stream), err, len(err))
        _check(rc, err)

    def tune(self, trials=3, stream=None) -> str:
        rep = ct.create_string_buffer(4096)
        rc = lib().mrs_plan_tune(self._h, self._stream(stream), trials, rep, len(rep))
        if rc:
            raise MrsError(rc, "tuning failed")
        return rep.value.decode()

    def set_tiles(self, x_tile, line_tiles, units_per_item=1):
        err = _errbuf()
        _check(lib().mrs_plan_set_tiles(self._h, x_tile, line_tiles, units_per_item, err, len(err)), err)

    def tiles(self):
        a, b, c = ct.c_int32(), ct.c_int32(), ct.c_int32()
        lib().mrs_plan_get_tiles(self._h, ct.byref(a), ct.byref(b), ct.byref(c))
        return a.value, b.value, c.value

    def set_basis(self, on: bool):
        """Run the volumes of qualifying slices in the slice-basis y-stage (True) or per unit (False)."""
        err = _errbuf()
        _check(lib().mrs_plan_set_basis(self._h, 1 if on else 0, err, len(err)), err)

    def basis(self):
        """(on, number of slices that qualify for the slice-basis y-stage)."""
        a, b = ct.c_int32(), ct.c_int32()
        lib().mrs_plan_get_basis(self._h, ct.byref(a), ct.byref(b))
        return bool(a.value), b.value

    def set_vjp_basis(self, on: bool):
        """Run the y-adjoint of qualifying slices' volumes in the slice-basis kernel (True) or per unit."""
        err = _errbuf()
        _check(lib().mrs_plan_set_vjp_basis(self._h, 1 if on else 0, err, len(err)), err)

    def vjp_basis(self) -> bool:
        a = ct.c_int32()
        lib().mrs_plan_get_vjp_basis(self._h, ct.byref(a))
        return bool(a.value)

    def set_zstage(self, sets: Sequence["ZSet"]):
        """Attach partition transforms that run after the 2D forward (and inside its graph)."""
        zstage_check(sets, adjoint=False)
        arr = _zset_array(sets)
        self._zkeep = list(sets)
        err = _errbuf()
        _check(lib().mrs_plan_set_zstage(self._h, arr, len(sets), err, len(err)), err)

    def stage_times(self, iters=5, stream=None):
        ms = (ct.c_double * 5)()
        err = _errbuf()
        _check(lib().mrs_plan_stage_times(self._h, self._stream(stream), iters, ms, err, len(err)), err)
        return {"prep": ms[0], "ystage": ms[1], "xstage": ms[2], "epilogue": ms[3], "zstage": ms[4]}

    def describe(self) -> str:
        buf = ct.create_string_buffer(1 << 16)
        lib().mrs_plan_describe(self._h, buf, len(buf))
        return buf.value.decode()

    def stats(self):
        t, l, u = ct.c_double(), ct.c_double(), ct.c_double()
        lib().mrs_plan_stats(self._h, ct.byref(t), ct.byref(l), ct.byref(u))
        return {"terms": t.value, "lines": l.value, "units": u.value}

    def close(self):
        if getattr(self, "_h", None):
            lib().mrs_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
