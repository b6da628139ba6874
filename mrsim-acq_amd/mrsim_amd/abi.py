"""ctypes mirror of include/mrsim_gpu.h (the C ABI of the k-space forward model).

The structs mirror the reference's `Acquisition` (src/kspace.rs:166-197), `T2Slice`
(src/kspace.rs:380-384), `SliceInput` (src/kspace.rs:458-489) and `LineTiming`
(src/kspace.rs:291-296), field for field, in the header's order.
"""
from __future__ import annotations

import ctypes as ct

MRS_OK = 0
MRS_E_INVALID = -1
MRS_E_DEVICE = -2
MRS_E_NOMEM = -3
MRS_E_UNSUPPORTED = -4
MRS_MAX_COMP = 4
MRS_MAX_COILS = 64

PF_FIBERFOX, PF_CONTIGUOUS, PF_SCANNER = 0, 1, 2
ECHO_SPIN, ECHO_GRADIENT = 0, 1
WINDOW_NONE, WINDOW_TUKEY, WINDOW_HANN, WINDOW_FERMI = 0, 1, 2, 3
PREC_SPLIT_BF16, PREC_FP32 = 0, 1


class C64(ct.Structure):
    _fields_ = [("re", ct.c_double), ("im", ct.c_double)]


class C32(ct.Structure):
    _fields_ = [("re", ct.c_float), ("im", ct.c_float)]


class Acquisition(ct.Structure):
    _fields_ = [
        ("t_line", ct.c_double),
        ("t_echo", ct.c_double),
        ("t_inhom", ct.c_double),
        ("signal_scale", ct.c_double),
        ("reverse_phase", ct.c_int32),
        ("do_distortions", ct.c_int32),
        ("do_relaxation", ct.c_int32),
        ("pf_mode", ct.c_int32),
        ("noise_variance", ct.c_double),
        ("partial_fourier", ct.c_double),
        ("ghost_offset", ct.c_double),
        ("eddy_strength", ct.c_double),
        ("eddy_quad", ct.c_double),
        ("eddy_phase", ct.c_double),
        ("eddy_tau", ct.c_double),
        ("spike_amplitude", ct.c_double),
        ("n_spikes", ct.c_uint64),
        ("window", ct.c_int32),
        ("echo", ct.c_int32),
        ("window_a", ct.c_double),
        ("window_b", ct.c_double),
        ("n_coils", ct.c_uint64),
        ("accel", ct.c_uint64),
        ("acs_lines", ct.c_uint64),
        ("seed", ct.c_uint64),
    ]

    @classmethod
    def default(cls, **kw) -> "Acquisition":
        """`Acquisition::default()` (src/kspace.rs:239-267) with keyword overrides."""
        a = cls(t_line=1.0, t_echo=90.0, t_inhom=50.0, signal_scale=100.0, reverse_phase=0,
                do_distortions=1, do_relaxation=1, pf_mode=PF_FIBERFOX, noise_variance=0.0,
                partial_fourier=1.0, ghost_offset=0.0, eddy_strength=0.0, eddy_quad=0.0,
                eddy_phase=0.0, eddy_tau=70.0, spike_amplitude=1.0, n_spikes=0,
                window=WINDOW_NONE, echo=ECHO_SPIN, window_a=0.0, window_b=0.0, n_coils=1,
                accel=1, acs_lines=24, seed=0)
        for k, v in kw.items():
            if not hasattr(a, k):
                raise AttributeError(f"Acquisition has no field {k!r}")
            setattr(a, k, v)
        return a

    def copy(self, **kw) -> "Acquisition":
        b = Acquisition()
        ct.pointer(b)[0] = self
        for k, v in kw.items():
            setattr(b, k, v)
        return b


class T2Slice(ct.Structure):
    _fields_ = [("map", ct.POINTER(ct.c_float)), ("value", ct.c_float)]


class SliceInput(ct.Structure):
    _fields_ = [
        ("compartments", ct.POINTER(ct.POINTER(ct.c_float))),
        ("n_compartments", ct.c_size_t),
        ("t2", ct.POINTER(T2Slice)),
        ("t_inhom", ct.POINTER(T2Slice)),
        ("fmap", ct.POINTER(ct.c_float)),
        ("phase0", ct.POINTER(ct.c_double)),
        ("sim", ct.c_size_t * 2),
        ("acq_matrix", ct.c_size_t * 2),
        ("z", ct.c_size_t),
        ("nz", ct.c_size_t),
        ("has_eddy_drive", ct.c_int32),
        ("eddy_drive", ct.c_double * 3),
        ("has_prep_drive", ct.c_int32),
        ("prep_magnitude", ct.c_double),
        ("prep_direction", ct.c_double * 3),
        ("slice_seed", ct.c_uint64),
    ]


class LineTiming(ct.Structure):
    _fields_ = [
        ("t_ms", ct.POINTER(ct.c_double)),
        ("trf_ms", ct.POINTER(ct.c_double)),
        ("tread_ms", ct.POINTER(ct.c_double)),
        ("polarity", ct.POINTER(ct.c_int8)),
        ("ny", ct.c_size_t),
    ]


class GroupDesc(ct.Structure):
    _fields_ = [
        ("snx", ct.c_size_t), ("sny", ct.c_size_t), ("nx", ct.c_size_t), ("ny", ct.c_size_t),
        ("n_coils", ct.c_size_t),
        ("n_compartments", ct.c_size_t),
        ("complex_images", ct.c_int32),
        ("acq", Acquisition),
        ("timing", ct.POINTER(LineTiming)),
    ]


class SliceDesc(ct.Structure):
    _fields_ = [
        ("group", ct.c_int32),
        ("fmap", ct.c_void_p),
        ("phase0", ct.c_void_p),
        ("t2", ct.c_void_p),
        ("t_inhom", ct.c_void_p),
        ("coil_maps", ct.c_void_p),
        ("log_amp", ct.c_void_p),
        ("t2_uniform", ct.c_float * MRS_MAX_COMP),
        ("t_inhom_uniform", ct.c_float * MRS_MAX_COMP),
        ("has_t_inhom", ct.c_int32),
        ("z", ct.c_size_t),
        ("nz", ct.c_size_t),
        ("has_eddy_drive", ct.c_int32),
        ("eddy_drive", ct.c_double * 3),
    ]


class UnitDesc(ct.Structure):
    _fields_ = [
        ("slice", ct.c_int32),
        ("images", ct.c_void_p),
        ("out", ct.c_void_p),
        ("slice_seed", ct.c_uint64),
        ("coil0", ct.c_uint64),
    ]


class ZSetDesc(ct.Structure):
    """mrs_zset: one set of the batched partition transform (device pointers)."""
    _fields_ = [
        ("inp", ct.c_void_p),
        ("out", ct.c_void_p),
        ("weights", ct.c_void_p),
        ("comp_stride", ct.c_size_t),
        ("nz", ct.c_uint32),
        ("n_comp", ct.c_uint32),
        ("n_coils", ct.c_uint32),
        ("ny", ct.c_uint32),
        ("nx", ct.c_uint32),
        ("ky_segments", ct.c_uint32),
        ("accumulate", ct.c_int32),
        ("reserved", ct.c_int32),
    ]


STRUCT_SIZES = {
    "mrs_acquisition": Acquisition,
    "mrs_t2_slice": T2Slice,
    "mrs_slice_input": SliceInput,
    "mrs_line_timing": LineTiming,
    "mrs_group_desc": GroupDesc,
    "mrs_slice_desc": SliceDesc,
    "mrs_unit_desc": UnitDesc,
    "mrs_zset": ZSetDesc,
}


def fptr(a):
    """A float* into a C-contiguous float32 numpy array."""
    return a.ctypes.data_as(ct.POINTER(ct.c_float))


def dptr(a):
    return a.ctypes.data_as(ct.POINTER(ct.c_double))


class SliceInputBuilder:
    """Keeps the numpy arrays alive that a `SliceInput` points into."""

    def __init__(self, comps, t2, fmap, sim, acq_matrix, z=0, nz=1, t_inhom=None, phase0=None,
                 eddy_drive=None, slice_seed=0):
        import numpy as np
        self._keep = []
        comps = [np.ascontiguousarray(c, dtype=np.float32) for c in comps]
        self._keep += comps
        arr = (ct.POINTER(ct.c_float) * len(comps))(*[fptr(c) for c in comps])
        self._keep.append(arr)

        def t2arr(spec):
            out = (T2Slice * len(spec))()
            for i, s in enumerate(spec):
                if isinstance(s, (int, float)):
                    out[i].map = ct.POINTER(ct.c_float)()
                    out[i].value = float(s)
                else:
                    m = np.ascontiguousarray(s, dtype=np.float32)
                    self._keep.append(m)
                    out[i].map = fptr(m)
                    out[i].value = 0.0
            self._keep.append(out)
            return out

        t2a = t2arr(t2)
        tia = t2arr(t_inhom) if t_inhom is not None else None
        fm = np.ascontiguousarray(fmap, dtype=np.float32)
        self._keep.append(fm)
        inp = SliceInput()
        inp.compartments = arr
        inp.n_compartments = len(comps)
        inp.t2 = t2a
        inp.t_inhom = tia if tia is not None else ct.POINTER(T2Slice)()
        inp.fmap = fptr(fm)
        if phase0 is not None:
            p0 = np.ascontiguousarray(phase0, dtype=np.float64)
            self._keep.append(p0)
            inp.phase0 = dptr(p0)
        else:
            inp.phase0 = ct.POINTER(ct.c_double)()
        inp.sim[0], inp.sim[1] = sim
        inp.acq_matrix[0], inp.acq_matrix[1] = acq_matrix
        inp.z, inp.nz = z, nz
        if eddy_drive is not None:
            inp.has_eddy_drive = 1
            for i in range(3):
                inp.eddy_drive[i] = float(eddy_drive[i])
        inp.slice_seed = slice_seed
        self.inp = inp
        self.comps = comps
        self.fmap = fm
        self.sim = tuple(sim)
        self.acq_matrix = tuple(acq_matrix)
