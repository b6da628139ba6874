"""torch.autograd wrapper of the GPU forward model and its adjoint (vector-Jacobian product).

``KspaceModel`` fixes one geometry and protocol (a group) and its static per-slice inputs (coil
maps, T2 maps, optional phase0).  Calling it with the compartment images, the fieldmap (Hz) and the
T2' maps (ms) runs ``mrs_plan_forward``; backward runs ``mrs_plan_vjp``, which returns

  d/d images   complex64 (Q, P): sum over lines of D_q conj(P) conj(S) T^H g   (conjugate Wirtinger)
  d/d fmap     float32 (P,):     2 pi Im(sum_q conj(m_q) sum_l tb0_l D_q,l t_l)
  d/d T2'      float32 (Q, P):   Re(conj(m_q) sum_l tau_l D_q,l t_l) / T2'^2

the gradient convention torch uses for complex tensors (grad = dL/dRe + i dL/dIm for real L).
"""
from __future__ import annotations

from typing import Optional

import torch

from . import abi
from .gpu import Group, Plan, Slice, Unit


class KspaceModel:
    def __init__(self, snx, sny, nx, ny, coil_maps: torch.Tensor, acq: abi.Acquisition, n_comp: int,
                 t2: Optional[torch.Tensor] = None, t2_uniform=(100.0,) * abi.MRS_MAX_COMP,
                 complex_images: bool = True, phase0: Optional[torch.Tensor] = None, timing=None,
                 precision: int = abi.PREC_SPLIT_BF16):
        self.geom = (snx, sny, nx, ny)
        self.C = coil_maps.shape[0]
        self.Q = n_comp
        self.cplx = complex_images
        self.group = Group(snx, sny, nx, ny, self.C, n_comp, acq, timing=timing, complex_images=complex_images)
        self.coil_maps = coil_maps.contiguous()
        self.t2 = None if t2 is None else t2.contiguous()
        self.t2_uniform = t2_uniform
        self.phase0 = phase0
        dev = coil_maps.device
        P = snx * sny
        self._fmap0 = torch.zeros(P, dtype=torch.float32, device=dev)
        self._ti0 = torch.full((n_comp, P), 50.0, dtype=torch.float32, device=dev)
        img_dtype = torch.complex64 if complex_images else torch.float32
        self._img0 = torch.zeros((n_comp, P), dtype=img_dtype, device=dev)
        self._out0 = torch.zeros((self.C, ny, nx), dtype=torch.complex64, device=dev)
        self.plan = Plan([self.group], [self._slice(self._fmap0, self._ti0)],
                         [Unit(0, self._img0, self._out0)], precision)

    def _slice(self, fmap, t_inhom):
        return Slice(0, fmap, self.coil_maps, phase0=self.phase0, t2=self.t2, t2_uniform=self.t2_uniform,
                     t_inhom=t_inhom)

    def _bind(self, images, fmap, t_inhom, out):
        # the plan keeps the last bound pointers; the same tensors again need no call
        key = (images.data_ptr(), fmap.data_ptr(), t_inhom.data_ptr(), out.data_ptr())
        if key == getattr(self, "_bound", None):
            return
        self.plan.set_pointers(slices=[self._slice(fmap, t_inhom)], units=[Unit(0, images, out)])
        self._bound = key

    def forward_raw(self, images, fmap, t_inhom):
        # The plan writes into one output buffer of the model's and the caller gets a copy: binding the
        # same pointers call after call leaves the plan's descriptors as they are (a changed pointer
        # rebinds them, which waits for the device).
        self._bind(images.detach().contiguous(), fmap.detach().contiguous(), t_inhom.detach().contiguous(),
                   self._out0)
        self.plan.forward()
        return self._out0.clone()

    def vjp_raw(self, images, fmap, t_inhom, gout, need=(True, True, True), after_forward=False):
        """after_forward: the inputs are unchanged since the last forward_raw (the adjoint reuses its
        slice planes; a changed fieldmap or T2' tensor still rebuilds them)."""
        self._bind(images.detach().contiguous(), fmap.detach().contiguous(), t_inhom.detach().contiguous(),
                   self._out0)
        gi = torch.empty_like(images) if need[0] else None
        gf = torch.empty_like(fmap) if need[1] else None
        gt = torch.empty_like(t_inhom) if need[2] else None
        self.plan.vjp([gout.contiguous()], None if gi is None else [gi], None if gf is None else [gf],
                      None if gt is None else [gt], after_forward=after_forward)
        return gi, gf, gt

    def __call__(self, images, fmap, t_inhom):
        return _KspaceFn.apply(images, fmap, t_inhom, self)


class _KspaceFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, images, fmap, t_inhom, model: KspaceModel):
        ctx.model = model
        ctx.save_for_backward(images, fmap, t_inhom)
        return model.forward_raw(images, fmap, t_inhom)

    @staticmethod
    def backward(ctx, gout):
        images, fmap, t_inhom = ctx.saved_tensors
        need = ctx.needs_input_grad[:3]
        gi, gf, gt = ctx.model.vjp_raw(images, fmap, t_inhom, gout, need)
        return gi, gf, gt, None


KspaceForward = _KspaceFn
