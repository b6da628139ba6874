"""GPU adjoint (VJP) of the forward model: 128^2 -> 64^2, 8 coils, 3 compartments, complex64 images,
T2/T2' maps, a seeded complex-Gaussian cotangent g.  L = Re sum conj(g) K, so K.backward(g) gives dL.

Checks:
  - adjoint identity <A m, g> = <m, A^H g> for the image (relative 1e-5);
  - gradients at sampled voxels against FP64 central differences of the oracle (image, fieldmap, T2');
  - gradcheck-style central differences through the GPU forward along random directions
    (step 1e-3 relative), tolerance 2e-3 relative.
"""
import numpy as np
import pytest

from mrsim_amd import abi, workloads
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from mrsim_amd.autograd import KspaceModel
    w = workloads.make("grad_128")
    sd = w.slices[0]
    dev = "cuda"
    model = KspaceModel(w.snx, w.sny, w.nx, w.ny, torch.from_numpy(sd.coil_maps).to(dev), w.acq, w.n_comp,
                        t2=torch.from_numpy(sd.t2).to(dev), complex_images=True)
    img = torch.from_numpy(w.units[0][1]).to(dev)
    fmap = torch.from_numpy(sd.fmap).to(dev)
    ti = torch.from_numpy(sd.t_inhom).to(dev)
    rng = np.random.default_rng(7)
    g = (rng.standard_normal((w.n_coils, w.ny, w.nx)) + 1j * rng.standard_normal((w.n_coils, w.ny, w.nx)))
    g = g.astype(np.complex64)
    return w, model, img, fmap, ti, torch.from_numpy(g).to(dev)


def _grads(model, img, fmap, ti, g):
    import torch
    a, b, c = (t.clone().requires_grad_(True) for t in (img, fmap, ti))
    k = model(a, b, c)
    k.backward(g)
    torch.cuda.synchronize()
    return k.detach(), a.grad, b.grad, c.grad


def _oracle_loss(w, img, fmap, ti, g):
    sd = w.slices[0]
    k = O.forward_direct(w.snx, w.sny, w.nx, w.ny, w.n_coils, w.n_comp, w.acq, img, fmap, sd.coil_maps, t2=sd.t2,
                         t_inhom=ti, complex_images=True)
    return float(np.real(np.vdot(g, k)))


def test_forward_matches_oracle(setup):
    w, model, img, fmap, ti, g = setup
    k = model.forward_raw(img, fmap, ti).cpu().numpy()
    want = O.forward_workload_unit(w, 0)
    assert np.linalg.norm(k - want) / np.linalg.norm(want) <= 1e-5


def test_vjp_after_forward_reuses_or_rebuilds_the_slice_planes(setup):
    # After a forward on the same inputs the adjoint may read that forward's slice planes; once a slice
    # input changes it must rebuild them.  Both must give the plain adjoint's gradients bit for bit, and
    # the two fieldmaps give different gradients, so stale planes could not pass.
    import torch
    w, model, img, fmap, ti, g = setup
    fmap2 = fmap + 3.0
    want1 = model.vjp_raw(img, fmap, ti, g)
    want2 = model.vjp_raw(img, fmap2, ti, g)
    model.forward_raw(img, fmap, ti)
    got1 = model.vjp_raw(img, fmap, ti, g, after_forward=True)
    got2 = model.vjp_raw(img, fmap2, ti, g, after_forward=True)
    torch.cuda.synchronize()
    for a, b in zip(got1 + got2, want1 + want2):
        assert torch.equal(a, b)
    assert not torch.equal(want1[1], want2[1])


def test_vjps_queued_on_a_side_stream_match_null_stream_results(setup):
    # Five adjoints with alternating cotangents queued back to back on a non-blocking stream, with no
    # synchronisation between them: each call's descriptors (cotangent and gradient pointers) go to
    # their own slot of the plan's descriptor ring, so no earlier call's kernels read a later call's
    # pointers.  Each result equals the same adjoint run alone on the current stream, bit for bit.
    import torch
    w, model, img, fmap, ti, g = setup
    g2 = (g * (0.5 - 1.5j)).contiguous()
    want1 = model.vjp_raw(img, fmap, ti, g)
    torch.cuda.synchronize()
    want2 = model.vjp_raw(img, fmap, ti, g2)
    torch.cuda.synchronize()
    assert not torch.equal(want1[0], want2[0])
    s = torch.cuda.Stream()
    torch.cuda.synchronize()
    with torch.cuda.stream(s):
        got = [model.vjp_raw(img, fmap, ti, gg) for gg in (g, g2, g, g2, g)]
    s.synchronize()
    for k, r in enumerate(got):
        want = want1 if k % 2 == 0 else want2
        for a, b in zip(r, want):
            assert torch.equal(a, b), f"queued adjoint {k} differs"


def test_adjoint_identity(setup):
    w, model, img, fmap, ti, g = setup
    k, gi, gf, gt = _grads(model, img, fmap, ti, g)
    lhs = np.vdot(g.cpu().numpy(), k.cpu().numpy())          # <g, A m>
    rhs = np.vdot(gi.cpu().numpy(), img.cpu().numpy())       # <A^H g, m>
    scale = np.linalg.norm(g.cpu().numpy()) * np.linalg.norm(k.cpu().numpy())
    print(f"adjoint identity: {lhs:.6e} vs {rhs:.6e} (|diff| / (|g||Am|) = {abs(lhs - rhs) / scale:.2e})", flush=True)
    assert abs(lhs - rhs) <= 1e-5 * scale


def test_gradients_match_fp64_central_differences(setup):
    w, model, img, fmap, ti, g = setup
    k, gi, gf, gt = _grads(model, img, fmap, ti, g)
    gi, gf, gt = gi.cpu().numpy(), gf.cpu().numpy(), gt.cpu().numpy()
    img_h, fmap_h, ti_h, g_h = img.cpu().numpy(), fmap.cpu().numpy(), ti.cpu().numpy(), g.cpu().numpy()
    P = w.snx * w.sny
    mag = np.abs(img_h).sum(axis=0)
    vox = [int(v) for v in np.argsort(-mag)[[0, 40, 400]]] + [64 * 128 + 64]
    for i in vox:
        # fieldmap (Hz): step 1e-3 relative
        h = 1e-3 * max(abs(float(fmap_h[i])), 1.0)
        fp, fm = fmap_h.copy(), fmap_h.copy()
        fp[i] += h
        fm[i] -= h
        d = (_oracle_loss(w, img_h, fp, ti_h, g_h) - _oracle_loss(w, img_h, fm, ti_h, g_h)) / (fp[i] - fm[i])
        print(f"voxel {i}: dL/dfmap gpu {gf[i]:.6e} fp64 {d:.6e}", flush=True)
        assert abs(gf[i] - d) <= 2e-3 * max(abs(d), 1e-3 * np.abs(gf).max())
        for q in range(w.n_comp):
            h = 1e-3 * float(ti_h[q, i])
            tp, tm = ti_h.copy(), ti_h.copy()
            tp[q, i] += h
            tm[q, i] -= h
            d = (_oracle_loss(w, img_h, fmap_h, tp, g_h) - _oracle_loss(w, img_h, fmap_h, tm, g_h)) / (tp[q, i] - tm[q, i])
            assert abs(gt[q, i] - d) <= 2e-3 * max(abs(d), 1e-3 * np.abs(gt[q]).max()), (i, q, gt[q, i], d)
            # image: linear, so the differences are exact up to FP64 rounding
            e = np.zeros_like(img_h)
            e[q, i] = 1.0
            dre = _oracle_loss(w, img_h + e, fmap_h, ti_h, g_h) - _oracle_loss(w, img_h, fmap_h, ti_h, g_h)
            dim = _oracle_loss(w, img_h + 1j * e, fmap_h, ti_h, g_h) - _oracle_loss(w, img_h, fmap_h, ti_h, g_h)
            want = dre + 1j * dim
            assert abs(gi[q, i] - want) <= 2e-4 * max(abs(want), 1e-3 * np.abs(gi[q]).max()), (i, q, gi[q, i], want)


def test_gradcheck_style_directional_derivatives(setup):
    # The differences are taken through the plain FP32 forward (rel-L2 ~3e-7), whose round-off over a
    # 1e-3 relative step stays well below the tolerance.
    import torch
    from mrsim_amd.autograd import KspaceModel
    w, model, img, fmap, ti, g = setup
    k, gi, gf, gt = _grads(model, img, fmap, ti, g)
    gen = torch.Generator(device="cpu").manual_seed(3)
    sd = w.slices[0]
    fp32 = KspaceModel(w.snx, w.sny, w.nx, w.ny, torch.from_numpy(sd.coil_maps).to("cuda"), w.acq, w.n_comp,
                       t2=torch.from_numpy(sd.t2).to("cuda"), complex_images=True, precision=abi.PREC_FP32)

    def loss(a, b, c):
        return float(torch.real(torch.vdot(g.reshape(-1), fp32.forward_raw(a, b, c).reshape(-1))).item())

    # fieldmap direction, step 1e-3 relative to the map's scale
    v = torch.randn(fmap.shape, generator=gen).to(fmap.device)
    h = 1e-3 * float(fmap.abs().max())
    fd = (loss(img, fmap + h * v, ti) - loss(img, fmap - h * v, ti)) / (2 * h)
    an = float((gf * v).sum())
    print(f"fmap direction: fd {fd:.6e} analytic {an:.6e}", flush=True)
    assert abs(fd - an) <= 2e-3 * abs(an)
    # T2' direction: its signal over a 1e-3 step (~1e-5 of L) is near FP32 round-off in L, so the
    # differences are taken through the FP64 oracle of the same function.
    v = torch.randn(ti.shape, generator=gen)
    h = 1e-3 * float(ti.abs().mean())
    ti_h, v_h = ti.cpu().numpy(), v.numpy()
    img_h, fmap_h, g_h = img.cpu().numpy(), fmap.cpu().numpy(), g.cpu().numpy()
    fd = (_oracle_loss(w, img_h, fmap_h, ti_h + h * v_h, g_h) - _oracle_loss(w, img_h, fmap_h, ti_h - h * v_h, g_h)) / (2 * h)
    an = float((gt.cpu() * v).sum())
    print(f"T2' direction: fd(fp64) {fd:.6e} analytic {an:.6e}", flush=True)
    assert abs(fd - an) <= 2e-3 * abs(an)
    v = (torch.randn(img.shape, generator=gen) + 1j * torch.randn(img.shape, generator=gen)).to(img.device)
    v = v.to(torch.complex64)
    h = 1e-3 * float(img.abs().max())
    fd = (loss(img + h * v, fmap, ti) - loss(img - h * v, fmap, ti)) / (2 * h)
    an = float(torch.real(torch.vdot(gi.reshape(-1), v.reshape(-1))))
    print(f"image direction: fd {fd:.6e} analytic {an:.6e}", flush=True)
    assert abs(fd - an) <= 2e-3 * abs(an)


def test_fieldmap_fit_reduces_the_loss(setup):
    import torch
    w, model, img, fmap, ti, g = setup
    target = model.forward_raw(img, fmap, ti).detach()
    f = (fmap + 5.0).clone().requires_grad_(True)
    opt = torch.optim.Adam([f], lr=0.5)
    losses = []
    for it in range(50):
        opt.zero_grad()
        k = model(img, f, ti)
        loss = (k - target).abs().pow(2).sum()
        loss.backward()
        opt.step()
        losses.append(float(loss))
    print(f"fieldmap fit: loss {losses[0]:.4e} -> {losses[-1]:.4e}", flush=True)
    assert losses[-1] < 0.5 * losses[0]
