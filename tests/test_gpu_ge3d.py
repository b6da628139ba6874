"""GPU 3D gradient-echo train: the batched partition transform and the whole forward against the
FP64 oracle (oracle/ge3d.py).

- the partition transform (mrs_zstage): forward against nz * centred inverse FFT with per-line
  weights, several compartments, accumulation; the adjoint through <Ax, y> = <x, A^H y>; partition
  counts 2..64 (every table size), ragged column tails;
- the batched forward at reduced sizes (two geometries of uneven size in one plan, two echoes, two
  compartments, z-varying coil maps, 2 ky segments): every (volume, echo) against the oracle,
  rel-L2 <= 1e-5, split-BF16 and plain FP32, eager and graph replay;
- the train's vector-Jacobian product (Ge3dForward.vjp) on the small train with T2' maps: the adjoint
  identity for the images, and the fieldmap, T2' and image gradients at sampled voxels against FP64
  central differences of the oracle's train;
- the full-size train (two volumes, 128^2 -> 64^2 and 192^2 -> 96^2, 24 partitions, 32 coils,
  4 echoes): one coil of each volume at the first and last echo against the oracle (rel-L2 <= 1e-5),
  and for every coil and echo the size-independent round trip: the output divided by its weights and
  taken back through the adjoint transform / nz equals the per-partition 2D k-spaces.
"""
import numpy as np
import pytest

from mrsim_amd import abi, gpu, workloads
from oracle import ge3d as G
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _zref(x, nz, kseg, ny, w):
    """FP64 forward transform: x (Q, nz, C, ny, nx) -> (nz, C, ny, nx), weights (nz, kseg, Q) or None."""
    zk = G.zkernel(nz)
    k = np.einsum("pz,qzcyx->qpcyx", zk, x.astype(np.complex128))
    if w is not None:
        sy = np.arange(ny) % kseg
        k = k * np.transpose(w[:, sy, :], (2, 0, 1))[:, :, None, :, None]
    return k.sum(axis=0)


@pytest.mark.parametrize("nz", [2, 3, 4, 5, 6, 16, 17, 24, 32, 33, 64])
def test_zstage_forward_and_adjoint(torch_dev, nz):
    torch = torch_dev
    rng = np.random.default_rng(100 + nz)
    Q, C, ny, nx, kseg = 2, 5, 12, 11, 4          # 660 columns: two full blocks and a ragged tail
    print(f"zstage nz={nz} Q={Q} C={C} ny={ny} nx={nx}", flush=True)
    x = (rng.standard_normal((Q, nz, C, ny, nx)) + 1j * rng.standard_normal((Q, nz, C, ny, nx))).astype(np.complex64)
    w = rng.uniform(0.3, 1.2, size=(nz, kseg, Q))
    xd = torch.from_numpy(x).cuda()
    wd = torch.from_numpy(w.astype(np.float32).reshape(-1)).cuda()
    out = torch.zeros((nz, C, ny, nx), dtype=torch.complex64, device="cuda")
    gpu.zstage([gpu.ZSet(xd, out, nz, C, ny, nx, n_comp=Q, ky_segments=kseg, weights=wd)])
    want = _zref(x, nz, kseg, ny, w.astype(np.float32).astype(np.float64))
    got = out.cpu().numpy()
    print(f"  forward rel-L2 {_rel(got, want):.2e}", flush=True)
    assert _rel(got, want) <= 1e-5
    # accumulate adds onto what is there
    gpu.zstage([gpu.ZSet(xd, out, nz, C, ny, nx, n_comp=Q, ky_segments=kseg, weights=wd, accumulate=True)])
    assert _rel(out.cpu().numpy(), 2 * want) <= 1e-5
    # adjoint: <A x, y> = <x, A^H y>
    y = (rng.standard_normal((nz, C, ny, nx)) + 1j * rng.standard_normal((nz, C, ny, nx))).astype(np.complex64)
    yd = torch.from_numpy(y).cuda()
    gx = torch.zeros_like(xd)
    gpu.zstage([gpu.ZSet(yd, gx, nz, C, ny, nx, n_comp=Q, ky_segments=kseg, weights=wd)], adjoint=True)
    lhs = np.vdot(y.astype(np.complex128), want)
    rhs = np.vdot(gx.cpu().numpy().astype(np.complex128), x.astype(np.complex128))
    scale = np.linalg.norm(y) * np.linalg.norm(want)
    print(f"  adjoint identity |diff|/(|y||Ax|) = {abs(lhs - rhs) / scale:.2e}", flush=True)
    assert abs(lhs - rhs) <= 1e-5 * scale


def test_zplan_runs_equal_mrs_zstage(torch_dev):
    # a z-plan (descriptors kept on the device, runs without waiting) gives mrs_zstage's results bit for
    # bit, forward and adjoint, over repeated runs with other work in between
    torch = torch_dev
    rng = np.random.default_rng(7)
    nz, Q, C, ny, nx, kseg = 24, 1, 3, 10, 9, 2
    x = (rng.standard_normal((Q, nz, C, ny, nx)) + 1j * rng.standard_normal((Q, nz, C, ny, nx))).astype(np.complex64)
    y = (rng.standard_normal((nz, C, ny, nx)) + 1j * rng.standard_normal((nz, C, ny, nx))).astype(np.complex64)
    wd = torch.from_numpy(rng.uniform(0.3, 1.2, size=nz * kseg * Q).astype(np.float32)).cuda()
    xd, yd = torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda()
    for adjoint, inp, shape in ((False, xd, (nz, C, ny, nx)), (True, yd, (Q, nz, C, ny, nx))):
        want = torch.zeros(shape, dtype=torch.complex64, device="cuda")
        got = torch.zeros_like(want)
        gpu.zstage([gpu.ZSet(inp, want, nz, C, ny, nx, n_comp=Q, ky_segments=kseg, weights=wd)], adjoint=adjoint)
        zp = gpu.ZPlan([gpu.ZSet(inp, got, nz, C, ny, nx, n_comp=Q, ky_segments=kseg, weights=wd)], adjoint=adjoint)
        for _ in range(3):
            got.zero_()
            zp.run()
            torch.cuda.synchronize()
            assert torch.equal(got, want)
        zp.close()


def test_zstage_several_sets_in_one_call(torch_dev):
    torch = torch_dev
    rng = np.random.default_rng(7)
    sets, wants, keep = [], [], []
    for nz, C, ny, nx in [(4, 3, 8, 8), (24, 2, 16, 12), (40, 1, 4, 4)]:
        x = (rng.standard_normal((1, nz, C, ny, nx)) + 1j * rng.standard_normal((1, nz, C, ny, nx))).astype(np.complex64)
        xd = torch.from_numpy(x).cuda()
        out = torch.zeros((nz, C, ny, nx), dtype=torch.complex64, device="cuda")
        keep += [xd, out]
        sets.append(gpu.ZSet(xd, out, nz, C, ny, nx))
        wants.append(_zref(x, nz, 1, ny, None))
    gpu.zstage(sets)
    for s, want in zip(sets, wants):
        assert _rel(s.out.cpu().numpy(), want) <= 1e-5


def _oracle_volume(w, v, e, coils=None):
    maps = v.coil_maps if coils is None else v.coil_maps[:, coils]
    return G.ge3d_kspace(v.snx, v.sny, v.nx, v.ny, v.nz, v.images, v.t2, v.t_inhom, v.fmap, None, maps, w.acq,
                         w.echo_times_ms[e], w.ky_segments, w.t_line_ms, v.weights.astype(np.float32))


@pytest.mark.parametrize("precision", [abi.PREC_SPLIT_BF16, abi.PREC_FP32])
def test_ge3d_small_matches_oracle(torch_dev, precision):
    torch = torch_dev
    w = workloads.make_ge3d("ge3d_small")
    fw, keep = workloads.to_ge3d(w, precision)
    fw.forward()
    torch.cuda.synchronize()
    for vi, v in enumerate(w.volumes):
        for e in range(len(w.echo_times_ms)):
            want = _oracle_volume(w, v, e)
            got = fw.out[vi][e].cpu().numpy()
            r = _rel(got, want)
            print(f"ge3d_small volume {vi} ({v.snx}->{v.nx}, nz {v.nz}) echo {e}: rel-L2 {r:.2e}", flush=True)
            assert r <= 1e-5
    if precision == abi.PREC_SPLIT_BF16:
        eager = [[o.clone() for o in ov] for ov in fw.out]
        for _ in range(3):
            for ov in fw.out:
                for o in ov:
                    o.fill_(0)
            fw.forward_graph()
            torch.cuda.synchronize()
            for a, b in zip(eager, fw.out):
                for x, y in zip(a, b):
                    assert torch.equal(x, y)
    fw.close()


def test_ge3d_refusals(torch_dev):
    torch = torch_dev
    from mrsim_amd.ge3d import Ge3dForward, Ge3dVolume
    P, nz = 24 * 24, 4
    one = torch.ones(nz * P, device="cuda")
    maps = torch.ones((nz, 1, P), dtype=torch.complex64, device="cuda")
    ok = dict(images=[one], t2=[80.0], fmap=torch.zeros(nz * P, device="cuda"), coil_maps=maps,
              weights=torch.ones(nz, device="cuda"))
    acq = abi.Acquisition.default(echo=abi.ECHO_GRADIENT)
    cases = [(dict(), acq.copy(echo=abi.ECHO_SPIN), "gradient-echo acquisition"),
             (dict(), acq.copy(eddy_strength=1.0), "eddy"),
             (dict(), acq.copy(partial_fourier=0.75), "partial Fourier"),
             (dict(images=[one, one]), acq, "compartment images"),
             (dict(images=[None]), acq, "no image but a nonzero weight")]
    for over, a, msg in cases:
        kw = dict(ok, **over)
        with pytest.raises(ValueError, match=msg):
            Ge3dForward([Ge3dVolume(24, 24, 12, 12, nz, **kw)], a, [12.0], 1, 0.4)


@pytest.fixture(scope="module")
def grad_setup(torch_dev):
    # the small train with T2' maps (the gradient is per voxel) and a seeded complex cotangent
    torch = torch_dev
    w = workloads.make_ge3d("ge3d_small")
    rng = np.random.default_rng(31)
    for v in w.volumes:
        v.t_inhom = [rng.uniform(35.0, 55.0, size=(v.nz, v.snx * v.sny)).astype(np.float32) for _ in v.images]
    fw, vols = workloads.to_ge3d(w)
    fw.forward()
    g = [[(rng.standard_normal(o.shape) + 1j * rng.standard_normal(o.shape)).astype(np.complex64) for o in ov]
         for ov in fw.out]
    gd = [[torch.from_numpy(x).cuda() for x in gv] for gv in g]
    grads = fw.vjp(gd)
    torch.cuda.synchronize()
    yield w, fw, vols, g, grads
    fw.close()


def test_ge3d_vjp_adjoint_identity(grad_setup):
    # the train is linear in the images: Re <g, A m> = <A^H g, m> summed over volumes and echoes
    w, fw, vols, g, grads = grad_setup
    lhs, n_out, n_g = 0.0, 0.0, 0.0
    for vi in range(len(vols)):
        for e in range(len(w.echo_times_ms)):
            k = fw.out[vi][e].cpu().numpy().astype(np.complex128)
            lhs += float(np.real(np.vdot(g[vi][e], k)))
            n_out += float(np.sum(np.abs(k) ** 2))
            n_g += float(np.sum(np.abs(g[vi][e]) ** 2))
    rhs = 0.0
    for vi, v in enumerate(vols):
        for c, im in enumerate(v.images):
            rhs += float(np.dot(grads[vi]["images"][c].cpu().numpy().astype(np.float64),
                                im.cpu().numpy().astype(np.float64)))
    scale = np.sqrt(n_out * n_g)
    print(f"3D vjp adjoint identity: {lhs:.6e} vs {rhs:.6e} (|diff| / (|g||Am|) = {abs(lhs - rhs) / scale:.2e})",
          flush=True)
    assert abs(lhs - rhs) <= 1e-5 * scale


def test_ge3d_vjp_after_forward_matches_the_plain_vjp(grad_setup):
    # the adjoint on the slice planes of the forward just run equals the one that rebuilds them
    import torch
    w, fw, vols, g, grads = grad_setup
    gd = [[torch.from_numpy(x).cuda() for x in gv] for gv in g]
    fw.forward()
    again = fw.vjp(gd, after_forward=True)
    torch.cuda.synchronize()
    for a, b in zip(again, grads):
        assert torch.equal(a["fmap"], b["fmap"])
        for key in ("images", "t_inhom"):
            for x, y in zip(a[key], b[key]):
                assert (x is None and y is None) or torch.equal(x, y)


def test_ge3d_vjp_again_with_the_same_cotangent_tensors_changed_in_place(torch_dev):
    # The adjoint partition transform keeps a plan over the first cotangent tensors it was given; a
    # second call with the same tensors, scaled by 2 in place, must read the new values (the
    # gradients double), and cotangents in new tensors (copied into the train's own buffers) give
    # the same gradients as the first call.
    torch = torch_dev
    w = workloads.make_ge3d("ge3d_small")
    fw, vols = workloads.to_ge3d(w)
    fw.forward()
    rng = np.random.default_rng(32)
    gd = [[torch.from_numpy((rng.standard_normal(o.shape) + 1j * rng.standard_normal(o.shape)).astype(np.complex64))
           .cuda() for o in ov] for ov in fw.out]
    first = fw.vjp(gd)
    for gv in gd:
        for t in gv:
            t.mul_(2.0)
    second = fw.vjp(gd)
    fresh = fw.vjp([[(t * 0.5).contiguous() for t in gv] for gv in gd])
    torch.cuda.synchronize()
    for a, b, c in zip(first, second, fresh):
        assert torch.allclose(b["fmap"], 2 * a["fmap"], rtol=1e-5, atol=1e-6 * float(a["fmap"].abs().max()))
        assert torch.equal(c["fmap"], a["fmap"])
        for key in ("images", "t_inhom"):
            for x, y, z in zip(a[key], b[key], c[key]):
                if x is None:
                    continue
                assert torch.allclose(y, 2 * x, rtol=1e-5, atol=1e-6 * float(x.abs().max()))
                assert torch.equal(z, x)
    fw.close()


def _ge3d_oracle_loss(w, v, gv, images=None, fmap=None, t_inhom=None):
    loss = 0.0
    for e in range(len(w.echo_times_ms)):
        k = G.ge3d_kspace(v.snx, v.sny, v.nx, v.ny, v.nz, v.images if images is None else images, v.t2,
                          v.t_inhom if t_inhom is None else t_inhom, v.fmap if fmap is None else fmap, None,
                          v.coil_maps, w.acq, w.echo_times_ms[e], w.ky_segments, w.t_line_ms,
                          v.weights.astype(np.float32))
        loss += float(np.real(np.vdot(gv[e], k)))
    return loss


def test_ge3d_vjp_matches_fp64_central_differences(grad_setup):
    # volume 1 (32^2 -> 16^2, 6 partitions, 2 compartments, 2 echoes): the fieldmap, T2' and image
    # gradients at bright voxels of two partitions against FP64 differences of the oracle's train
    w, fw, vols, g, grads = grad_setup
    vi = 1
    v = w.volumes[vi]
    P = v.snx * v.sny
    gf = grads[vi]["fmap"].cpu().numpy()
    gt = [t.cpu().numpy() for t in grads[vi]["t_inhom"]]
    gi = [t.cpu().numpy() for t in grads[vi]["images"]]
    mag = np.abs(v.images[0]) + np.abs(v.images[1])
    for z in (1, 4):
        p = int(np.argmax(mag[z]))
        i = z * P + p
        h = 1e-3 * max(abs(float(v.fmap[z, p])), 1.0)
        fp, fm = v.fmap.copy(), v.fmap.copy()
        fp[z, p] += h
        fm[z, p] -= h
        d = (_ge3d_oracle_loss(w, v, g[vi], fmap=fp) - _ge3d_oracle_loss(w, v, g[vi], fmap=fm)) / (fp[z, p] - fm[z, p])
        print(f"partition {z} voxel {p}: dL/dfmap gpu {gf[i]:.6e} fp64 {d:.6e}", flush=True)
        assert abs(gf[i] - d) <= 2e-3 * max(abs(d), 1e-3 * np.abs(gf).max())
        base = _ge3d_oracle_loss(w, v, g[vi])
        for c in range(len(v.images)):
            h = 1e-3 * float(v.t_inhom[c][z, p])
            tp = [t.copy() for t in v.t_inhom]
            tm = [t.copy() for t in v.t_inhom]
            tp[c][z, p] += h
            tm[c][z, p] -= h
            d = (_ge3d_oracle_loss(w, v, g[vi], t_inhom=tp) - _ge3d_oracle_loss(w, v, g[vi], t_inhom=tm)) / \
                (tp[c][z, p] - tm[c][z, p])
            print(f"  compartment {c}: dL/dT2' gpu {gt[c][i]:.6e} fp64 {d:.6e}", flush=True)
            assert abs(gt[c][i] - d) <= 2e-3 * max(abs(d), 1e-3 * np.abs(gt[c]).max())
            # image: linear, so a unit step gives the derivative up to FP64 rounding
            im = [x.copy() for x in v.images]
            im[c][z, p] += 1.0
            d = _ge3d_oracle_loss(w, v, g[vi], images=im) - base
            print(f"  compartment {c}: dL/dimage gpu {gi[c][i]:.6e} fp64 {d:.6e}", flush=True)
            assert abs(gi[c][i] - d) <= 2e-4 * max(abs(d), 1e-3 * np.abs(gi[c]).max())


@pytest.fixture(scope="module")
def full(torch_dev):
    torch = torch_dev
    w = workloads.make_ge3d("ge3d")
    fw, keep = workloads.to_ge3d(w)
    fw.tune(2)
    fw.forward()
    torch.cuda.synchronize()
    yield w, fw
    fw.close()


def test_ge3d_full_size_one_coil_matches_oracle(full):
    w, fw = full
    for vi, v in enumerate(w.volumes):
        coil = 5 + 11 * vi
        for e in (0, len(w.echo_times_ms) - 1):
            want = _oracle_volume(w, v, e, coils=[coil])[:, 0]
            got = fw.out[vi][e][:, coil].cpu().numpy()
            r = _rel(got, want)
            print(f"ge3d volume {vi} ({v.snx}->{v.nx}) echo {e} coil {coil}: rel-L2 {r:.2e}", flush=True)
            assert r <= 1e-5


def test_ge3d_full_size_partition_round_trip(full):
    import torch
    w, fw = full
    for vi, v in enumerate(w.volumes):
        C, nz, ny, nx, kseg = w.n_coils, v.nz, v.ny, v.nx, w.ky_segments
        wt = torch.from_numpy(v.weights[:, :, 0].astype(np.float32)).cuda()           # (nz, kseg)
        sy = torch.arange(ny, device="cuda") % kseg
        wl = wt[:, sy][:, None, :, None]                                                 # (nz, 1, ny, 1)
        for e in range(len(w.echo_times_ms)):
            k3 = fw.out[vi][e]
            back = torch.zeros_like(k3)
            gpu.zstage([gpu.ZSet((k3 / wl).contiguous(), back, nz, C, ny, nx)], adjoint=True)
            back /= nz
            k2 = fw.k2[vi][e][0]
            r = float(torch.linalg.vector_norm(back - k2) / torch.linalg.vector_norm(k2))
            print(f"ge3d volume {vi} echo {e}: partition round trip rel-L2 {r:.2e}", flush=True)
            assert r <= 1e-5


def test_ge3d_full_size_vjp_matches_fp64_central_differences(full):
    # Both full-size volumes (128^2 -> 64^2 and 192^2 -> 96^2, 24 partitions, 32 coils, 4 echoes): a
    # seeded cotangent on two coils of every echo, and the fieldmap, T2' and image gradients at the
    # brightest voxel of two partitions against FP64 central differences of the oracle's train.  The
    # train is a sum over voxels of image[v] times a function of voxel v's own fieldmap and T2', so
    # each derivative is that of the oracle run on an image holding voxel v alone (the oracle skips
    # the partitions whose image is zero, which keeps this to a few seconds of CPU).
    import torch
    w, fw = full
    rng = np.random.default_rng(41)
    coils = [3, 20]
    g, gd = [], []
    for ov in fw.out:
        gv, gdv = [], []
        for o in ov:
            x = np.zeros(o.shape, dtype=np.complex64)
            sub = (rng.standard_normal((o.shape[0], len(coils)) + tuple(o.shape[2:])) +
                   1j * rng.standard_normal((o.shape[0], len(coils)) + tuple(o.shape[2:]))).astype(np.complex64)
            x[:, coils] = sub
            gv.append(sub)
            gdv.append(torch.from_numpy(x).cuda())
        g.append(gv)
        gd.append(gdv)
    grads = fw.vjp(gd)
    torch.cuda.synchronize()

    def loss(v, gv, image, fmap=None, t_inhom=None):
        total = 0.0
        for e in range(len(w.echo_times_ms)):
            k = G.ge3d_kspace(v.snx, v.sny, v.nx, v.ny, v.nz, [image], v.t2, v.t_inhom if t_inhom is None else
                              [t_inhom], v.fmap if fmap is None else fmap, None, v.coil_maps[:, coils], w.acq,
                              w.echo_times_ms[e], w.ky_segments, w.t_line_ms, v.weights.astype(np.float32))
            total += float(np.real(np.vdot(gv[e], k)))
        return total

    for vi, v in enumerate(w.volumes):
        P = v.snx * v.sny
        gf = grads[vi]["fmap"].cpu().numpy()
        gt = grads[vi]["t_inhom"][0].cpu().numpy()
        gi = grads[vi]["images"][0].cpu().numpy()
        for z in (5, 17):
            p = int(np.argmax(np.abs(v.images[0][z])))
            i = z * P + p
            one = np.zeros_like(v.images[0])
            one[z, p] = v.images[0][z, p]
            h = 1e-3 * max(abs(float(v.fmap[z, p])), 1.0)
            fp, fm = v.fmap.copy(), v.fmap.copy()
            fp[z, p] += h
            fm[z, p] -= h
            d = (loss(v, g[vi], one, fmap=fp) - loss(v, g[vi], one, fmap=fm)) / (float(fp[z, p]) - float(fm[z, p]))
            print(f"volume {vi} ({v.snx}->{v.nx}) partition {z} voxel {p}: dL/dfmap gpu {gf[i]:.6e} fp64 {d:.6e}",
                  flush=True)
            assert abs(gf[i] - d) <= 2e-3 * max(abs(d), 1e-3 * np.abs(gf).max())
            ti0 = float(v.t_inhom[0])
            tp = np.full((v.nz, P), ti0, dtype=np.float32)
            tm = tp.copy()
            tp[z, p] += 1e-3 * ti0
            tm[z, p] -= 1e-3 * ti0
            d = (loss(v, g[vi], one, t_inhom=tp) - loss(v, g[vi], one, t_inhom=tm)) / (float(tp[z, p]) - float(tm[z, p]))
            print(f"  dL/dT2' gpu {gt[i]:.6e} fp64 {d:.6e}", flush=True)
            assert abs(gt[i] - d) <= 2e-3 * max(abs(d), 1e-3 * np.abs(gt).max())
            # the image: the loss is linear in it, so the voxel at 1 and every other at 0 is the derivative
            unit = np.zeros_like(v.images[0])
            unit[z, p] = 1.0
            d = loss(v, g[vi], unit)
            print(f"  dL/dimage gpu {gi[i]:.6e} fp64 {d:.6e}", flush=True)
            assert abs(gi[i] - d) <= 2e-4 * max(abs(d), 1e-3 * np.abs(gi).max())


def _oracle_pipeline(w, v, e, acq, seed, g, k3=None, kernels=None):
    """The oracle's acquired k-space -> partitions_to_slices -> reconstruct_slices: (nz, ny*nx)."""
    if k3 is None:
        k3 = _oracle_volume(w, v, e)
    mask = O.sampling_mask(1, v.ny, acq)
    sl = G.partitions_to_slices(k3.copy(), mask, g, seed ^ G.echo_salt(e), acq)
    return G.reconstruct_slices(sl, acq, kernels)


def test_ge3d_acquire_matches_oracle_pipeline(torch_dev):
    # spikes, receiver noise on the (volume, partition, echo) streams, inverse z-DFT, reconstruction:
    # the GPU against the oracle's restatement on the same seeds.  Two split-BF16 stages (forward and
    # reconstruction) compose, so the bound is 2e-5.
    w = workloads.make_ge3d("ge3d_small")
    acq = w.acq.copy(noise_variance=0.3, n_spikes=2, spike_amplitude=1.7)
    w.acq = acq
    fw, keep = workloads.to_ge3d(w)
    fw.forward()
    imgs = fw.acquire(seed=1234, volume_index=[3, 8])
    for vi, v in enumerate(w.volumes):
        for e in range(len(w.echo_times_ms)):
            want = _oracle_pipeline(w, v, e, acq, 1234, [3, 8][vi])
            got = imgs[vi][e].cpu().numpy().reshape(v.nz, -1)
            r = _rel(got, want)
            print(f"acquire volume {vi} echo {e}: rel-L2 vs oracle {r:.2e}", flush=True)
            assert r <= 2e-5
    fw.close()


def test_ge3d_acquire_with_grappa(torch_dev):
    # Accelerated acquisition: kernels per (echo, slice) fitted on a reference scan (no acceleration,
    # no spikes, volume index 0, seed ^ echo_salt(e) ^ the calibration salt).
    # 1. the oracle's kernels fed to the GPU pipeline vs the oracle's pipeline with the same kernels
    #    (one linear map on both sides): rel-L2 <= 2e-5;
    # 2. the GPU's own calibration and pipeline vs the oracle's, bounded by 3x the image error the
    #    oracle's pipeline shows when its k-spaces carry random relative errors of the GPU's measured
    #    k-space size (the regularised fit amplifies them).
    torch = torch_dev
    w = workloads.make_ge3d("ge3d_small")
    w.volumes = w.volumes[:1]
    acq = w.acq.copy(noise_variance=0.2, accel=2, acs_lines=6)
    w.acq = acq
    fw, keep = workloads.to_ge3d(w)
    fw.forward()
    v = w.volumes[0]
    ne = len(w.echo_times_ms)
    acq_cal = acq.copy(accel=1, n_spikes=0)
    lines = O.sampling_mask(1, v.ny, acq_cal)
    shape = (acq.accel - 1, w.n_coils, 6 * w.n_coils)
    k_cal, ker = [], []
    for e in range(ne):
        k_cal.append(G.ge3d_kspace(v.snx, v.sny, v.nx, v.ny, v.nz, v.images, v.t2, v.t_inhom, v.fmap, None,
                                   v.coil_maps, acq_cal, w.echo_times_ms[e], w.ky_segments, w.t_line_ms,
                                   v.weights.astype(np.float32)))
        sl = G.partitions_to_slices(k_cal[e].copy(), lines, 0, 99 ^ G.echo_salt(e) ^ G.SEED_SALT_GRAPPA_CAL,
                                    acq_cal)
        ker.append([O.grappa_fit(sl[z].reshape(w.n_coils, -1), v.nx, v.ny, acq.accel, acq.acs_lines, lines)
                    for z in range(v.nz)])
    given = [[[torch.from_numpy(k.reshape(shape)).cuda() for k in ker[e]] for e in range(ne)]]
    imgs = fw.acquire(seed=99, grappa=given)
    wants = []
    for e in range(ne):
        wants.append(_oracle_pipeline(w, v, e, acq, 99, 0, kernels=ker[e]))
        r = _rel(imgs[0][e].cpu().numpy().reshape(v.nz, -1), wants[e])
        print(f"grappa echo {e}: GPU pipeline with the oracle's kernels rel-L2 {r:.2e}", flush=True)
        assert r <= 2e-5
    # the whole accelerated pipeline on the GPU, its calibration included
    cal = fw.grappa_calibration(seed=99)
    imgs = fw.acquire(seed=99, grappa=cal)
    rng = np.random.default_rng(1)
    for e in range(ne):
        k3 = _oracle_volume(w, v, e)
        ek = _rel(fw.out[0][e].cpu().numpy(), k3)
        pert = lambda k: k + ek * np.sqrt(np.mean(np.abs(k[k != 0]) ** 2)) * (k != 0) * (
            rng.standard_normal(k.shape) + 1j * rng.standard_normal(k.shape)) / np.sqrt(2)
        sl = G.partitions_to_slices(pert(k_cal[e]), lines, 0, 99 ^ G.echo_salt(e) ^ G.SEED_SALT_GRAPPA_CAL, acq_cal)
        kp = [O.grappa_fit(sl[z].reshape(w.n_coils, -1), v.nx, v.ny, acq.accel, acq.acs_lines, lines)
              for z in range(v.nz)]
        e_pert = _rel(_oracle_pipeline(w, v, e, acq, 99, 0, k3=pert(k3), kernels=kp), wants[e])
        r = _rel(imgs[0][e].cpu().numpy().reshape(v.nz, -1), wants[e])
        print(f"grappa echo {e}: whole GPU pipeline rel-L2 {r:.2e}; oracle with k-space errors of "
              f"{ek:.1e}: {e_pert:.2e}", flush=True)
        assert r <= max(2e-5, 3.0 * e_pert)
    fw.close()
