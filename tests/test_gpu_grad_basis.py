"""GPU slice-basis y-adjoint: the volumes of a slice take their y-adjoint in one kernel that generates
the slice's per-line phase and decay factors once and contracts every unit's x-adjoint result against
them on MFMA (mrs_plan_set_vjp_basis).  The adjoint of the forward in src/kspace.rs:607-1042 (the
reference has none; the per-unit adjoint kernels are the comparison).

Checks, on workloads with 20, 7 and 45 volumes per slice (128^2 -> 64^2, 8 coils, 3 compartments) and
small grids (40^2 -> 20^2, 3 coils, Q = 1, 2 complex, 4; gradient echoes with the default timing and
with a caller's timing table), covering the three ways the adjoint forms its T2' and fieldmap sums:
  - every unit's image, fieldmap and T2' gradients against the per-unit y-adjoint (both split-BF16
    products of the same FP32 sums);
  - every output element written (the buffers start as NaN);
  - FP64 pins: the adjoint identity <g, A m> = <A^H g, m> against the oracle's FP64 forward, and on
    the small grids the fieldmap and T2' gradients against FP64 central differences of the oracle
    and the image gradient against the oracle's FP64 forward of single-voxel images.
"""
import numpy as np
import pytest

from mrsim_amd import workloads
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300))


def _workload(name):
    from mrsim_amd.workloads import Stream, Workload, _slice, _volume_images
    from mrsim_amd import abi
    if name == "batch_q3_v20":
        return workloads.make("batch_128", n_slices=2, n_volumes=20)
    if name == "batch_q3_v7":        # an odd set: the last pair has one unit
        return workloads.make("batch_128", n_slices=2, n_volumes=7)
    if name == "batch_q3_v45":       # three sets of 15 units
        return workloads.make("batch_128", n_slices=1, n_volumes=45)
    Q = {"small_q1": 1, "small_q2_cplx": 2, "small_q4": 4, "small_ge": 2, "small_ge_timing": 1}[name]
    cplx = name.endswith("cplx")
    C = 3
    acq = abi.Acquisition.default(t_line=0.5, t_echo=30.0, signal_scale=1.0, n_coils=C)
    if name.startswith("small_ge"):
        acq.echo = abi.ECHO_GRADIENT
    w = Workload(name, 40, 40, 20, 20, C, Q, acq, complex_images=cplx)
    if name == "small_ge_timing":
        # a caller's timing table whose RF times are not TE + t: the T2' weights are no multiple of
        # the fieldmap ones, so the adjoint keeps separate products
        from mrsim_amd import gpu
        t, trf, tread, pol = gpu.line_timing(acq, 20, 20)
        trf = trf + 0.3 * np.sin(np.arange(20))
        w.timing = (t, trf, tread, pol)
    for s in range(2):
        st = Stream(0x5EED0B00 + 31 * s + Q)
        sd = _slice(st, 40, 40, Q, C, 16.0, 24.0)
        w.slices.append(sd)
        for g in range(5 + s):
            img = _volume_images(st, sd, g)
            if cplx:
                img = (img * np.exp(1j * sd.phase0)[None, :]).astype(np.complex64)
            w.units.append((s, img))
    return w


def _cotangents(w, seed=11):
    rng = np.random.default_rng(seed)
    shp = (w.n_coils, w.ny, w.nx)
    return [(rng.standard_normal(shp) + 1j * rng.standard_normal(shp)).astype(np.complex64) for _ in w.units]


def _gradients(w, gouts, basis):
    import torch
    plan, outs, keep = workloads.to_plan(w)
    plan.set_vjp_basis(basis)
    assert plan.vjp_basis() == basis
    P, Q = w.snx * w.sny, w.n_comp
    dt = torch.complex64 if w.complex_images else torch.float32
    g = [torch.from_numpy(x).cuda() for x in gouts]
    gi = [torch.full((Q, P), float("nan"), dtype=dt, device="cuda") for _ in w.units]
    gf = [torch.full((P,), float("nan"), dtype=torch.float32, device="cuda") for _ in w.units]
    gt = [torch.full((Q, P), float("nan"), dtype=torch.float32, device="cuda") for _ in w.units]
    plan.forward()
    plan.vjp(g, gi, gf, gt, after_forward=True)
    torch.cuda.synchronize()
    res = [(a.cpu().numpy(), b.cpu().numpy(), c.cpu().numpy()) for a, b, c in zip(gi, gf, gt)]
    k = [o.cpu().numpy() for o in outs]
    plan.close()
    return res, k


def _oracle_k(w, u, images=None, fmap=None, t_inhom=None):
    si, img = w.units[u]
    sd = w.slices[si]
    return O.forward_direct(w.snx, w.sny, w.nx, w.ny, w.n_coils, w.n_comp, w.acq,
                            img if images is None else images, sd.fmap if fmap is None else fmap, sd.coil_maps,
                            phase0=None if w.complex_images else sd.phase0, t2=sd.t2,
                            t_inhom=sd.t_inhom if t_inhom is None else t_inhom, complex_images=w.complex_images,
                            timing=w.timing)


# The line-weight mode the plan picks for each workload: spin echoes split the fieldmap product by
# the sign of the line time (2), gradient echoes scale it (1), a caller's timing table with other RF
# times keeps separate products (0).
MODES = {"batch_q3_v20": 2, "batch_q3_v7": 2, "batch_q3_v45": 2, "small_q1": 2, "small_q2_cplx": 2, "small_q4": 2,
         "small_ge": 1, "small_ge_timing": 0}
CASES = list(MODES)


@pytest.mark.parametrize("name", CASES)
def test_adjoint_weight_mode(name):
    w = _workload(name)
    plan, outs, keep = workloads.to_plan(w)
    desc = plan.describe()
    plan.close()
    assert f"adjoint weights: mode {MODES[name]} " in desc, desc


@pytest.mark.parametrize("name", CASES)
def test_slice_basis_adjoint_matches_per_unit_and_fp64(name):
    w = _workload(name)
    gouts = _cotangents(w)
    got, k = _gradients(w, gouts, True)
    want, _ = _gradients(w, gouts, False)
    worst = [0.0, 0.0, 0.0]
    for u, (a, b) in enumerate(zip(got, want)):
        for j in range(3):
            assert np.all(np.isfinite(a[j])), f"unit {u}: gradient {j} has unwritten (NaN) elements"
            worst[j] = max(worst[j], rel(a[j], b[j]))
    print(f"{name}: max rel-L2 basis vs per-unit adjoint: image {worst[0]:.2e} fieldmap {worst[1]:.2e} "
          f"T2' {worst[2]:.2e}", flush=True)
    assert max(worst) <= 5e-5
    # FP64 adjoint identity of the image gradient on the first and the last unit:
    # Re <g, A m> = Re <A^H g, m> (real images: sum of gimg * m)
    for u in sorted({0, len(w.units) - 1}):
        km = _oracle_k(w, u)
        g = gouts[u].astype(np.complex128)
        lhs = np.real(np.vdot(g, km))
        m = w.units[u][1].reshape(w.n_comp, -1)
        gi = got[u][0]
        rhs = np.real(np.vdot(gi, m)) if w.complex_images else float(np.sum(gi.astype(np.float64) * m))
        scale = np.linalg.norm(g) * np.linalg.norm(km)
        print(f"  unit {u}: adjoint identity {lhs:.6e} vs {rhs:.6e} ({abs(lhs - rhs) / scale:.1e} of |g||Am|)",
              flush=True)
        assert abs(lhs - rhs) <= 1e-5 * scale


@pytest.mark.parametrize("name", ["small_q1", "small_q2_cplx", "small_q4", "small_ge", "small_ge_timing"])
def test_slice_basis_adjoint_matches_fp64_central_differences(name):
    # L_u = Re <g_u, A_u(m, f, T2')>.  The fieldmap and T2' derivatives at a few voxels of a unit against
    # FP64 central differences of the oracle (step 1e-3 relative), and the image gradient against the
    # oracle's forward of a single-voxel image (linear: exact up to FP64 rounding).
    w = _workload(name)
    gouts = _cotangents(w, seed=5)
    got, _ = _gradients(w, gouts, True)
    P, Q = w.snx * w.sny, w.n_comp
    for u in (1, len(w.units) - 1):
        si, img = w.units[u]
        sd = w.slices[si]
        g = gouts[u].astype(np.complex128)
        gi, gf, gt = got[u]
        loss = lambda **kw: float(np.real(np.vdot(g, _oracle_k(w, u, **kw))))
        mag = np.abs(img).reshape(Q, P).sum(axis=0)
        for v in [int(i) for i in np.argsort(-mag)[[0, 25, 300]]]:
            h = 1e-3 * max(abs(float(sd.fmap[v])), 1.0)
            fp, fm = sd.fmap.copy(), sd.fmap.copy()
            fp[v] += h
            fm[v] -= h
            d = (loss(fmap=fp) - loss(fmap=fm)) / (float(fp[v]) - float(fm[v]))
            assert abs(gf[v] - d) <= 2e-3 * max(abs(d), 1e-3 * np.abs(gf).max()), (u, v, gf[v], d)
            for q in range(Q):
                ti = np.ascontiguousarray(sd.t_inhom).reshape(Q, P)
                h = 1e-3 * float(ti[q, v])
                tp, tm = ti.copy(), ti.copy()
                tp[q, v] += h
                tm[q, v] -= h
                d = (loss(t_inhom=tp) - loss(t_inhom=tm)) / (float(tp[q, v]) - float(tm[q, v]))
                assert abs(gt[q, v] - d) <= 2e-3 * max(abs(d), 1e-3 * np.abs(gt[q]).max()), (u, v, q, gt[q, v], d)
                e = np.zeros((Q, P), dtype=img.dtype)
                e[q, v] = 1.0
                want = loss(images=e)
                if w.complex_images:
                    want = want + 1j * loss(images=(1j * e).astype(img.dtype))
                assert abs(gi[q, v] - want) <= 2e-4 * max(abs(want), 1e-3 * np.abs(gi[q]).max()), (u, v, q, gi[q, v], want)
        print(f"{name} unit {u}: fieldmap, T2' and image gradients match FP64 at 3 voxels", flush=True)


def test_slice_basis_adjoint_switch_round_trip():
    # Switching the slice-basis adjoint off and on again on one plan rebuilds its launch layout; the
    # results equal those of plans that never switched, bit for bit.
    import torch
    w = _workload("batch_q3_v7")
    gouts = _cotangents(w, seed=3)
    ref_on, _ = _gradients(w, gouts, True)
    ref_off, _ = _gradients(w, gouts, False)
    plan, outs, keep = workloads.to_plan(w)
    P, Q = w.snx * w.sny, w.n_comp
    g = [torch.from_numpy(x).cuda() for x in gouts]
    gi = [torch.empty((Q, P), dtype=torch.float32, device="cuda") for _ in w.units]
    gf = [torch.empty((P,), dtype=torch.float32, device="cuda") for _ in w.units]
    gt = [torch.empty((Q, P), dtype=torch.float32, device="cuda") for _ in w.units]
    plan.forward()
    for on, ref in ((True, ref_on), (False, ref_off), (True, ref_on)):
        plan.set_vjp_basis(on)
        plan.vjp(g, gi, gf, gt, after_forward=True)
        torch.cuda.synchronize()
        for u in range(len(w.units)):
            for t, r in zip((gi[u], gf[u], gt[u]), ref[u]):
                assert np.array_equal(t.cpu().numpy(), r), (on, u)
    plan.close()
