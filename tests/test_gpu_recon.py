"""GPU reconstruction (mrs_recon_*, mrs_reconstruct_coils) against the oracle's restatement of
reconstruct_coils_full (src/kspace.rs:1301-1378) and the reference's golden images.

- the oracle's own FP64 coil k-spaces reconstructed on the GPU vs by the oracle: every
  non-accelerated golden configuration and shape, rel-L2 <= 1e-5 (the GPU computes in FP32 with
  split-BF16 products);
- simulate_slice end to end on the GPU (k-space and reconstruction) vs the golden images;
- windows (Hann, Tukey, Fermi) and odd matrices vs the oracle;
- a batch: the 600-unit forward's k-spaces reconstructed in one launch, sampled units vs the oracle.
"""
import os

import numpy as np
import pytest

from mrsim_amd import abi, gpu, workloads
from oracle import oracle as O

pytestmark = pytest.mark.gpu
TOL = 1e-5


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


CASES = [(s, c[0]) for s in range(3) for c in O.golden_cases() if c[1].accel <= 1]


@pytest.mark.parametrize("shape,name", CASES, ids=[f"{O.GOLDEN_SHAPES[s][0]}x{O.GOLDEN_SHAPES[s][1]}-{n}"
                                                     for s, n in CASES])
def test_gpu_reconstruction_matches_oracle_and_golden(shape, name):
    nx, ny, o = O.GOLDEN_SHAPES[shape]
    _, acq, ed, maps = {c[0]: c for c in O.golden_cases()}[name]
    b = O.golden_slice_builder(shape, ed, maps)
    coils = O.coil_kspaces(b, acq)
    want = O.reconstruct(coils.copy(), nx, ny, acq, None)
    got = gpu.reconstruct_coils(coils, nx, ny, acq)
    e_rec = rel(got, want)
    # the whole of simulate_slice on the GPU: k-space, then reconstruction
    g = gpu.build_coil_kspaces(b, acq)
    full = gpu.reconstruct_coils(g, nx, ny, acq)
    gold = np.fromfile(os.path.join(os.path.dirname(__file__), "golden", "p4_baseline", "default",
                                    f"slice-{nx}x{ny}o{o}-{name}.bin"), dtype=np.float32)
    gold = gold[0::2] + 1j * gold[1::2]
    e_gold = rel(full, gold)
    print(f"{name} {nx}x{ny}: recon vs oracle {e_rec:.2e}, GPU simulate_slice vs golden {e_gold:.2e}", flush=True)
    assert e_rec <= TOL
    assert e_gold <= TOL


@pytest.mark.parametrize("window,a,b", [(abi.WINDOW_HANN, 0.0, 0.0), (abi.WINDOW_TUKEY, 0.4, 0.0),
                                        (abi.WINDOW_FERMI, 0.8, 0.05)])
@pytest.mark.parametrize("nx,ny,nc", [(17, 70, 4), (64, 64, 8), (128, 97, 2), (9, 8, 1)])
def test_windows_and_shapes_match_oracle(window, a, b, nx, ny, nc):
    rng = np.random.default_rng(nx * 1000 + ny + window)
    k = (rng.standard_normal((nc, nx * ny)) + 1j * rng.standard_normal((nc, nx * ny))) * 10.0
    acq = abi.Acquisition.default(n_coils=nc, window=window, window_a=a, window_b=b)
    want = O.reconstruct(k.copy(), nx, ny, acq, None)
    got = gpu.reconstruct_coils(k, nx, ny, acq)
    e = rel(got, want)
    print(f"window {window} {nx}x{ny} C={nc}: rel-L2 {e:.2e}", flush=True)
    assert e <= TOL


def test_batched_reconstruction_of_the_forward():
    import torch
    w = workloads.make("batch_128", n_slices=3, n_volumes=4)
    plan, outs, keep = workloads.to_plan(w)
    plan.forward()
    rec = gpu.Recon(w.nx, w.ny, w.n_coils, w.acq)
    imgs = [torch.empty((w.ny, w.nx), dtype=torch.complex64, device="cuda") for _ in outs]
    rec.run(outs, imgs)
    for u in (0, 5, 11):
        k = outs[u].cpu().numpy().reshape(w.n_coils, -1).astype(np.complex128)
        want = O.reconstruct(k.copy(), w.nx, w.ny, w.acq, None)
        e = rel(imgs[u].cpu().numpy().reshape(-1), want)
        print(f"unit {u}: batched recon vs oracle recon of the same k-space {e:.2e}", flush=True)
        assert e <= TOL
    rec.close()
    plan.close()


GRAPPA_CASES = [(s, c[0]) for s in range(3) for c in O.golden_cases() if c[1].accel > 1]


def _ref_scan(shape, name):
    nx, ny, o = O.GOLDEN_SHAPES[shape]
    _, acq, ed, maps = {c[0]: c for c in O.golden_cases()}[name]
    ref_b = O.golden_slice_builder(shape, None, maps, slice_seed=7 ^ 0x47525043414C)
    acq_ref = acq.copy(accel=1, n_spikes=0)
    return nx, ny, acq, acq_ref, O.coil_kspaces(ref_b, acq_ref)


@pytest.mark.parametrize("shape,name", GRAPPA_CASES, ids=[f"{O.GOLDEN_SHAPES[s][0]}x{O.GOLDEN_SHAPES[s][1]}-{n}"
                                                            for s, n in GRAPPA_CASES])
def test_grappa_fit_and_apply_match_oracle(shape, name):
    # The same FP32-rounded reference scan into both fits: FP64 arithmetic in the reference's operation
    # order on both sides, so the weights agree to FP64 round-off.
    import torch
    nx, ny, acq, acq_ref, ref = _ref_scan(shape, name)
    ref32 = ref.astype(np.complex64)
    lines = O.sampling_mask(1, ny, acq_ref)
    want = O.grappa_fit(ref32.astype(np.complex128), nx, ny, acq.accel, acq.acs_lines, lines)
    w = gpu.grappa_fit(torch.from_numpy(ref32.reshape(-1, ny, nx)).cuda(), acq.accel, acq.acs_lines, lines)
    got = w.cpu().numpy().reshape(-1)
    e = rel(got, want)
    print(f"{name}: GRAPPA weights rel diff {e:.2e}", flush=True)
    assert e <= 1e-12
    # apply + reconstruction with those weights vs the oracle's
    b = O.golden_slice_builder(shape, *{c[0]: c for c in O.golden_cases()}[name][2:4])
    k = O.coil_kspaces(b, acq).astype(np.complex64).astype(np.complex128)
    img_want = O.reconstruct(k.copy(), nx, ny, acq, want)
    img_got = gpu.reconstruct_coils(k, nx, ny, acq, grappa_weights=want)
    e2 = rel(img_got, img_want)
    print(f"{name}: GRAPPA reconstruction vs oracle {e2:.2e}", flush=True)
    assert e2 <= TOL


ALL_CASES = [(s, c[0]) for s in range(3) for c in O.golden_cases()]


@pytest.mark.parametrize("shape,name", ALL_CASES, ids=[f"{O.GOLDEN_SHAPES[s][0]}x{O.GOLDEN_SHAPES[s][1]}-{n}"
                                                         for s, n in ALL_CASES])
def test_gpu_simulate_slice_matches_golden(shape, name):
    # simulate_slice (src/kspace.rs:1229-1284) entirely on the GPU against the reference's records.
    # GRAPPA cases: bounded by the image error the oracle's own pipeline gives when its k-spaces carry
    # random relative errors of the GPU's measured k-space size (the regularised fit amplifies them).
    nx, ny, o = O.GOLDEN_SHAPES[shape]
    _, acq, ed, maps = {c[0]: c for c in O.golden_cases()}[name]
    b = O.golden_slice_builder(shape, ed, maps)
    gold = np.fromfile(os.path.join(os.path.dirname(__file__), "golden", "p4_baseline", "default",
                                    f"slice-{nx}x{ny}o{o}-{name}.bin"), dtype=np.float32)
    gold = gold[0::2] + 1j * gold[1::2]
    got = gpu.simulate_slice(b, acq)
    e = rel(got, gold)
    if acq.accel <= 1:
        print(f"{name} {nx}x{ny}: GPU simulate_slice vs golden {e:.2e}", flush=True)
        assert e <= TOL
        return
    _, _, _, acq_ref, o_ref = _ref_scan(shape, name)
    ref_b = O.golden_slice_builder(shape, None, maps, slice_seed=7 ^ 0x47525043414C)
    g_ref, g = gpu.build_coil_kspaces(ref_b, acq_ref), gpu.build_coil_kspaces(b, acq)
    o_k = O.coil_kspaces(b, acq)
    ek = max(rel(g_ref, o_ref), rel(g, o_k))
    rng = np.random.default_rng(0)
    pert = lambda k: k + ek * np.sqrt(np.mean(np.abs(k[k != 0]) ** 2)) * (k != 0) * (
        rng.standard_normal(k.shape) + 1j * rng.standard_normal(k.shape)) / np.sqrt(2)
    lines = O.sampling_mask(1, ny, acq_ref)
    w = O.grappa_fit(pert(o_ref), nx, ny, acq.accel, acq.acs_lines, lines)
    e_pert = rel(O.reconstruct(pert(o_k), nx, ny, acq, w), gold)
    print(f"{name} {nx}x{ny}: GPU simulate_slice vs golden {e:.2e}, oracle with perturbed k-spaces {e_pert:.2e}",
          flush=True)
    assert e <= max(TOL, 3.0 * e_pert)
