"""The NUFFT comparator on the GPU (nufft.hip, FP64) against the oracle's restatement of the
`kspace` feature build (orc_nufft1 / orc_build_coil_kspace_nufft, pinned bit for bit to the
reference's kspace golden record) and against the exact forward.

- mrs_nufft1 on the reference's `matches_the_direct_sum` configurations (src/nufft.rs:325-347):
  within 1e-12 of the oracle's NUFFT (same algorithm, FP64 both; the GPU's fine-grid sums run in
  atomic order and its FFT is a radix-2 transform, the oracle evaluates the bins directly), and
  within the reference's bound 10 tol x total of the direct sum; power-of-two and other fine grids;
  integer positions reduce to the DFT;
- mrs_build_coil_kspaces_nufft: every golden configuration the NUFFT path takes, every coil,
  against the oracle's feature build (rel-L2 <= 1e-12, FP64 both), with spikes and noise as well;
  the GPU k-spaces reconstructed by the oracle reproduce the kspace golden images (rel-L2 <= 1e-6 of
  f32 images); against the exact GPU forward (split-BF16, rel-L2 <= 1e-5) and the oracle's rotor
  path (1e-12: the two y-stages are the same sum).
"""
import ctypes as ct
import os

import numpy as np
import pytest

from mrsim_amd import abi, gpu
from oracle import oracle as O

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "p4_baseline")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _cases():
    st = ct.c_uint64(42)
    L = O.lib()
    unit = lambda: L.orc_rng_unit(ct.byref(st))
    out = []
    for n, k_lo, n_out, tol in [(48, -12, 24, 1e-12), (140, -35, 70, 1e-12), (16, -8, 16, 1e-12), (64, -16, 32, 1e-8)]:
        nsrc = 3 * n
        pos = np.array([(unit() - 0.5) * 3.0 * n for _ in range(nsrc)])
        re = np.array([unit() - 0.5 for _ in range(nsrc)])
        im = np.array([unit() - 0.5 for _ in range(nsrc)])
        re2 = np.array([unit() for _ in range(nsrc)])
        out.append((n, k_lo, n_out, tol, pos, [re + 1j * im, re2 + 0j]))
    return out


def test_nufft1_matches_oracle_and_direct_sum():
    import torch
    for n, k_lo, n_out, tol, pos, ws in _cases():
        want = O.nufft1(n, k_lo, n_out, tol, pos, ws)
        # three columns sharing the positions, the weight vectors permuted per column
        W = np.stack([np.stack(ws), np.stack(ws[::-1]), 2.0 * np.stack(ws)])
        got = gpu.nufft1(torch.from_numpy(pos[None, :]).cuda(), torch.from_numpy(W).cuda(), n, k_lo, n_out,
                         tol).cpu().numpy()
        for col, order, scale in [(0, (0, 1), 1.0), (1, (1, 0), 1.0), (2, (0, 1), 2.0)]:
            for q, wq in enumerate(order):
                e = rel(got[col, q], scale * want[wq])
                d = float(np.max(np.abs(got[col, q] - scale * O.nufft_direct(n, k_lo, n_out, pos, ws[wq]))))
                bound = 10.0 * tol * float(np.sum(np.abs(scale * ws[wq])))
                print(f"n={n} nf={2 * n} col {col} q {q}: vs oracle {e:.1e}, vs direct {d:.1e} (bound {bound:.1e})",
                      flush=True)
                assert e <= 1e-12
                assert d <= bound


def test_nufft1_integer_positions_reduce_to_the_dft():
    import torch
    n = 32
    pos = np.arange(n, dtype=np.float64)
    w = np.sin(0.3 * pos) + 0j
    got = gpu.nufft1(torch.from_numpy(pos[None, :]).cuda(), torch.from_numpy(w[None, None, :]).cuda(), n, -16, 32,
                     1e-12).cpu().numpy()[0, 0]
    want = np.array([np.sum(w * np.exp(2j * np.pi * k * pos / n)) for k in range(-16, 16)])
    assert np.max(np.abs(got - want)) < 1e-10


NUFFT_CASES = [(s, c[0]) for s in range(3) for c in O.golden_cases() if not c[3] and c[1].eddy_strength == 0.0]


@pytest.mark.parametrize("shape,name", NUFFT_CASES, ids=[f"{O.GOLDEN_SHAPES[s][0]}x{O.GOLDEN_SHAPES[s][1]}-{n}"
                                                         for s, n in NUFFT_CASES])
def test_nufft_forward_matches_the_oracle_feature_build(shape, name):
    nx, ny, o = O.GOLDEN_SHAPES[shape]
    _, acq, ed, maps = {c[0]: c for c in O.golden_cases()}[name]
    b = O.golden_slice_builder(shape, ed, maps)
    nc = max(int(acq.n_coils), 1)
    got = gpu.build_coil_kspaces_nufft(b, acq)
    want = []
    for c in range(nc):
        k, used = O.coil_kspace_nufft(b, acq, c, nc)
        assert used
        want.append(k)
    want = np.stack(want)
    e = rel(got, want)
    # the exact paths: the oracle's rotor y-stage (same sum) and the GPU's split-BF16 forward
    e_rot = rel(got, O.coil_kspaces(b, acq))
    e_gpu = rel(got, gpu.build_coil_kspaces(b, acq))
    print(f"{name} {nx}x{ny}: vs oracle NUFFT build {e:.1e}, vs rotor path {e_rot:.1e}, vs exact GPU {e_gpu:.1e}",
          flush=True)
    assert e <= 1e-12
    assert e_rot <= 1e-12
    assert e_gpu <= 1e-5
    if acq.accel <= 1:
        # the GPU's k-spaces through the oracle's reconstruction reproduce the kspace-build image
        img = O.reconstruct(got.copy(), nx, ny, acq, None)
        gold = np.fromfile(os.path.join(GOLDEN, "kspace", f"slice-{nx}x{ny}o{o}-{name}.bin"), dtype=np.float32)
        gold = gold[0::2] + 1j * gold[1::2]
        assert rel(img, gold) <= 1e-6


def test_nufft_forward_with_spikes_and_noise():
    nx, ny, o = O.GOLDEN_SHAPES[0]
    _, acq, ed, maps = {c[0]: c for c in O.golden_cases()}["distortion+relaxation"]
    acq = acq.copy(n_coils=3, noise_variance=4.0, n_spikes=3, spike_amplitude=2.5)
    b = O.golden_slice_builder(0, ed, maps, slice_seed=1234)
    got = gpu.build_coil_kspaces_nufft(b, acq)
    want = np.stack([O.coil_kspace_nufft(b, acq, c, 3)[0] for c in range(3)])
    e = rel(got, want)
    print(f"spikes and noise: vs oracle NUFFT build {e:.1e}", flush=True)
    assert e <= 1e-12
