"""mrs_grappa_fit_noalloc: the GRAPPA fit with its scratch in the library's fixed per-GPU device
array (no allocation), against mrs_grappa_fit on the same inputs.

Two calls in a row, the second with a smaller problem (fewer coils, another acceleration and
matrix), each must give the weights mrs_grappa_fit gives, bit for bit (the same kernels read the same
values; nothing left in the shared array from the first call may leak into the second).  A problem
beyond the array's limit is refused with a message before any GPU work.
"""
import numpy as np
import pytest

from mrsim_amd import abi, gpu
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    torch.cuda.set_device(0)
    gpu.lib()


def _ref_scan(shape, name, n_coils=None):
    nx, ny, o = O.GOLDEN_SHAPES[shape]
    _, acq, ed, maps = {c[0]: c for c in O.golden_cases()}[name]
    if n_coils is not None:
        acq = acq.copy(n_coils=n_coils)
    acq_ref = acq.copy(accel=1, n_spikes=0)
    ref_b = O.golden_slice_builder(shape, None, maps, slice_seed=7 ^ 0x47525043414C)
    ref = O.coil_kspaces(ref_b, acq_ref).astype(np.complex64)
    return nx, ny, acq, O.sampling_mask(1, ny, acq_ref), ref


def test_noalloc_fit_matches_the_allocating_fit_on_two_calls():
    import torch
    # 4 coils R3 at 17x70, then a smaller problem: 2 coils R2 at 16x40
    cases = [(0, "grappa3", None), (1, "grappa-coils", 2)]
    for shape, name, nc in cases:
        nx, ny, acq, lines, ref = _ref_scan(shape, name, nc)
        C = ref.shape[0]
        k = torch.from_numpy(ref.reshape(C, ny, nx)).cuda()
        want = gpu.grappa_fit(k, acq.accel, acq.acs_lines, lines)
        got = torch.full_like(want, complex(np.nan, np.nan))
        gpu.grappa_fit_noalloc(k, acq.accel, acq.acs_lines, lines, got)
        torch.cuda.synchronize()
        print(f"{name} {nx}x{ny}: {C} coils, R{acq.accel}", flush=True)
        assert torch.equal(got, want), f"{name}: no-alloc weights differ"


def test_noalloc_fit_refuses_a_problem_beyond_its_scratch():
    import torch
    C, ny, nx, accel = 64, 32, 8, 5   # 4 offsets of 64 coils: 16 * 4 * 42 * 64^2 bytes = 11 MB > 8 MiB
    k = torch.zeros((C, ny, nx), dtype=torch.complex64, device="cuda")
    w = torch.zeros((accel - 1, C, 6 * C), dtype=torch.complex128, device="cuda")
    with pytest.raises(gpu.MrsError) as ei:
        gpu.grappa_fit_noalloc(k, accel, 16, np.ones(ny, np.uint8), w)
    assert "no-alloc limit" in str(ei.value)
