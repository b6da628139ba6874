"""GPU parity of the forward model against the FP64 oracle and the reference's golden fixtures.

Tolerances (relative L2 over each coil's k-space, or the reconstructed image):
  split-BF16 MFMA path  <= 1e-5   (three-term split products, FP32 accumulation, FP64 phases)
  plain FP32 path       <= 1e-5
Unacquired rows must be exactly zero.
"""
import sys

import numpy as np
import pytest

from mrsim_amd import abi, gpu, workloads
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-5


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300))


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    gpu.lib()


def test_phase_evaluation_accuracy():
    rng = np.random.default_rng(1)
    cyc = np.concatenate([rng.uniform(-300, 300, 200000), np.linspace(-2, 2, 4097),
                          np.array([0.0, 0.25, 0.5, -0.25, 0.125, 1e-9])])
    r = cyc - np.rint(cyc)
    for mode, name in [(1, "poly"), (2, "hardware")]:
        s, c = gpu.probe_sincos(cyc, mode)
        es = np.abs(s - np.sin(2 * np.pi * r)).max()
        ec = np.abs(c - np.cos(2 * np.pi * r)).max()
        rs = np.sqrt(np.mean((s - np.sin(2 * np.pi * r)) ** 2))
        print(f"sincos {name}: max abs error sin {es:.2e} cos {ec:.2e}, rms {rs:.2e}", flush=True)
    s, c = gpu.probe_sincos(cyc, 0)
    es = np.abs(s - np.sin(2 * np.pi * r)).max()
    ec = np.abs(c - np.cos(2 * np.pi * r)).max()
    assert es < 1e-6 and ec < 1e-6


CASES = [(s, c[0]) for s in range(3) for c in O.golden_cases()]


@pytest.mark.parametrize("shape,name", CASES, ids=[f"{O.GOLDEN_SHAPES[s][0]}x{O.GOLDEN_SHAPES[s][1]}-{n}"
                                                     for s, n in CASES])
def test_drop_in_coil_kspaces_match_oracle(shape, name):
    _, acq, ed, maps = {c[0]: c for c in O.golden_cases()}[name]
    b = O.golden_slice_builder(shape, ed, maps)
    got = gpu.build_coil_kspaces(b, acq)
    want = O.coil_kspaces(b, acq)
    nx, ny = b.acq_matrix
    mask = O.sampling_mask(nx, ny, acq)
    for c in range(got.shape[0]):
        e = rel(got[c], want[c])
        assert e <= TOL, f"coil {c}: rel-L2 {e:.2e}"
        assert np.all(got[c][~mask] == 0)


@pytest.mark.parametrize("shape,name", CASES, ids=[f"{O.GOLDEN_SHAPES[s][0]}x{O.GOLDEN_SHAPES[s][1]}-{n}"
                                                     for s, n in CASES])
def test_gpu_kspace_reconstructs_to_golden_image(shape, name):
    # simulate_slice (src/kspace.rs:1229-1284) with every coil k-space from the GPU, the
    # reconstruction from the oracle, against the reference's golden record.
    import os
    nx, ny, o = O.GOLDEN_SHAPES[shape]
    _, acq, ed, maps = {c[0]: c for c in O.golden_cases()}[name]
    b = O.golden_slice_builder(shape, ed, maps)
    gold = np.fromfile(os.path.join(os.path.dirname(__file__), "golden", "p4_baseline", "default",
                                    f"slice-{nx}x{ny}o{o}-{name}.bin"), dtype=np.float32)
    gold = gold[0::2] + 1j * gold[1::2]

    def recon(ref_coils, coils):
        w = None
        if acq.accel > 1:
            lines = O.sampling_mask(1, ny, acq.copy(accel=1, n_spikes=0))
            w = O.grappa_fit(ref_coils, nx, ny, acq.accel, acq.acs_lines, lines)
        return O.reconstruct(coils, nx, ny, acq, w)

    ref_b = O.golden_slice_builder(shape, None, maps, slice_seed=7 ^ 0x47525043414C)
    acq_ref = acq.copy(accel=1, n_spikes=0)
    g_ref = gpu.build_coil_kspaces(ref_b, acq_ref) if acq.accel > 1 else None
    g = gpu.build_coil_kspaces(b, acq)
    e = rel(recon(g_ref, g), gold)
    if acq.accel <= 1:
        assert e <= TOL, f"rel-L2 {e:.2e}"
        return
    # GRAPPA (src/kspace.rs:1540-1633) fits a regularised least-squares kernel on the reference
    # scan, which amplifies k-space error.  Bound: the image error the same fit gives when the
    # oracle's own k-spaces carry random relative errors of the GPU's measured size.
    o_ref, o = O.coil_kspaces(ref_b, acq_ref), O.coil_kspaces(b, acq)
    ek = max(rel(g_ref, o_ref), rel(g, o))
    rng = np.random.default_rng(0)
    pert = lambda k: k + ek * np.sqrt(np.mean(np.abs(k[k != 0]) ** 2)) * (k != 0) * (
        rng.standard_normal(k.shape) + 1j * rng.standard_normal(k.shape)) / np.sqrt(2)
    e_pert = rel(recon(pert(o_ref), pert(o)), gold)
    print(f"{name}: k-space rel-L2 {ek:.2e}, image rel-L2 {e:.2e}, oracle perturbed by the same: {e_pert:.2e}")
    assert e <= max(TOL, 3.0 * e_pert), f"rel-L2 {e:.2e} vs GRAPPA amplification bound {3 * e_pert:.2e}"


def test_single_coil_call_spikes_timing_and_log_amp():
    comps, fmap, phase0, t2m, tim = O.golden_inputs(0)
    nx, ny, o = O.GOLDEN_SHAPES[0]
    b = abi.SliceInputBuilder(list(comps), [t2m, 100.0, t2m], fmap, (nx * o, ny * o), (nx, ny), z=3, nz=9,
                              t_inhom=[tim, 45.0, tim], phase0=phase0, eddy_drive=(0.3, -0.8, 0.5), slice_seed=11)
    acq = abi.Acquisition.default(n_coils=4, eddy_strength=1.0, ghost_offset=0.01, n_spikes=5, noise_variance=1.0,
                                  echo=abi.ECHO_GRADIENT, t_echo=40.0)
    got = gpu.build_coil_kspace(b, acq, coil=2, ncoils=4)
    want = O.coil_kspace(b, acq, 2, 4)
    # spikes (full-peak samples) at other positions would leave an O(1) relative error
    assert rel(got, want) <= TOL
    nxny = nx * ny
    # explicit timing table (a shifted echo) and a log-amplitude map
    t, trf, tread, pol = O.line_timing(acq, nx, ny)
    timing = (t, trf + 3.0, tread, pol)
    la = 0.1 * np.sin(np.arange(nx * o * ny * o) / 17.0)
    acq2 = acq.copy(n_spikes=0, noise_variance=0.0)
    got = gpu.build_coil_kspace(b, acq2, coil=1, ncoils=4, timing=timing, log_amp=la)
    want = O.coil_kspace(b, acq2, 1, 4, timing=timing, log_amp=la)
    assert rel(got, want) <= TOL
    assert nxny == got.size


def _plan_vs_oracle(w, precision, units):
    import torch
    plan, outs, keep = workloads.to_plan(w, precision)
    plan.forward()
    torch.cuda.synchronize()
    errs = []
    for u in units:
        want = O.forward_workload_unit(w, u)
        errs.append(rel(outs[u].cpu().numpy(), want))
    plan.close()
    return errs


@pytest.mark.parametrize("precision", [abi.PREC_SPLIT_BF16, abi.PREC_FP32])
def test_batch_128_units_match_fp64_oracle(precision):
    w = workloads.make("batch_128", n_slices=2, n_volumes=2)
    errs = _plan_vs_oracle(w, precision, [0, 3])
    print(f"batch_128 precision={precision} rel-L2 vs FP64: {errs}", flush=True)
    assert max(errs) <= TOL


def test_full_600_unit_batch_matches_oracle_and_scales_exactly():
    # The benchmark's whole batch (30 slices x 20 volumes): 12 units spread over slices and volumes
    # against the FP64 oracle, both y-stage paths on every unit, and doubled images giving exactly
    # doubled k-spaces on every unit (the forward is linear in the images and a power-of-two scale
    # passes exactly through the BF16 split and the FP32 accumulation).
    import torch
    w = workloads.make("batch_128")
    assert len(w.units) == 600
    plan, outs, keep = workloads.to_plan(w)
    res = {}
    for basis in (True, False):
        plan.set_basis(basis)
        plan.set_tiles(8, 4, 1 if basis else 4)
        plan.forward_graph()
        torch.cuda.synchronize()
        res[basis] = [o.clone() for o in outs]
    worst = max(float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))
                for a, b in zip(res[True], res[False]))
    print(f"600 units: max rel diff slice basis vs per-unit sets {worst:.2e}", flush=True)
    assert worst <= TOL
    for u in np.linspace(0, 599, 12).astype(int):
        e = rel(res[True][u].cpu().numpy(), O.forward_workload_unit(w, int(u)))
        print(f"  unit {u}: rel-L2 vs FP64 {e:.2e}", flush=True)
        assert e <= TOL
    # the units' image tensors (the slices' inputs are dicts) doubled in place: the y-stage reads
    # them directly, so every output must come out exactly doubled
    n = 0
    for t in keep:
        if isinstance(t, torch.Tensor):
            t.mul_(2.0)
            n += 1
    assert n == 600
    plan.set_basis(True)
    plan.set_tiles(8, 4, 1)
    plan.forward()
    torch.cuda.synchronize()
    bad = sum(int(not torch.equal(o, 2 * r)) for o, r in zip(outs, res[True]))
    print(f"  doubled images: {600 - bad} of 600 units exactly doubled", flush=True)
    assert bad == 0
    plan.close()


def test_slice_128_matches_fp64_oracle():
    w = workloads.make("slice_128")
    for prec in (abi.PREC_SPLIT_BF16, abi.PREC_FP32):
        errs = _plan_vs_oracle(w, prec, [0])
        print(f"slice_128 precision={prec} rel-L2 vs FP64: {errs}", flush=True)
        assert max(errs) <= TOL


def test_pf_256_matches_fp64_oracle():
    w = workloads.make("pf_256", n_slices=1)
    for prec in (abi.PREC_SPLIT_BF16, abi.PREC_FP32):
        errs = _plan_vs_oracle(w, prec, [0])
        print(f"pf_256 precision={prec} rel-L2 vs FP64: {errs}", flush=True)
        assert max(errs) <= TOL


@pytest.mark.parametrize("name", ["tile_96", "tile_192"])
def test_three_row_tile_x_stage_matches_fp64_oracle(name):
    # 48 kx in one block of three 16-row tiles; 96 kx in one block of six tiles and eight waves
    w = workloads.make(name)
    errs = _plan_vs_oracle(w, abi.PREC_SPLIT_BF16, [0])
    print(f"{name} rel-L2 vs FP64: {errs}", flush=True)
    assert max(errs) <= TOL


@pytest.mark.parametrize("name,ghost,mt,nw", [("tile_192", 0.0, 6, 8), ("tile_192", 0.015, 3, 8),
                                              ("slice_128", 0.015, 4, 8), ("pf_256", 0.015, 1, 4)])
def test_x_stage_block_shapes_match_fp64_oracle(name, ghost, mt, nw):
    # The planner's x-stage blocks hold 16*MT kx rows of twiddles (both ghost tables where the group
    # has a ghost): four waves within 80 KB of LDS, or eight waves within 156 KB where that needs fewer
    # kx blocks.  Covered: 96 kx in one eight-wave block; with ghost tables 96 kx in two eight-wave
    # blocks of 48 rows (restaging between kx blocks), 64 kx in one, and 128 kx of a 256-column grid
    # in eight four-wave blocks of 16 rows.  Each against the FP64 oracle.
    w = workloads.make(name, n_slices=1) if name == "pf_256" else workloads.make(name)
    if ghost:
        w.acq = w.acq.copy(ghost_offset=ghost)
    plan, outs, keep = workloads.to_plan(w, abi.PREC_SPLIT_BF16)
    desc = plan.describe()
    plan.close()
    assert f"x-stage class 0: {mt} kx tiles and {nw} waves per block" in desc, desc
    errs = _plan_vs_oracle(w, abi.PREC_SPLIT_BF16, [0])
    print(f"{name} ghost {ghost}: MT {mt}, {nw} waves, rel-L2 vs FP64: {errs}", flush=True)
    assert max(errs) <= TOL


def test_complex_images_match_fp64_oracle():
    w = workloads.make("grad_128")
    errs = _plan_vs_oracle(w, abi.PREC_SPLIT_BF16, [0])
    print(f"grad_128 rel-L2 vs FP64: {errs}", flush=True)
    assert max(errs) <= TOL


def test_graph_replay_tuning_and_determinism():
    import torch
    w = workloads.make("batch_128", n_slices=2, n_volumes=3)
    plan, outs, keep = workloads.to_plan(w)
    plan.forward()
    torch.cuda.synchronize()
    ref = [o.clone() for o in outs]
    for rep in range(3):
        for o in outs:
            o.zero_()
        torch.cuda.synchronize()
        plan.forward_graph()
        torch.cuda.synchronize()
        for o, r in zip(outs, ref):
            assert torch.equal(o, r), f"graph replay {rep} differs"
        _ = (ref[0].abs() * 2).sum().item()  # eager work between replays
    report = plan.tune(trials=2)
    print(report, flush=True)
    # the pick is the finalist (graph replays, round-robin) with the lowest median, and the plan keeps it
    best = [ln for ln in report.splitlines() if ln.startswith("best ")][-1]
    kv = dict(t.split("=") for t in best.split()[1:5])
    assert f"finalist basis={kv['basis']} xt={kv['xt']} gp={kv['gp']} nv={kv['nv']} graph median" in report, report
    assert plan.tiles() == (int(kv["xt"]), int(kv["gp"]), int(kv["nv"]))
    plan.forward()
    torch.cuda.synchronize()
    # the x tile only redistributes work (bitwise equal); line tiles and unit sets change the
    # evaluation path, so the tuned result agrees to round-off
    for o, r in zip(outs, ref):
        assert rel(o.cpu().numpy(), r.cpu().numpy()) <= 1e-6, "tile choice changed the result"
    plan.close()


def _set_plan_outputs(w, tiles, basis=False):
    import torch
    plan, outs, keep = workloads.to_plan(w)
    plan.set_basis(basis)
    plan.set_tiles(*tiles)
    assert plan.tiles() == tuple(tiles)
    plan.forward()
    torch.cuda.synchronize()
    res = [o.cpu().numpy() for o in outs]
    desc = plan.describe()
    plan.close()
    return res, desc


@pytest.mark.parametrize("name,kw,nv", [("batch_128", dict(n_slices=3, n_volumes=6), 2),
                                        ("batch_128", dict(n_slices=2, n_volumes=7), 4),
                                        ("unitsets_q1", dict(), 4), ("unitsets_q1", dict(), 2),
                                        ("unitsets_cplx", dict(), 4)])
def test_unit_sets_match_single_units_and_oracle(name, kw, nv):
    # Units of one slice evaluated together in sets of nv (the phase and decay shared): every unit,
    # in full sets and leftovers, equals its single-unit evaluation to round-off and the FP64 oracle.
    if name == "unitsets_q1":
        w = workloads.make("batch_128", n_slices=2, n_volumes=7)
        w.n_comp = 1
        w.units = [(si, img[:1].copy()) for si, img in w.units]
        for sd in w.slices:
            sd.t2, sd.t_inhom = sd.t2[:1].copy(), sd.t_inhom[:1].copy()
    elif name == "unitsets_cplx":
        w = workloads.make("batch_128", n_slices=2, n_volumes=5)
        w.n_comp, w.complex_images = 1, True
        w.units = [(si, (img[:1] * np.exp(1j * w.slices[si].phase0)[None, :]).astype(np.complex64))
                   for si, img in w.units]
        for sd in w.slices:
            sd.t2, sd.t_inhom = sd.t2[:1].copy(), sd.t_inhom[:1].copy()
    else:
        w = workloads.make(name, **kw)
    single, _ = _set_plan_outputs(w, (16, 4, 1))
    sets, desc = _set_plan_outputs(w, (16, 4, nv))
    assert f"units per item {nv}" in desc
    worst = max(rel(a, b) for a, b in zip(sets, single))
    print(f"{name} sets of {nv}: max rel diff to single units {worst:.2e}", flush=True)
    assert worst <= 1e-6
    for u in (0, len(w.units) // 2, len(w.units) - 1):
        e = rel(sets[u], O.forward_workload_unit(w, u))
        print(f"  unit {u}: rel-L2 vs FP64 {e:.2e}", flush=True)
        assert e <= TOL


def test_chunked_streams_batch_matches_oracle_and_replays():
    # 64 units: the plan splits them into 4 chunks whose y -> x stages run on forked streams
    import torch
    w = workloads.make("batch_128", n_slices=4, n_volumes=16)
    plan, outs, keep = workloads.to_plan(w)
    plan.forward()
    torch.cuda.synchronize()
    ref = [o.clone() for o in outs]
    for u in (0, 37, 63):
        want = O.forward_workload_unit(w, u)
        e = rel(outs[u].cpu().numpy(), want)
        assert e <= TOL, f"unit {u}: rel-L2 {e:.2e}"
    for rep in range(2):
        for o in outs:
            o.zero_()
        torch.cuda.synchronize()
        plan.forward_graph()
        torch.cuda.synchronize()
        assert all(torch.equal(o, r) for o, r in zip(outs, ref)), f"replay {rep}"
    plan.close()


def test_echoes_share_static_planes_until_an_input_changes():
    # Two gradient-echo groups (TE 20 and 32 ms, distortions on) whose slices bind the same fieldmap,
    # coil maps and T2: the plan gives them one set of static planes, with each echo's f*TE phase on
    # its lines.  Rebinding one slice's fieldmap must split the set again.  Both states against the
    # FP64 oracle at the split-BF16 tolerance.
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    rng = np.random.default_rng(12)
    snx = sny = 32
    nx = ny = 16
    C, Q = 4, 1
    P = snx * sny
    acqs = [abi.Acquisition.default(echo=abi.ECHO_GRADIENT, t_echo=te, t_line=0.5, signal_scale=1.0, n_coils=C)
            for te in (20.0, 32.0)]
    fmap = (rng.standard_normal(P) * 20.0).astype(np.float32)
    fmap2 = (rng.standard_normal(P) * 20.0).astype(np.float32)
    maps = (rng.standard_normal((C, P)) + 1j * rng.standard_normal((C, P))).astype(np.complex64)
    t2 = rng.uniform(60.0, 90.0, size=(Q, P)).astype(np.float32)
    img = rng.uniform(0.0, 1.0, size=(Q, P)).astype(np.float32)
    d = lambda a: torch.from_numpy(a).cuda()
    f_d, f2_d, m_d, t2_d, img_d = d(fmap), d(fmap2), d(maps), d(t2), d(img)
    outs = [torch.empty((C, ny, nx), dtype=torch.complex64, device="cuda") for _ in range(2)]
    groups = [gpu.Group(snx, sny, nx, ny, C, Q, a) for a in acqs]
    slices = [gpu.Slice(g, f_d, m_d, t2=t2_d) for g in range(2)]
    units = [gpu.Unit(g, img_d, outs[g]) for g in range(2)]
    plan = gpu.Plan(groups, slices, units)

    def check(fmaps):
        plan.forward()
        torch.cuda.synchronize()
        for g in range(2):
            want = O.forward_direct(snx, sny, nx, ny, C, Q, acqs[g], img, fmaps[g], maps, t2=t2)
            got = outs[g].cpu().numpy()
            err = np.linalg.norm(got - want) / np.linalg.norm(want)
            print(f"echo {g}: rel-L2 {err:.2e}", flush=True)
            assert err <= 1e-5

    check([fmap, fmap])
    plan.set_pointers(slices=[slices[0], gpu.Slice(1, f2_d, m_d, t2=t2_d)])
    check([fmap, fmap2])
    plan.close()


def _basis_workload(name):
    """Workloads of the slice-basis y-stage: the volumes of each slice in one kernel."""
    from mrsim_amd.workloads import Stream, Workload, _slice, _volume_images
    if name == "batch_q3_v20":
        return workloads.make("batch_128", n_slices=2, n_volumes=20)
    if name == "batch_q3_v7":      # an odd set: the last pair has one unit
        return workloads.make("batch_128", n_slices=2, n_volumes=7)
    if name == "batch_q3_v45":     # three sets of 15 units per slice
        return workloads.make("batch_128", n_slices=1, n_volumes=45)
    if name == "batch_q3_v18":     # 9 pairs: the last round's single pair split one line tile per wave
        return workloads.make("batch_128", n_slices=2, n_volumes=18)
    if name == "batch_q3_v11":     # 6 pairs, the last with one unit: the last round split two tiles per wave
        return workloads.make("batch_128", n_slices=2, n_volumes=11)
    # small grids: 40^2 -> 20^2 (three 16-row chunks, the last one ragged; 20 lines padded to 32: the
    # upper line tiles are idle),
    # 3 coils (rows of absent coils), Q = 1, 2 or 4, real or complex images
    Q = {"small_q1": 1, "small_q2_cplx": 2, "small_q4": 4}[name]
    cplx = name.endswith("cplx")
    C = 3
    acq = abi.Acquisition.default(t_line=0.5, t_echo=30.0, signal_scale=1.0, n_coils=C)
    w = Workload(name, 40, 40, 20, 20, C, Q, acq, complex_images=cplx)
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


@pytest.mark.parametrize("name", ["batch_q3_v20", "batch_q3_v7", "batch_q3_v45", "batch_q3_v18", "batch_q3_v11",
                                  "small_q1", "small_q2_cplx", "small_q4"])
def test_slice_basis_matches_per_unit_and_oracle(name):
    # The volumes of a slice in the slice-basis y-stage (per-slice phase and decay factors shared by
    # every unit, per-unit coil-weighted images on the MFMA) against the per-unit y-stage and the
    # FP64 oracle, both at the split-BF16 tolerance.
    w = _basis_workload(name)
    per_unit, _ = _set_plan_outputs(w, (8, 4, 1), basis=False)
    basis, desc = _set_plan_outputs(w, (8, 4, 1), basis=True)
    assert "slice basis on" in desc and " 0 unit sets" not in desc, desc
    worst = max(rel(a, b) for a, b in zip(basis, per_unit))
    print(f"{name}: max rel diff basis vs per-unit {worst:.2e}", flush=True)
    assert worst <= TOL
    for u in sorted({0, 1, len(w.units) // 2, len(w.units) - 1}):
        e = rel(basis[u], O.forward_workload_unit(w, u))
        print(f"  unit {u}: rel-L2 vs FP64 {e:.2e}", flush=True)
        assert e <= TOL
    nx, ny = w.nx, w.ny
    mask = O.sampling_mask(nx, ny, w.acq).reshape(ny, nx)
    for k in basis:
        assert np.all(k[:, ~mask] == 0)


def test_log_amp_closed_forms_on_the_gpu():
    # src/kspace.rs:959-964: log_amp adds to the exponent of every voxel's relaxation weight.  On the
    # GPU drop-in: log_amp = c everywhere gives e^c times the plain k-space, and log_amp = ln s gives
    # the k-space of the compartments scaled by s (s a power of two: the scaled f32 inputs are exact).
    comps, fmap, phase0, t2m, tim = O.golden_inputs(1)
    nx, ny, o = O.GOLDEN_SHAPES[1]
    P = nx * o * ny * o
    s = 2.0 ** np.random.default_rng(5).integers(-2, 3, size=P)

    def mk(scale=None):
        cs = [(c * scale).astype(np.float32) for c in comps] if scale is not None else list(comps)
        return abi.SliceInputBuilder(cs, [t2m] * len(cs), fmap, (nx * o, ny * o), (nx, ny), z=3, nz=9,
                                     t_inhom=[tim] * len(cs), phase0=phase0, slice_seed=7)

    acq = abi.Acquisition.default(n_coils=3)
    k0 = gpu.build_coil_kspace(mk(), acq, coil=1, ncoils=3)
    kc = gpu.build_coil_kspace(mk(), acq, coil=1, ncoils=3, log_amp=np.full(P, 0.37))
    assert rel(kc, np.exp(0.37) * k0) <= TOL
    k_la = gpu.build_coil_kspace(mk(), acq, coil=2, ncoils=3, log_amp=np.log(s))
    k_sc = gpu.build_coil_kspace(mk(s), acq, coil=2, ncoils=3)
    assert rel(k_la, k_sc) <= TOL
    # and both against the oracle's forward with the same log-amplitude map
    assert rel(k_la, O.coil_kspace(mk(), acq, 2, 3, log_amp=np.log(s))) <= TOL
