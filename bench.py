"""Benchmark of the batched exact k-space forward on one MI355X.

Workload (default): 30 slices x 20 volumes x 8 coils x 3 compartments, 128^2 sim grid -> 64^2
acquired, T2/T2' maps per compartment, spin-echo EPI (t_line 0.5 ms, TE 30 ms), synthetic inputs.
A step is one forward of the whole batch (600 slice units) replayed from a HIP graph, inputs
resident in HBM.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "mrsim-acq_amd"))
sys.path.insert(0, ROOT)

# MI355X published peaks
PEAK_HBM_GBS = 8000.0
PEAK_FP32_TF = 157.3        # FP32 (vector and dense matrix)
PEAK_BF16_MFMA_TF = 2500.0  # dense bf16 MFMA


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _pool_rate(fn, n, cores):
    """Units per second of fn(i) for i < n over `cores` forked worker processes."""
    import multiprocessing as mp
    ctx = mp.get_context("fork")
    with ctx.Pool(cores) as pool:
        pool.map(fn, range(cores))   # warm-up: every worker loads the oracle
        t0 = time.perf_counter()
        pool.map(fn, range(n))
        return n / (time.perf_counter() - t0)


def cpu_baseline(budget_s: float = 8.0):
    """The oracle's FP64 restatement of the reference's CPU path, timed on the host cores:
      default_exact      config-1 units (rotor y-stage, twiddle-table x-stage: the default build)
      kspace_fft_xstage  the same units with the `kspace` feature's FFT x-stage (its NUFFT y-stage needs
                         uniform relaxation, which T2/T2' maps rule out: the rotor y-stage stays)
      config0_1core / config0_all_cores  one 128^2 -> 64^2 slice, 1 coil, 1 compartment, T2/T2' maps
    One process per unit, as the reference's `par` runs volumes, on min(nproc, 16) processes (the
    CPU share of one GPU on the benchmark machines).  Each leg is a bounded sample of about
    `budget_s` seconds."""
    nproc = os.cpu_count() or 1
    cores = min(nproc, 16)
    legs = {}
    for name, fn in (("default_exact", _cpu_unit), ("kspace_fft_xstage", _cpu_unit_fft)):
        t0 = time.perf_counter()
        fn(0)
        t1 = time.perf_counter() - t0
        n = max(cores, min(8 * cores, int(budget_s * cores / max(t1, 1e-3)) // cores * cores))
        legs[name] = {"slices_per_s": round(_pool_rate(fn, n, cores), 2), "units": n,
                      "one_unit_one_core_s": round(t1, 3)}
    _cpu_config0(0)
    t0 = time.perf_counter()
    k = 0
    while time.perf_counter() - t0 < budget_s / 4 or k < 3:
        _cpu_config0(k)
        k += 1
    c0_one = k / (time.perf_counter() - t0)
    n0 = max(cores, min(64 * cores, int(c0_one * cores * budget_s / 2) // cores * cores))
    legs["config0_1core"] = {"slices_per_s": round(c0_one, 2), "slices": k}
    legs["config0_all_cores"] = {"slices_per_s": round(_pool_rate(_cpu_config0, n0, cores), 2), "slices": n0}
    d = legs["default_exact"]
    return {"value": d["slices_per_s"], "unit": "slices/s", "cores": cores, "kind": "port",
            "sample": f"{d['units']} slice units of the batch workload (128^2->64^2, 8 coils, 3 compartments, "
                      f"T2/T2' maps): the oracle's FP64 restatement of build_coil_kspace_timed x 8 coils, "
                      f"default exact-DFT build, {cores} processes",
            "nproc": nproc, "cpu_model": _cpu_model(), "legs": legs}


def _cpu_inputs(i):
    import numpy as np

    from mrsim_amd import abi
    from mrsim_amd import workloads as W
    st = W.Stream(0xC0DE + i)
    snx = sny = 128
    P = snx * sny
    comps = [st.unit(P).astype(np.float32) for _ in range(3)]
    t2 = [(40 + 80 * st.unit(P)).astype(np.float32) for _ in range(3)]
    ti = [(20 + 60 * st.unit(P)).astype(np.float32) for _ in range(3)]
    fmap = (100 * (st.unit(P) - 0.5)).astype(np.float32)
    b = abi.SliceInputBuilder(comps, t2, fmap, (snx, sny), (64, 64), t_inhom=ti, phase0=st.unit(P) * 6.28)
    acq = abi.Acquisition.default(t_line=0.5, t_echo=30.0, n_coils=8)
    return b, acq


def _cpu_unit(i):
    from oracle import oracle as O
    b, acq = _cpu_inputs(i)
    for c in range(8):
        O.coil_kspace(b, acq, c, 8)
    return i


def _cpu_unit_fft(i):
    from oracle import oracle as O
    b, acq = _cpu_inputs(i)
    with O.fft_xstage():
        for c in range(8):
            O.coil_kspace(b, acq, c, 8)
    return i


_C0 = None


def _cpu_config0(i):
    """Config 0 on the CPU: one slice, 128^2 -> 64^2, 1 coil, 1 compartment, T2/T2' maps, complex128."""
    global _C0
    from mrsim_amd import abi, workloads
    from oracle import oracle as O
    if _C0 is None:
        w = workloads.make("slice_128")
        sd = w.slices[0]
        _C0 = (abi.SliceInputBuilder([w.units[0][1][0]], [sd.t2[0]], sd.fmap, (128, 128), (64, 64),
                                     t_inhom=[sd.t_inhom[0]], phase0=sd.phase0), w.acq)
    O.coil_kspace(_C0[0], _C0[1], 0, 1)
    return i


_PF = {}


def _oracle_pf256_unit(args):
    """FP64 oracle k-space of one unit of the pf_256 workload (for the per-slice error of config 2)."""
    from mrsim_amd import workloads
    from oracle import oracle as O
    n_slices, u = args
    if n_slices not in _PF:
        _PF[n_slices] = workloads.make("pf_256", n_slices=n_slices)
    return O.forward_workload_unit(_PF[n_slices], u)


def config2_pass(steps: int, check: bool = True):
    """Config 2: 24 slices of 256^2 -> 128^2, 16 coils, partial Fourier 6/8 (97 of 128 lines), Hann
    recon window.  The batched split-BF16 forward (graph replay) and the plain FP32 per-slice forward,
    with the relative L2 error of every slice's k-space against the FP64 oracle for both; the window is
    a reconstruction factor (src/kspace.rs:1327-1340), applied by the reconstruction, timed here too."""
    import multiprocessing as mp

    import numpy as np
    import torch

    from mrsim_amd import abi, workloads
    from mrsim_amd import gpu as G
    w = workloads.make("pf_256")
    ns = len(w.units)
    res = {"workload": "pf_256: 24 slices, 256^2->128^2, 16 coils, 1 compartment, PF 6/8 (97 lines), Hann window"}
    outs_by = {}
    for prec, name in ((abi.PREC_SPLIT_BF16, "split_bf16"), (abi.PREC_FP32, "plain_fp32")):
        plan, outs, keep = workloads.to_plan(w, prec)
        run = plan.forward_graph if prec == abi.PREC_SPLIT_BF16 else plan.forward
        if prec == abi.PREC_SPLIT_BF16:
            plan.tune(trials=3)
        for _ in range(2):
            run()
        torch.cuda.synchronize()
        n = max(3, steps)
        t0 = time.perf_counter()
        for _ in range(n):
            run()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / n
        res[name] = {"ms_per_batch": round(1e3 * dt, 4), "slices_per_s": round(ns / dt, 1)}
        outs_by[name] = [o.cpu().numpy() for o in outs]
        plan.close()
    rec = G.Recon(w.nx, w.ny, w.n_coils, w.acq.copy(window=abi.WINDOW_HANN))
    plan, outs, keep = workloads.to_plan(w, abi.PREC_SPLIT_BF16)
    plan.forward()
    imgs = [torch.empty((w.ny, w.nx), dtype=torch.complex64, device="cuda") for _ in outs]
    rec.run(outs, imgs)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(max(3, steps)):
        rec.launch()
    torch.cuda.synchronize()
    res["recon_hann_ms_per_batch"] = round(1e3 * (time.perf_counter() - t0) / max(3, steps), 4)
    rec.close()
    plan.close()
    if check:
        cores = min(os.cpu_count() or 1, 16)
        with mp.get_context("fork").Pool(cores) as pool:
            want = pool.map(_oracle_pf256_unit, [(ns, u) for u in range(ns)])
        for name in ("split_bf16", "plain_fp32"):
            e = [float(np.linalg.norm(g - r) / np.linalg.norm(r)) for g, r in zip(outs_by[name], want)]
            res[name]["rel_l2_vs_fp64_per_slice"] = [float(f"{x:.3e}") for x in e]
            res[name]["rel_l2_max"] = float(f"{max(e):.3e}")
    return res


def nufft_pass(iters: int):
    """The NUFFT comparator (uniform relaxation: the reference's `kspace` feature y-stage) on one slice,
    128^2 -> 64^2, 8 coils, 3 compartments of uniform T2: the FP64 NUFFT y-stage kernel alone (8 coils
    x 128 columns of 128 displaced sources, 6 weight vectors each, 64 outputs) and the two host-buffer
    drop-ins, the NUFFT forward (FP64) and the exact forward (split-BF16 plan), on the same slice."""
    import torch

    from mrsim_amd import abi
    from mrsim_amd import gpu as G
    from mrsim_amd import workloads
    w = workloads.make("batch_128", n_slices=1, n_volumes=1)
    sd = w.slices[0]
    acq = w.acq.copy()
    b = abi.SliceInputBuilder([im for im in w.units[0][1]], [80.0, 45.0, 120.0], sd.fmap, (w.snx, w.sny),
                              (w.nx, w.ny), phase0=sd.phase0)
    def timed(fn, n):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(n):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / n * 1e3
    ncols, nw, nsrc = w.n_coils * w.snx, 2 * w.n_comp, w.sny
    pos = torch.randn((w.snx, nsrc), dtype=torch.float64, device="cuda") + torch.arange(nsrc, device="cuda")
    wts = torch.randn((ncols, nw, nsrc), dtype=torch.complex128, device="cuda")
    ms_kernel = timed(lambda: G.nufft1(pos, wts, nsrc, -(w.ny // 2), w.ny, 1e-13), iters)
    ms_nufft = timed(lambda: G.build_coil_kspaces_nufft(b, acq), max(3, iters // 4))
    ms_exact = timed(lambda: G.build_coil_kspaces(b, acq), max(3, iters // 4))
    return {"workload": f"1 slice {w.snx}^2->{w.nx}^2, {w.n_coils} coils, {w.n_comp} compartments, uniform T2",
            "nufft_ystage_kernel_ms": round(ms_kernel, 4), "nufft_forward_dropin_ms": round(ms_nufft, 3),
            "exact_forward_dropin_ms": round(ms_exact, 3),
            "note": "FP64 NUFFT (tol 1e-13, width 14) vs split-BF16 exact; drop-ins include host transfers"}


def grad_batch_pass(plan, w, outs, iters: int):
    """The headline batch (600 units) forward and forward + adjoint: gradients for every unit's images,
    fieldmap and T2' from a seeded complex cotangent, the adjoint reusing the forward's slice planes
    (mrs_plan_vjp_after_forward).  Both timed as whole-batch passes."""
    import torch
    gen = torch.Generator(device="cuda").manual_seed(7)
    gout = [torch.randn(o.shape, dtype=o.dtype, device="cuda", generator=gen) for o in outs]
    P, Q = w.snx * w.sny, w.n_comp
    gi = [torch.empty((Q, P), dtype=torch.float32, device="cuda") for _ in outs]
    gf = [torch.empty(P, dtype=torch.float32, device="cuda") for _ in outs]
    gt = [torch.empty((Q, P), dtype=torch.float32, device="cuda") for _ in outs]
    arrays = plan.pointer_arrays(gout, gi, gf, gt)
    for _ in range(2):
        plan.forward_graph()
        plan.vjp_arrays(arrays, after_forward=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(iters):
        plan.forward_graph()
    torch.cuda.synchronize()
    t_fwd = (time.perf_counter() - t0) / iters
    t0 = time.perf_counter()
    for _ in range(iters):
        plan.forward_graph()
        plan.vjp_arrays(arrays, after_forward=True)
    torch.cuda.synchronize()
    t_both = (time.perf_counter() - t0) / iters
    return {"workload": f"the forward's batch: {len(outs)} units 128^2->64^2, 8 coils, 3 compartments, real images; "
                        "gradients for images, fieldmap and T2' per unit",
            "forward_ms": round(1e3 * t_fwd, 4), "forward_adjoint_ms": round(1e3 * t_both, 4),
            "ratio": round(t_both / t_fwd, 3)}


def grad_pass(iters: int):
    """One slice, 128^2 -> 64^2, 8 coils, 3 compartments, complex64 images: forward time, forward +
    adjoint (gradients for images, fieldmap and T2') time, and a timed 50-iteration Adam fit of the
    fieldmap to the k-space of the true one."""
    import torch

    from mrsim_amd import workloads
    from mrsim_amd.autograd import KspaceModel
    w = workloads.make("grad_128")
    sd = w.slices[0]
    dev = "cuda"
    model = KspaceModel(w.snx, w.sny, w.nx, w.ny, torch.from_numpy(sd.coil_maps).to(dev), w.acq, w.n_comp,
                        t2=torch.from_numpy(sd.t2).to(dev), complex_images=True)
    img = torch.from_numpy(w.units[0][1]).to(dev)
    fmap = torch.from_numpy(sd.fmap).to(dev)
    ti = torch.from_numpy(sd.t_inhom).to(dev)
    g = torch.randn((w.n_coils, w.ny, w.nx), dtype=torch.complex64, device=dev)
    for _ in range(3):
        model.forward_raw(img, fmap, ti)
        model.vjp_raw(img, fmap, ti, g)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(iters):
        model.forward_raw(img, fmap, ti)
    torch.cuda.synchronize()
    t_fwd = (time.perf_counter() - t0) / iters
    t0 = time.perf_counter()
    for _ in range(iters):
        model.forward_raw(img, fmap, ti)
        model.vjp_raw(img, fmap, ti, g)
    torch.cuda.synchronize()
    t_both = (time.perf_counter() - t0) / iters
    target = model.forward_raw(img, fmap, ti).detach()

    def fit(f, n):
        opt = torch.optim.Adam([f], lr=0.5, fused=True)
        for _ in range(n):
            opt.zero_grad()
            loss = (model(img, f, ti) - target).abs().pow(2).sum()
            loss.backward()
            opt.step()
        return loss

    # untimed: the first iterations load torch's loss and optimizer kernels
    fit((fmap + 5.0).clone().requires_grad_(True), 2)
    f = (fmap + 5.0).clone().requires_grad_(True)
    opt = torch.optim.Adam([f], lr=0.5, fused=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(50):
        opt.zero_grad()
        loss = (model(img, f, ti) - target).abs().pow(2).sum()
        loss.backward()
        opt.step()
    l_end = float(loss.detach())
    torch.cuda.synchronize()
    t_fit = time.perf_counter() - t0
    return {"workload": "grad_128: 1 slice 128^2->64^2, 8 coils, 3 compartments, complex64",
            "forward_ms": round(1e3 * t_fwd, 4), "forward_adjoint_ms": round(1e3 * t_both, 4),
            "ratio": round(t_both / t_fwd, 3), "fit50_ms": round(1e3 * t_fit, 2), "fit50_final_loss": l_end}


def ge3d_pass(steps: int, check: bool = True, tiles: str = "", stages_only: bool = False):
    """The 3D gradient-echo train (two volumes, 128^2 -> 64^2 and 192^2 -> 96^2, 24 partitions each,
    32 coils, 1 compartment, 4 echoes): one hipGraph replay per step runs every per-partition 2D
    forward and both volumes' partition transforms with their per-excitation weights."""
    import numpy as np
    import torch

    from mrsim_amd import workloads
    w = workloads.make_ge3d("ge3d")
    fw, keep = workloads.to_ge3d(w)
    if tiles:
        fw.plan.set_tiles(*(int(v) for v in tiles.split(",")))
    else:
        log(fw.tune(5).strip())
    log(fw.describe().strip())
    if stages_only:
        st = fw.stage_times(iters=steps)
        fw.close()
        return {"stage_ms": st, "tiles": tiles}
    for _ in range(2):
        fw.forward_graph()
    torch.cuda.synchronize()

    def fwd_loop(n):
        t0 = time.perf_counter()
        for _ in range(n):
            fw.forward_graph()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / n

    # best of three repetitions, as for forward + adjoint below
    dt = min(fwd_loop(max(4, steps)) for _ in range(3))
    st = fw.stage_times(iters=3)
    # forward + adjoint: the forward, then the vector-Jacobian product of a seeded complex cotangent
    # (partition-transform adjoint, per-unit adjoints, sums over echoes); the adjoint reads the slice
    # planes the forward just built (as a backward pass reads saved activations), and the plain
    # adjoint, which rebuilds them, is timed beside it
    rng = np.random.default_rng(5)
    gout = [[torch.from_numpy((rng.standard_normal(o.shape) + 1j * rng.standard_normal(o.shape))
                              .astype(np.complex64)).cuda() for o in ov] for ov in fw.out]

    def fa_loop(reuse, n):
        t0 = time.perf_counter()
        for _ in range(n):
            fw.forward_graph()
            fw.vjp(gout, after_forward=reuse)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / n

    fa_loop(True, 2)
    fa_loop(False, 2)
    # three interleaved repetitions of each path, the best of each (a step is a few ms)
    nfa = max(4, steps)
    reps = [(fa_loop(True, nfa), fa_loop(False, nfa)) for _ in range(3)]
    dt_fa = min(r[0] for r in reps)
    dt_fa_rebuild = min(r[1] for r in reps)
    n_e = len(w.echo_times_ms)
    parts = sum(v.nz for v in w.volumes) * n_e
    # partition transform traffic: read K2 and write K3, complex64, every (volume, echo)
    zbytes = sum(2 * 8 * v.nz * w.n_coils * v.ny * v.nx for v in w.volumes) * n_e
    err = None
    if check:
        from oracle import ge3d as G
        v = w.volumes[0]
        want = G.ge3d_kspace(v.snx, v.sny, v.nx, v.ny, v.nz, v.images, v.t2, v.t_inhom, v.fmap, None,
                             v.coil_maps[:, [7]], w.acq, w.echo_times_ms[0], w.ky_segments, w.t_line_ms,
                             v.weights.astype(np.float32))[:, 0]
        got = fw.out[0][0][:, 7].cpu().numpy()
        err = float(np.linalg.norm(got - want) / np.linalg.norm(want))
    fw.close()
    return {"workload": "ge3d: 2 volumes (128^2->64^2, 192^2->96^2) x 24 partitions x 4 echoes (TE 12/24/36/48), "
                        "32 coils in 2 rings along z, 4 ky segments, per-excitation weights",
            "ms_per_train": round(1e3 * dt, 4), "partition_slices_per_s": round(parts / dt, 1),
            "volumes_x_echoes_per_s": round(len(w.volumes) * n_e / dt, 2), "stage_ms": st,
            "forward_adjoint_ms": round(1e3 * dt_fa, 4), "forward_adjoint_over_forward": round(dt_fa / dt, 3),
            "forward_adjoint_rebuilding_planes_ms": round(1e3 * dt_fa_rebuild, 4),
            "zstage_gbs": round(zbytes / (st["zstage"] * 1e-3) / 1e9, 1) if st["zstage"] > 0 else None,
            "rel_l2_vs_fp64_coil7_vol0_echo0": err}


def _pmc_per_dispatch(outdir, counter, match):
    """Average per dispatch of `counter` over kernels whose name contains `match` (rocprofv3 csv)."""
    import csv
    import glob
    tot, ids = 0.0, set()
    for path in glob.glob(os.path.join(outdir, "**", "*counter_collection.csv"), recursive=True):
        for r in csv.DictReader(open(path)):
            if match in r["Kernel_Name"] and r["Counter_Name"].startswith(counter):
                tot += float(r["Counter_Value"])
                ids.add(r["Dispatch_Id"])
    return (tot / len(ids), len(ids)) if ids else (None, 0)


def measure_traffic(tiles: str, basis: bool, kernel: str, timeout: float = 300.0):
    """HBM bytes per dispatch of `kernel` in the batch forward run at `tiles` and `basis` (the
    configuration the timed run used): rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE, each run on its
    own in a child process, each scaled by what the same counter reads for a device copy of 256 MiB.
    None when rocprofv3 is missing or a pass fails."""
    import shutil
    import subprocess
    import tempfile
    if not shutil.which("rocprofv3"):
        return None
    me = [sys.executable, "-u", os.path.abspath(__file__)]
    child = me + ["--steps", "2", "--warmup", "1", "--no-cpu", "--no-plain", "--no-grad", "--no-ge3d",
                  "--no-nufft", "--no-config2", "--check-units", "0", "--no-traffic", "--tiles", tiles,
                  "--basis", "1" if basis else "0"]
    nbytes = 256 << 20
    copy = me + ["--calibrate-copy", str(nbytes)]
    env = dict(os.environ, TMPDIR=os.environ.get("TMPDIR", "/tmp"))
    got = {}
    with tempfile.TemporaryDirectory() as d:
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            for name, cmd, match in (("bench", child, kernel), ("copy", copy, "copyBuffer")):
                out = os.path.join(d, f"{name}_{counter}")
                log(f"traffic pass: {counter} on {name}")
                try:
                    r = subprocess.run(["rocprofv3", "--pmc", counter, "--output-format", "csv", "-d", out, "-o",
                                        "r", "--"] + cmd, capture_output=True, text=True, timeout=timeout, env=env)
                except subprocess.TimeoutExpired:
                    log(f"traffic pass {counter}/{name} timed out")
                    return None
                v, n = _pmc_per_dispatch(out, counter, match)
                if v is None:
                    log(f"traffic pass {counter}/{name} gave no counters (rc {r.returncode}): {r.stderr[-400:]}")
                    return None
                got[(counter, name)] = (v, n)
    # counters in KiB per dispatch; the copy moves nbytes each way
    f_scale = (nbytes / 1024) / got[("FETCH_SIZE", "copy")][0]
    w_scale = (nbytes / 1024) / got[("WRITE_SIZE", "copy")][0]
    fetch = got[("FETCH_SIZE", "bench")][0] * f_scale * 1024
    write = got[("WRITE_SIZE", "bench")][0] * w_scale * 1024
    return {"fetch_bytes": fetch, "write_bytes": write, "fetch_scale": round(f_scale, 4),
            "write_scale": round(w_scale, 4), "dispatches": got[("FETCH_SIZE", "bench")][1]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--slices", type=int, default=30)
    ap.add_argument("--volumes", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-grad", action="store_true")
    ap.add_argument("--no-nufft", action="store_true")
    ap.add_argument("--no-ge3d", action="store_true")
    ap.add_argument("--no-config2", action="store_true", help="skip the 256^2 16-coil partial-Fourier leg")
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 HBM-traffic passes")
    ap.add_argument("--basis", type=int, choices=(0, 1), default=None,
                    help="with --tiles: run qualifying slices in the slice-basis y-stage (1) or per unit (0)")
    ap.add_argument("--only-ge3d", action="store_true", help="run only the 3D gradient-echo train (profiling)")
    ap.add_argument("--only-grad-batch", action="store_true",
                    help="run only the batch forward + adjoint (profiling); --vjp-basis picks the y-adjoint")
    ap.add_argument("--vjp-basis", type=int, choices=(0, 1), default=1,
                    help="slice-basis y-adjoint (1) or the per-unit one (0) in --only-grad-batch")
    ap.add_argument("--ge3d-tiles", type=str, default="", help="tiles of the 3D train (skips its autotuner)")
    ap.add_argument("--stages-only", action="store_true",
                    help="only the per-stage serial launches (profiling: one dispatch per chunk at a time)")
    ap.add_argument("--check-units", type=int, default=2)
    ap.add_argument("--tiles", type=str, default="", help="x_tile,line_tiles[,units_per_item] (skips the autotuner)")
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="only run 5 device copies of this many bytes (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        a = torch.empty(args.calibrate_copy // 4, dtype=torch.float32, device="cuda").fill_(1.0)
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": args.calibrate_copy, "copies": 5}))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the forward has no cross-GPU step, and multi-GPU runs are not "
              "part of this build (each GPU would take its own share of units).", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    from mrsim_amd import abi, workloads
    torch.cuda.set_device(0)
    if args.only_ge3d:
        print(json.dumps({"ge3d": ge3d_pass(args.steps, check=args.check_units > 0, tiles=args.ge3d_tiles,
                                            stages_only=args.stages_only)}), flush=True)
        return
    t_gen = time.perf_counter()
    w = workloads.make("batch_128", n_slices=args.slices, n_volumes=args.volumes)
    log(f"workload generated in {time.perf_counter() - t_gen:.1f} s: {len(w.units)} units")
    plan, outs, keep = workloads.to_plan(w, abi.PREC_SPLIT_BF16)
    st = torch.cuda.current_stream()
    if args.tiles:
        if args.basis is not None:
            plan.set_basis(bool(args.basis))
        plan.set_tiles(*(int(v) for v in args.tiles.split(",")))
    else:
        log(plan.tune(trials=5).strip())
    nunits = len(w.units)
    if args.only_grad_batch:
        plan.set_vjp_basis(bool(args.vjp_basis))
        print(json.dumps({"grad_pass_batch": grad_batch_pass(plan, w, outs, args.steps),
                          "vjp_basis": plan.vjp_basis(), "tiles": plan.tiles()}), flush=True)
        plan.close()
        return
    if args.stages_only:
        st_ms = plan.stage_times(iters=args.steps)
        print(json.dumps({"stage_ms": st_ms, "tiles": plan.tiles(), "describe": plan.describe()}), flush=True)
        plan.close()
        return

    # warm-up, then K graph replays bracketed by synchronisation
    for _ in range(args.warmup):
        plan.forward_graph()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plan.forward_graph()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms_step = 1e3 * dt / args.steps
    sps = nunits * args.steps / dt
    log(f"split-bf16 batched: {ms_step:.3f} ms/step, {sps:.0f} slices/s")

    # per-stage kernel durations with HIP events on the launch stream
    stage_ms = plan.stage_times(iters=max(3, args.steps // 2))
    log(f"stage times (ms): {stage_ms}")
    # HBM traffic of the dominant kernel (the y-stage of the configuration just timed) from counter
    # passes in child processes, while this process waits
    basis_on = plan.basis()[0]
    y_kernel = "ystage_basis" if basis_on else "ystage_mfma"
    traffic = None
    if not args.no_traffic:
        traffic = measure_traffic(",".join(str(v) for v in plan.tiles()), basis_on, y_kernel)
        log(f"y-stage traffic per dispatch: {traffic}")

    # parity of sampled units against the FP64 oracle
    errs = []
    if args.check_units > 0:
        from oracle import oracle as O
        idx = np.linspace(0, nunits - 1, args.check_units).astype(int)
        for u in idx:
            want = O.forward_workload_unit(w, int(u))
            got = outs[int(u)].cpu().numpy()
            errs.append(float(np.linalg.norm(got - want) / np.linalg.norm(want)))
        log(f"rel-L2 vs FP64 oracle on units {list(idx)}: {errs}")

    # the plain FP32 per-slice path (one launch pair per unit) on the same batch
    plain = None
    errs_plain = []
    if not args.no_plain:
        pplan, pouts, pkeep = workloads.to_plan(w, abi.PREC_FP32)
        pplan.forward()
        torch.cuda.synchronize()
        ksteps = max(1, args.steps // 4)
        t0 = time.perf_counter()
        for _ in range(ksteps):
            pplan.forward()
        torch.cuda.synchronize()
        pdt = time.perf_counter() - t0
        plain = nunits * ksteps / pdt
        log(f"plain fp32 per-slice: {1e3 * pdt / ksteps:.3f} ms/step, {plain:.0f} slices/s")
        if args.check_units > 0:
            from oracle import oracle as O
            for u in np.linspace(0, nunits - 1, args.check_units).astype(int):
                want = O.forward_workload_unit(w, int(u))
                errs_plain.append(float(np.linalg.norm(pouts[int(u)].cpu().numpy() - want) / np.linalg.norm(want)))
            log(f"plain fp32 rel-L2 vs FP64 oracle: {errs_plain}")
        pplan.close()

    # reconstruction of the batch's k-spaces (window, inverse 2D DFT per coil, Roemer combine)
    from mrsim_amd import gpu as G
    rec = G.Recon(w.nx, w.ny, w.n_coils, w.acq)
    imgs = [torch.empty((w.ny, w.nx), dtype=torch.complex64, device="cuda") for _ in outs]
    rec.run(outs, imgs)
    torch.cuda.synchronize()
    nrec = max(3, args.steps // 2)
    t0 = time.perf_counter()
    for _ in range(nrec):
        rec.launch()
    torch.cuda.synchronize()
    rdt = (time.perf_counter() - t0) / nrec
    rbytes = nunits * (w.n_coils + 1) * w.nx * w.ny * 8
    # the separable inverse DFT of inverse_2d: per coil ny*nx*nx + nx*ny*ny complex MACs, 8 flops each
    rflops = nunits * w.n_coils * 8.0 * (w.ny * w.nx * w.nx + w.nx * w.ny * w.ny)
    recon = {"workload": f"{nunits} units, {w.n_coils} coils, {w.nx}x{w.ny}: inverse 2D DFT per coil + Roemer combine",
             "ms_per_batch": round(1e3 * rdt, 4), "slices_per_s": round(nunits / rdt, 1),
             "hbm_gbs": round(rbytes / rdt / 1e9, 1), "hbm_frac": round(rbytes / rdt / 1e9 / PEAK_HBM_GBS, 4),
             "dft_tflops": round(rflops / rdt / 1e12, 2), "dft_frac_of_fp32_peak": round(rflops / rdt / 1e12 / PEAK_FP32_TF, 4),
             "note": "the DFT runs as split-BF16 products (3 per complex product) on the bf16 MFMA with FP32 "
                     "accumulation; its FP32-accurate flops are priced at the published FP32 rate"}
    rec.close()
    log(f"reconstruction: {recon}")

    grad = None if args.no_grad else grad_pass(max(5, args.steps))
    grad_batch = None if args.no_grad else grad_batch_pass(plan, w, outs, max(5, args.steps // 2))
    if grad_batch:
        log(f"batched gradient pass: {grad_batch}")
    nufft = None if args.no_nufft else nufft_pass(max(5, args.steps))
    if nufft:
        log(f"NUFFT comparator: {nufft}")
    if grad:
        log(f"gradient pass: {grad}")
    ge3d = None if args.no_ge3d else ge3d_pass(max(3, args.steps // 4), check=args.check_units > 0,
                                               tiles=args.ge3d_tiles)
    if ge3d:
        log(f"3D gradient-echo train: {ge3d}")

    c2 = None if args.no_config2 else config2_pass(max(3, args.steps // 4), check=args.check_units > 0)
    if c2:
        log(f"config 2: {c2}")

    stats = plan.stats()
    terms = stats["terms"]
    lines = stats["lines"]
    P = w.snx * w.sny
    Q, C = w.n_comp, w.n_coils
    Ly = lines / nunits
    # The dominant kernel is the y-stage, launched once per unit chunk; stage_ms["ystage"] times the
    # chunks' launches back to back with HIP events on their stream, so a launch averages
    # y_ms / nchunks over nunits / nchunks units.
    import re
    nchunks = int(re.search(r"(\d+) chunks", plan.describe()).group(1))
    y_ms = stage_ms["ystage"]
    t_launch = y_ms / nchunks * 1e-3
    upl = nunits / nchunks
    # algorithmic work per unit: F_y = 8 C P Ly real flops of the FP32-accurate complex y-contraction
    f_y = 8.0 * C * P * Ly
    achieved = f_y * upl / t_launch / 1e12
    # algorithmic HBM bytes per unit: images in, G out, plus each slice's static planes and coil maps
    alg_bytes = (Q * P * 4 + C * w.snx * Ly * 8) + len(w.slices) / nunits * P * (4 + 8 + 8 * Q + 8 * C)
    traffic_bytes = None
    if traffic:
        traffic_bytes = round(traffic["fetch_bytes"] + traffic["write_bytes"])
    kq = Q if basis_on else 1   # MFMA K per y: the compartments in the slice basis, E per unit
    x_ms = stage_ms["xstage"]
    x_bytes = nunits * C * Ly * (w.snx + w.nx) * 8
    roofline = {
        "bound": "mfma",
        "achieved": round(achieved, 2),
        "peak": PEAK_FP32_TF,
        "unit": "TFLOP/s",
        "frac": round(achieved / PEAK_FP32_TF, 4),
        "traffic": traffic_bytes,
        "kernel": y_kernel,
        "kernel_ms_per_launch": round(y_ms / nchunks, 4),
        "launches_per_step": nchunks,
        "per_unit": f"F_y = 8*C*P*Ly = 8*{C}*{P}*{Ly:.0f} flops per slice unit, {upl:.0f} units per launch; "
                    f"peak: the published FP32 rate (the contraction is FP32-accurate, run as split-BF16 "
                    f"products on the bf16 MFMA)",
        "traffic_detail": traffic,
        "views": {
            "hbm": {"achieved": round(alg_bytes * upl / t_launch / 1e9, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                    "frac": round(alg_bytes * upl / t_launch / 1e9 / PEAK_HBM_GBS, 4),
                    "algorithmic_bytes_per_launch": round(alg_bytes * upl)},
            "bf16_mfma_executed": {
                "achieved": round(3 * f_y * kq * upl / t_launch / 1e12, 1), "peak": PEAK_BF16_MFMA_TF,
                "unit": "TFLOP/s", "frac": round(3 * f_y * kq * upl / t_launch / 1e12 / PEAK_BF16_MFMA_TF, 4),
                "note": ("three split-BF16 products of the slice-basis contraction (K = compartments x y): 3*Q*F_y"
                         if basis_on else
                         "three split-BF16 products of the per-unit contraction (K = y, compartments summed "
                         "into E on the VALU): 3*F_y")},
            "literal_transcendentals": {"achieved": round(Ly * P * (Q + 2) * upl / t_launch / 1e12, 3),
                                        "unit": "T/s", "note": "Q exp + sin + cos per voxel and line; no peak claimed"},
            "xstage": {"ms_per_step": round(x_ms, 4),
                       "hbm_gbs": round(x_bytes / (x_ms * 1e-3) / 1e9, 1),
                       "hbm_frac": round(x_bytes / (x_ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
                       "bf16_mfma_tflops": round(3 * 8.0 * C * lines * w.nx * w.snx / (x_ms * 1e-3) / 1e12, 2)},
        },
    }
    cpu = None if args.no_cpu else cpu_baseline()
    if cpu:
        log(f"cpu baseline: {cpu}")
    out = {
        "metric": "slices/s (128^2->64^2, 8 coils) batched split-BF16 forward",
        "value": round(sps, 1),
        "unit": "slices/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "split-bf16 mfma / fp32 acc / fp64 phase",
        "data": "synthetic",
        "config": {"workload": "batch_128: 30 slices x 20 volumes x 8 coils x 3 compartments, 128^2->64^2, "
                               "T2/T2' maps, one hipGraph per step",
                   "units_per_step": nunits, "terms_per_step": terms},
        "terms_per_s": round(terms * args.steps / dt, 1),
        "plain_fp32_slices_per_s": None if plain is None else round(plain, 1),
        "batched_over_plain": None if plain is None else round(sps / plain, 2),
        "rel_l2_vs_fp64": errs,
        "plain_fp32_rel_l2_vs_fp64": errs_plain,
        "stage_ms": stage_ms,
        "tiles": plan.tiles(),
        "basis": plan.basis()[0],
        "grad_pass": grad,
        "grad_pass_batch": grad_batch,
        "nufft_comparator": nufft,
        "reconstruction": recon,
        "ge3d": ge3d,
        "config2": c2,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    plan.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
