#!/usr/bin/env python3
"""Benchmark: symmetric 5-level deeds registrations per second on one MI355X.

A step is one full symmetric registration (deedsBCV0.cpp:183-356: five levels,
grid 8..4, radius 8..4, quantisation 5..1, alpha 1.6, forward + backward with
consistent mapping, final warp of the moving image) of the seeded 256x256x128
CT-like pair, with both volumes already resident in HBM when the timed region
starts.  Prints one JSON line: the headline figure, its data-cost roofline,
the reference CPU baseline and whether the GPU output equals the reference's,
then under "per_config" the other workloads (96x96x80 three-level pair,
192x160x192 cross-modal pair, 16-pair multi-atlas batch, soft level) and
under "kernel_rooflines" the HBM-bound kernels (MIND, warps).

    python bench.py [--steps K] [--warmup W] [--workload full|small|crossmodal|atlas|soft|linear|dice]
                    [--no-per-config] [--no-cpu-baseline]
"""
from __future__ import annotations

import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

VALU_INT_PEAK_TOPS = 39.3  # 256 CU x 64 lanes x 2.4 GHz, 32-bit integer VALU ops
OPS_PER_HAMMING = 4        # 2 x v_xor_b32 + 2 x v_bcnt_u32_b32 per 64-bit code pair
HBM_PEAK_GBS = 8000.0      # MI355X HBM3E
TRAFFIC_JSON = ROOT / "profiles" / "registration_256x256x128_hbm_traffic.json"
REF_BIN = ROOT / "oracle" / "_ref"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_threads():
    return max(1, min(16, len(os.sched_getaffinity(0))))


def cpu_model():
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


# --------------------------------------------------------------- profiles ---

def _profile_kernels():
    try:
        return json.loads(TRAFFIC_JSON.read_text())["kernels"]
    except (OSError, ValueError, KeyError):
        return {}


def profiled_traffic(prefixes, per_unit):
    """HBM bytes (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated) per unit of
    work from the committed registration profile: the bytes of every kernel whose
    name starts with one of `prefixes`, over the dispatches of kernel `per_unit`
    (the one that runs once per unit).  None when the profile lacks them."""
    k = _profile_kernels()
    rows = {n: v for n, v in k.items() if any(re.sub(r"^void ", "", n).startswith(p) for p in prefixes)}
    units = sum(v["dispatches"] for n, v in rows.items() if re.sub(r"^void ", "", n).startswith(per_unit))
    if not rows or not units:
        return None
    return sum((v["fetch_bytes"] + v["write_bytes"]) * v["dispatches"] for v in rows.values()) / units


def datacost_traffic():
    return profiled_traffic(["k_datacost"], "k_datacost")


# ------------------------------------------------------------ workloads ---

def workload(name):
    from deedsbcv_amd import phantom as ph
    if name == "full":
        f, m = ph.workload_full(1)
        sched = ph.SCHEDULE_FULL
    elif name == "small":
        f, m = ph.workload_small(0)
        sched = ph.SCHEDULE_SMALL
    elif name == "crossmodal":
        f, m = ph.workload_crossmodal(2)
        sched = ph.SCHEDULE_FULL
    else:
        raise SystemExit(f"unknown workload {name}")
    return f, m, sched


def _timed(fn, steps, warmup):
    """Mean wall seconds of fn() over `steps` calls after `warmup` calls, GPU drained on both sides."""
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


# ------------------------------------------------------- reference (CPU) ---

def _deeds_cmd(exe, d, tag, sched):
    return [str(exe), "-F", f"{d}/f.nii.gz", "-M", f"{d}/m.nii.gz", "-O", f"{d}/{tag}",
            "-G", "x".join(map(str, sched["grid"])), "-L", "x".join(map(str, sched["radius"])),
            "-Q", "x".join(map(str, sched["quant"])), "-a", str(sched["alpha"])]


def reference_run(f, m, sched, threads, budget_s, cli=False):
    """The unmodified reference deedsBCV (oracle/_ref, compiled from the reference
    sources) on the same pair with `threads` OpenMP threads: its 'Total time' span,
    its per-level data-cost seconds, its process wall time and its
    _displacements.dat / _deformed.nii.gz.  With `cli`, this library's deedsBCV
    command line (deedsbcv_amd/bin) then runs on the same files: its wall time
    and whether its output files decompress to the reference's bytes.  None
    where the reference build is absent."""
    ref = REF_BIN / "deedsBCV"
    if not ref.exists():
        return None
    from deedsbcv_amd import nifti
    with tempfile.TemporaryDirectory() as d:
        nifti.write_nifti(f"{d}/f.nii.gz", f)
        nifti.write_nifti(f"{d}/m.nii.gz", m)
        t0 = time.perf_counter()
        r = subprocess.run(_deeds_cmd(ref, d, "o", sched), capture_output=True, text=True,
                           env=dict(os.environ, OMP_NUM_THREADS=str(threads)), timeout=budget_s)
        wall = time.perf_counter() - t0
        mt = re.search(r"Total time: ([0-9.eE+-]+) sec", r.stdout)
        if r.returncode != 0 or not mt:
            return None
        out = {"total_s": float(mt.group(1)), "wall_s": wall,
               "datacost_s": sum(float(x) for x in re.findall(r"data=([0-9.eE+-]+)", r.stdout)),
               "flow": nifti.read_displacements(f"{d}/o_displacements.dat"),
               "warped": nifti.read_nifti(f"{d}/o_deformed.nii.gz")[0]}
        exe = ROOT / "deedsbcv_amd" / "bin" / "deedsBCV"
        if cli and exe.exists():
            import gzip
            t0 = time.perf_counter()
            c = subprocess.run(_deeds_cmd(exe, d, "g", sched), capture_output=True, text=True, timeout=budget_s)
            cwall = time.perf_counter() - t0
            if c.returncode == 0:
                same = (Path(f"{d}/g_displacements.dat").read_bytes() == Path(f"{d}/o_displacements.dat").read_bytes()
                        and gzip.open(f"{d}/g_deformed.nii.gz").read() == gzip.open(f"{d}/o_deformed.nii.gz").read())
                out["cli"] = {"deeds_hip_wall_s": cwall, "reference_wall_s": wall, "speedup": wall / cwall,
                              "identical_files": same,
                              "note": "process wall time of the deedsBCV command line on the same .nii.gz pair: "
                                      "start-up, NIfTI read, registration, .dat and deformed .nii.gz written "
                                      "(decompressed bytes compared)"}
        return out


def _same(a, b):
    return bool(a is not None and b is not None and np.asarray(a).shape == np.asarray(b).shape
                and np.array_equal(a, b))


def cpu_baseline_entry(runs, threads, evals, sample):
    """cpu_baseline object from reference runs (list of reference_run results)."""
    if not runs or any(r is None for r in runs):
        return None
    total = sum(r["total_s"] for r in runs)
    data = sum(r["datacost_s"] for r in runs)
    out = {"value": len(runs) / total, "unit": "registrations/s", "cores": threads, "kind": "reference",
           "sample": sample, "cpu_model": cpu_model(), "total_s": total, "datacost_s": data}
    if evals and data > 0:
        out["hamming_gevals_per_s"] = evals / data / 1e9
    return out


# ------------------------------------------------------ plain registration ---

def plain_bench(a, name, steps, warmup, cpu=True):
    """One plain workload (full / small / crossmodal): registrations/s with the
    volumes in HBM, the data-cost rate and roofline, and -- with `cpu` -- the
    reference binary on the same pair, its outputs compared with the GPU's."""
    import torch

    from deedsbcv_amd import api
    log(f"generating workload {name}")
    f, m, sched = workload(name)
    o, n, mm = f.shape
    dev = torch.device("cuda:0")
    fixed = torch.from_numpy(f.astype(np.float32)).to(dev)
    moving = torch.from_numpy(m.astype(np.float32)).to(dev)
    g = sched["grid"][-1]
    sz1 = (mm // g) * (n // g) * (o // g)
    flow = torch.empty(3 * sz1, dtype=torch.float32, device=dev)
    warped = torch.empty_like(fixed)
    ctx = api.Context(0)
    kw = dict(grid=sched["grid"], radius=sched["radius"], quant=sched["quant"], alpha=sched["alpha"])
    for i in range(warmup):
        ctx.register(fixed, moving, flow, warped, **kw)
    torch.cuda.synchronize()
    timings = []
    t0 = time.perf_counter()
    for _ in range(steps):
        timings.append(ctx.register(fixed, moving, flow, warped, **kw))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    log(f"{name}: {steps} timed steps in {wall:.3f} s")
    fl = flow.cpu().numpy()
    wp = warped.cpu().numpy()
    if not np.all(np.isfinite(fl)):
        raise SystemExit("non-finite displacement field")
    evals = timings[0]["hamming_evals"]
    dc_s = sum(t["datacost_kernel_s"] for t in timings)
    dc_n = sum(t["datacost_launches"] for t in timings)
    achieved = OPS_PER_HAMMING * evals * steps / dc_s / 1e12
    mean = lambda key: [1e3 * x for x in np.mean([t[key] for t in timings], axis=0)]  # noqa: E731
    out = {
        "value": steps / wall, "unit": "registrations/s", "ms_per_step": 1e3 * wall / steps, "steps": steps,
        "warmup": warmup,
        "workload": f"{name}: {mm}x{n}x{o} int16 pair, grid {sched['grid']}, radius {sched['radius']}, "
                    f"quant {sched['quant']}, alpha {sched['alpha']}, symmetric",
        "dims": [mm, n, o], "levels": len(sched["grid"]),
        "hamming_gevals_per_s": evals * steps / wall / 1e9,
        "datacost_kernel_gevals_per_s": evals * steps / dc_s / 1e9,
        "hamming_evals_per_registration": evals,
        "stage_ms_fwd_stream": {k: 1e3 * float(np.mean([sum(t[k]) for t in timings]))
                                for k in ("mind", "data", "mst_reg", "transf")},
        "host_tree_build_ms": 1e3 * float(np.mean([t["tree_build_s"] for t in timings])),
        "host_tree_wait_ms": 1e3 * float(np.mean([t["tree_wait_s"] for t in timings])),
        "host_tree_wait_level_ms": mean("tree_wait_level_s"),
        "host_tree_build_level_ms": mean("tree_build_level_s"),
        "host_stats_ms": 1e3 * float(np.mean([t["stats_s"] for t in timings])),
        "gpu_idle_before_bp_level_ms": mean("bp_idle_level_s"),
        "bp_level_ms_fwd_stream": mean("bp_level_s"),
        "host_tree_ready_level_ms": mean("tree_ready_level_s"),
        "host_tree_prim_level_ms": mean("tree_shape_level_s"),
        "datacost_level_ms": mean("datacost_level_s"),
        "roofline": {
            "kernel": "k_datacost (Hamming cost volume)", "bound": "valu", "achieved": achieved,
            "peak": VALU_INT_PEAK_TOPS, "unit": "TOP/s (int32 VALU)", "frac": achieved / VALU_INT_PEAK_TOPS,
            "traffic": datacost_traffic() if name == "full" else None,
            "traffic_unit": "HBM bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated)",
            "launches_timed": dc_n, "avg_launch_ms": 1e3 * dc_s / dc_n,
        },
    }
    out["_timings"] = timings
    out["cpu_baseline"] = None
    if cpu:
        threads = a.cpu_threads or host_threads()
        log(f"{name}: reference deedsBCV with {threads} OpenMP threads")
        r = reference_run(f, m, sched, threads, a.cpu_budget, cli=name == "full")
        if r is not None and "cli" in r:
            out["cli_end_to_end"] = r["cli"]
        out["cpu_baseline"] = cpu_baseline_entry(
            [r], threads, evals, "one full symmetric registration of the same pair by the unmodified reference "
                                 "(its 'Total time' span, deedsBCV0.cpp:217-336)")
        if r is not None:
            out["identical_to_reference"] = {"flow": _same(fl, r["flow"]), "deformed": _same(wp, r["warped"]),
                                             "compared": "_displacements.dat and _deformed.nii.gz, bit for bit"}
    return out


# ------------------------------------------------------------ multi-atlas ---

def atlas_bench(a, steps, cpu=True):
    """Multi-atlas: 16 pairs of drawn sizes, the level-synchronous batch against
    one-at-a-time registrations of the same pairs; the reference binary on the
    first two pairs as the CPU sample."""
    import concurrent.futures as cf
    import multiprocessing as mp

    import torch

    from deedsbcv_amd import api
    from deedsbcv_amd import phantom as ph
    sched = ph.SCHEDULE_FULL
    dims = ph.atlas_dims(100, 16)
    log("generating 16 atlas pairs")
    if a.serial_data:
        vols = [ph.workload_crossmodal(100 + i, d) for i, d in enumerate(dims)]
    else:
        # workers forked from a fresh server process, not from this one (its GPU runtime has threads)
        with cf.ProcessPoolExecutor(max_workers=host_threads(), mp_context=mp.get_context("forkserver")) as ex:
            vols = list(ex.map(ph.workload_crossmodal, [100 + i for i in range(16)], dims))
    g = sched["grid"][-1]
    pairs = []
    for (f, m), (mm, n, o) in zip(vols, dims):
        pairs.append(dict(fixed=torch.from_numpy(f.astype(np.float32)).cuda(),
                          moving=torch.from_numpy(m.astype(np.float32)).cuda(),
                          flow=torch.empty(3 * (mm // g) * (n // g) * (o // g), device="cuda"),
                          warped=torch.empty(f.shape, device="cuda")))
    kw = dict(grid=sched["grid"], radius=sched["radius"], quant=sched["quant"], alpha=sched["alpha"])
    from deedsbcv_amd import _lib
    lib = _lib.load()

    def run_batches(graphs):
        # the first call runs eagerly, the second captures the level graphs: both untimed
        lib.deeds_set_batch_graphs(1 if graphs else 0)
        for _ in range(2):
            api.register_batch_device(pairs, concurrency=a.concurrency, **kw)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        times = None
        for _ in range(steps):
            times = api.register_batch_device(pairs, concurrency=a.concurrency, **kw)
        torch.cuda.synchronize()
        return time.perf_counter() - t0, times, [(p["flow"].cpu().numpy(), p["warped"].cpu().numpy()) for p in pairs]

    wall_eager, _, batch_eager = run_batches(False)
    wall, times, batch = run_batches(True)
    log(f"atlas: batch {1e3 * wall / steps:.1f} ms with level graphs, {1e3 * wall_eager / steps:.1f} ms without")
    ctx = api.Context(0)
    ctx.register(pairs[0]["fixed"], pairs[0]["moving"], pairs[0]["flow"], pairs[0]["warped"], **kw)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for p in pairs:
        ctx.register(p["fixed"], p["moving"], p["flow"], p["warped"], **kw)
    torch.cuda.synchronize()
    seq = time.perf_counter() - t1
    identical = all(_same(p["flow"].cpu().numpy(), b[0]) and _same(p["warped"].cpu().numpy(), b[1])
                    and _same(b[0], e[0]) and _same(b[1], e[1]) for p, b, e in zip(pairs, batch, batch_eager))
    n = len(pairs)
    evals = sum(t["hamming_evals"] for t in times)
    out = {"value": n * steps / wall, "unit": "registrations/s", "ms_per_step": 1e3 * wall / steps, "steps": steps,
           "workload": f"atlas: 16 pairs, dims U[128,224]xU[128,192]xU[96,176], grid {sched['grid']}, "
                       f"radius {sched['radius']}, quant {sched['quant']}, level-synchronous groups of "
                       f"{a.concurrency}", "dims": dims,
           "hamming_evals_per_batch": evals,
           "sequential_registrations_per_s": n / seq, "batch_speedup_vs_sequential": (n * steps / wall) / (n / seq),
           "batch_without_graphs_registrations_per_s": n * steps / wall_eager,
           "graph_speedup": wall_eager / wall,
           "identical_to_sequential": identical,
           "compared": "flows and deformed images of every pair, batch (with and without level graphs) vs one "
                       "at a time, bit for bit",
           "autotuner": api.tuning_report()}
    if cpu:
        threads = a.cpu_threads or host_threads()
        log(f"atlas: reference deedsBCV on pairs 0 and 1 with {threads} threads")
        runs = [reference_run(*vols[i], sched, threads, a.cpu_budget) for i in (0, 1)]
        out["cpu_baseline"] = cpu_baseline_entry(
            runs, threads, sum(times[i]["hamming_evals"] for i in (0, 1)),
            "pairs 0 and 1 of the 16 (202x182x106 and 185x133x119), one after the other, by the unmodified "
            "reference (its 'Total time' spans)")
        if all(r is not None for r in runs):
            out["identical_to_reference"] = {
                "pairs": [0, 1],
                "flow": all(_same(batch[i][0], runs[i]["flow"]) for i in (0, 1)),
                "deformed": all(_same(batch[i][1], runs[i]["warped"]) for i in (0, 1))}
    return out


# ---------------------------------------------------------- soft variant ---

def soft_bench(a, steps, warmup):
    """Differentiable soft level (160^3 float32 pair, grid 6, radius 6, quantisation 2,
    T = 0.5): forward + backward of the expected displacements with respect to both
    images, against the plain GPU forward of the same level (MIND of both images,
    Hamming cost, MST, min-sum BP).  Both build the MST of the device image on a
    host thread while their features and cost volume run on the GPU."""
    import concurrent.futures as cf
    import ctypes

    import torch

    from deedsbcv_amd import _lib, api, soft
    from deedsbcv_amd import phantom as ph
    log("generating the 160^3 soft pair")
    f, m = ph.workload_soft(4)
    o, n, mm = f.shape
    grid, hw, quant, alpha, T = 6, 6, 2, 1.6, 0.5
    qs = soft.mind_step(quant)
    N = (mm // grid) * (n // grid) * (o // grid)
    L3 = (2 * hw + 1) ** 3
    dev = torch.device("cuda:0")
    x = torch.from_numpy(f).to(dev)
    y = torch.from_numpy(m).to(dev)
    wgt = torch.randn((3, o // grid, n // grid, mm // grid), generator=torch.Generator().manual_seed(0)).to(dev)
    V = ctypes.c_void_p

    # plain forward of the same level on one context (device pointers)
    ctx = api.Context(0)
    mF = torch.empty(f.shape, dtype=torch.int64, device=dev)
    mM = torch.empty(f.shape, dtype=torch.int64, device=dev)
    cost = torch.empty((N, L3), dtype=torch.float32, device=dev)
    u0 = torch.zeros((3, N), dtype=torch.float32, device=dev)
    u1 = torch.empty((3, N), dtype=torch.float32, device=dev)
    pool = cf.ThreadPoolExecutor(max_workers=1)
    mst_ctx = api.Context(0)
    torch.cuda.synchronize()

    def _p32(arr):
        return arr.ctypes.data_as(V)

    def plain():
        # the MST (from the device image) is built on a host thread while MIND and the cost volume run,
        # as the registration driver does
        pending = pool.submit(mst_ctx.prims_graph, x, grid)
        _lib.ctx_call(ctx.h, "deeds_ctx_descriptor", V(mF.data_ptr()), V(x.data_ptr()), mm, n, o, qs)
        _lib.ctx_call(ctx.h, "deeds_ctx_descriptor", V(mM.data_ptr()), V(y.data_ptr()), mm, n, o, qs)
        _lib.ctx_call(ctx.h, "deeds_ctx_datacost", V(mF.data_ptr()), V(mM.data_ptr()), V(cost.data_ptr()), mm, n, o,
                      grid, hw, float(quant), alpha, 1)
        ordered, parents, edgemst = pending.result()
        _lib.ctx_call(ctx.h, "deeds_ctx_regularisation", V(cost.data_ptr()), V(u0[0].data_ptr()),
                      V(u0[1].data_ptr()), V(u0[2].data_ptr()), V(u1[0].data_ptr()), V(u1[1].data_ptr()),
                      V(u1[2].data_ptr()), hw, grid, float(quant), _p32(ordered), _p32(parents), _p32(edgemst),
                      mm, n, o, None)
        _lib.ctx_call(ctx.h, "deeds_ctx_synchronize")

    xs = x.clone().requires_grad_(True)
    ys = y.clone().requires_grad_(True)

    def soft_fb():
        xs.grad = ys.grad = None
        loss = (soft.soft_register(xs, ys, grid, hw, quant, alpha, T) * wgt).sum()
        loss.backward()

    def soft_f():
        with torch.no_grad():
            soft.soft_register(x, y, grid, hw, quant, alpha, T)

    t_plain = _timed(plain, steps, max(1, warmup))
    t_sf = _timed(soft_f, steps, max(1, warmup))
    t_sfb = _timed(soft_fb, steps, max(1, warmup))
    log(f"soft: plain forward {1e3 * t_plain:.2f} ms, soft forward {1e3 * t_sf:.2f} ms, "
        f"forward+backward {1e3 * t_sfb:.2f} ms")

    # stage split of the soft step (each stage timed alone)
    tree = mst_ctx.prims_graph(x, grid)
    F = soft.soft_features(xs, qs)
    M = soft.soft_features(ys, qs)
    C = soft.soft_cost(F, M, grid, hw, quant, alpha)
    U = soft.soft_bp(C, torch.zeros((3, N), device=dev), tree, f.shape, grid, hw, quant, T)
    gU = wgt.reshape(3, N).contiguous()
    gC = torch.autograd.grad(U, C, gU, retain_graph=True)[0]
    gF = torch.autograd.grad(C, (F, M), gC, retain_graph=True)
    st = {
        "mst_ms": _timed(lambda: mst_ctx.prims_graph(x, grid), steps, 1),
        "features_fwd_ms": _timed(lambda: soft.soft_features(x, qs), steps, 1) * 2,
        "cost_fwd_ms": _timed(lambda: soft.soft_cost(F.detach(), M.detach(), grid, hw, quant, alpha), steps, 1),
        "bp_fwd_ms": _timed(lambda: soft.soft_bp(C.detach(), torch.zeros((3, N), device=dev), tree, f.shape, grid,
                                                 hw, quant, T), steps, 1),
        "bp_bwd_ms": _timed(lambda: torch.autograd.grad(U, C, gU, retain_graph=True), steps, 1),
        "cost_bwd_ms": _timed(lambda: torch.autograd.grad(C, (F, M), gC, retain_graph=True), steps, 1),
        "features_bwd_ms": _timed(lambda: torch.autograd.grad(F, xs, gF[0], retain_graph=True), steps, 1) * 2,
    }
    st = {k: 1e3 * v for k, v in st.items()}
    terms = N * L3 * 54 * 12  # SSD terms of the cost volume (54 samples per label at grid 6)
    pool.shutdown()
    return {"value": 1.0 / t_sfb, "unit": "forward+backward/s", "ms_per_step": 1e3 * t_sfb, "steps": steps,
            "workload": "soft: 160x160x160 float32 pair, grid 6, radius 6 (2197 labels), quant 2, alpha 1.6, "
                        "T 0.5, gradient w.r.t. both images", "nodes": N, "labels": L3,
            "soft_forward_ms": 1e3 * t_sf, "plain_forward_ms": 1e3 * t_plain,
            "soft_fwd_bwd_over_plain_forward": t_sfb / t_plain, "soft_stage_ms": st,
            "cost_fwd_gterms_per_s": terms / (st["cost_fwd_ms"] * 1e-3) / 1e9,
            "cpu_baseline": None,
            "parity": "no reference counterpart; gradient checked against float64 central differences "
                      "(h=1e-4) on the 48^3 crop in tests/test_gpu_soft.py"}


# ------------------------------------------------ HBM-bound kernel rooflines ---

def kernel_rooflines(steps):
    """MIND (descriptor(): four kernels per call) and the warps on the 256x256x128
    pair, timed with HIP events on the stream they run on, against their
    algorithmic bytes: MIND 12 B per voxel (4 B image in, 8 B codes out); the
    fused upsample + trilinear warp (applyBCVfloat, and the registration's
    k_warp_affine_grid) 8 B per voxel; warpAffine with a full-resolution field
    20 B per voxel; upsampleDeformationsCL to full resolution 12 B per voxel."""
    import torch

    from deedsbcv_amd import api
    f, m, _ = workload("full")
    o, n, mm = f.shape
    sz = mm * n * o
    dev = torch.device("cuda:0")
    side = torch.cuda.Stream()
    out = {}
    with torch.cuda.stream(side):
        ctx = api.Context(0)
        im = torch.from_numpy(f.astype(np.float32) + 1024).to(dev)
        mv = torch.from_numpy(m.astype(np.float32) + 1024).to(dev)
        codes = torch.empty(f.shape, dtype=torch.int64, device=dev)
        rng = np.random.default_rng(3)
        g = 4
        flow = torch.from_numpy((rng.normal(size=3 * (mm // g) * (n // g) * (o // g)) * 3).astype(np.float32)).to(dev)
        grid = flow.reshape(3, o // g, n // g, mm // g)
        full = ctx.upsample_deformations(grid, (o, n, mm))
        X = np.eye(4, dtype=np.float32).T.ravel().copy()

        def timed(fn, reps):
            fn()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(side)
            for _ in range(reps):
                fn()
            e1.record(side)
            e1.synchronize()
            return e0.elapsed_time(e1) / reps * 1e-3

        def entry(kernel, seconds, bytes_per_voxel, traffic, note):
            gbs = bytes_per_voxel * sz / seconds / 1e9
            return {"kernel": kernel, "bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": gbs / HBM_PEAK_GBS, "traffic": traffic, "algorithmic_bytes": bytes_per_voxel * sz,
                    "avg_launch_ms": 1e3 * seconds, "note": note}

        reps = max(5, steps)
        for qs in (1, 2, 3):
            t = timed(lambda: ctx.descriptor(im, qs, out=codes), reps)
            out[f"mind_qs{qs}"] = entry(
                "MIND-SSC descriptor (all its kernels per call)", t, 12,
                profiled_traffic(["k_ssd_box", "k_box_lines", "k_box_cols", "k_mind"], "k_mind_code"),
                "traffic: mean HBM bytes per descriptor call over the registration profile (all qs)")
        t = timed(lambda: ctx.apply_float(mv, flow, g, X), reps)
        out["warp_fused_grid"] = entry("k_warp_affine_grid (upsample in registers + trilinear warp)", t, 8,
                                       profiled_traffic(["k_warp_affine_grid"], "k_warp_affine_grid"),
                                       "the registration's warp kernel; traffic from the registration profile")
        t = timed(lambda: ctx.warp_affine(mv, X, full), reps)
        out["warp_affine_fullres"] = entry("k_warp_affine (full-resolution field)", t, 20, None,
                                           "warpAffine with three full-resolution field volumes")
        t = timed(lambda: ctx.upsample_deformations(grid, (o, n, mm)), reps)
        out["upsample_fullres"] = entry("k_upsample (grid 4 to full resolution)", t, 12, None,
                                        "upsampleDeformationsCL writing three full-resolution fields")
    torch.cuda.synchronize()
    return out


# ------------------------------------------------------ linearBCV / Dice ---

def linear_bench(a):
    """linearBCV (linearBCV.cpp:144-263) on the 256x256x128 pair with its default
    schedule (grid 7,6,5,4; radius 5,4,3,2; quantisation 4,3,2,1; affine), inputs in
    HBM; the reference linearBCV binary on the same pair is the CPU baseline."""
    import ctypes

    import torch

    from deedsbcv_amd import _lib, api
    log("generating the 256x256x128 pair")
    f, m, _ = workload("full")
    o, n, mm = f.shape
    grid, radius, quant = (7, 6, 5, 4), (5, 4, 3, 2), (4, 3, 2, 1)
    dev = torch.device("cuda:0")
    x = torch.from_numpy(f.astype(np.float32)).to(dev)
    y = torch.from_numpy(m.astype(np.float32)).to(dev)
    ctx = api.Context(0)
    sched = _lib.make_schedule(grid, radius, quant, 1.0)
    X = np.zeros(16, np.float32)
    used = np.zeros(4, np.int32)
    V = ctypes.c_void_p
    torch.cuda.synchronize()

    def step():
        _lib.ctx_call(ctx.h, "deeds_ctx_linear_register", V(x.data_ptr()), V(y.data_ptr()), mm, n, o,
                      ctypes.byref(sched), 0, X.ctypes.data_as(V), used.ctypes.data_as(V))

    t = _timed(step, a.steps, max(1, a.warmup))
    cb = None
    ref = REF_BIN / "linearBCV"
    if not a.no_cpu_baseline and ref.exists():
        from deedsbcv_amd import nifti
        threads = a.cpu_threads or host_threads()
        with tempfile.TemporaryDirectory() as d:
            nifti.write_nifti(f"{d}/f.nii.gz", f)
            nifti.write_nifti(f"{d}/m.nii.gz", m)
            r = subprocess.run([str(ref), "-F", f"{d}/f.nii.gz", "-M", f"{d}/m.nii.gz", "-O", f"{d}/o"],
                               capture_output=True, text=True, env=dict(os.environ, OMP_NUM_THREADS=str(threads)),
                               timeout=a.cpu_budget)
            mt = re.search(r"Total time: ([0-9.eE+-]+) sec", r.stdout)
            if r.returncode == 0 and mt:
                total = float(mt.group(1))
                ref_X = (Path(d) / "o_matrix.txt").read_text()
                ours = "".join("  ".join("%g" % X[i + 4 * c] for c in range(4)) + "\n" for i in range(4))
                cb = {"value": 1.0 / total, "unit": "alignments/s", "cores": threads, "kind": "reference",
                      "cpu_model": cpu_model(),
                      "sample": "one affine alignment of the same pair by the unmodified reference linearBCV "
                                "(its 'Total time' span, linearBCV.cpp:187-263)",
                      "total_s": total, "same_matrix_text": ref_X == ours}
    return {
        "metric": "affine pre-alignments/s @256x256x128 (linearBCV, 4 levels)",
        "value": 1.0 / t, "unit": "alignments/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": 1e3 * t, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u64 Hamming / f32", "data": "synthetic (seeded CT-like phantom pair, B-spline warp)",
        "config": {"workload": "linear: 256x256x128 pair, grid (7, 6, 5, 4), radius (5, 4, 3, 2), "
                               "quant (4, 3, 2, 1), affine", "dims": [mm, n, o]},
        "points_used_per_level": used.tolist(), "cpu_baseline": cb}


def dice_bench(a):
    """diceMulti (diceMulti.cpp:17-79) on two 256x256x128 int16 label volumes with
    14 labels, inputs in HBM: 6 bytes per voxel (presence pass over the first
    volume, count pass over both) against the 8 TB/s HBM peak."""
    import ctypes

    import torch

    from deedsbcv_amd import _lib, api
    rng = np.random.default_rng(5)
    shape = (128, 256, 256)
    a_lab = np.zeros(shape, np.int16)
    for lab in range(1, 15):
        lo = [rng.integers(0, s // 2) for s in shape]
        a_lab[lo[0]:lo[0] + shape[0] // 3, lo[1]:lo[1] + shape[1] // 3, lo[2]:lo[2] + shape[2] // 3] = lab
    b_lab = np.roll(a_lab, 3, axis=2)
    dev = torch.device("cuda:0")
    A = torch.from_numpy(a_lab).to(dev)
    B = torch.from_numpy(b_lab).to(dev)
    ctx = api.Context(0)
    cnt = ctypes.c_int()
    V = ctypes.c_void_p
    torch.cuda.synchronize()

    def step():
        _lib.ctx_call(ctx.h, "deeds_ctx_dice", V(A.data_ptr()), V(B.data_ptr()), a_lab.size, 0, None, None,
                      ctypes.byref(cnt))

    t = _timed(step, a.steps, max(1, a.warmup))
    gbs = 6.0 * a_lab.size / t / 1e9
    return {
        "metric": "Dice evaluations/s @256x256x128 int16 (14 labels)", "value": 1.0 / t, "unit": "evaluations/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": 1e3 * t, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "int16", "data": "synthetic label volumes",
        "config": {"workload": "dice: 256x256x128 int16, 14 labels"}, "labels": cnt.value,
        "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
                     "traffic": None, "note": "whole call (two kernels, two small host round trips)"},
        "cpu_baseline": None}


# ------------------------------------------------------------------ main ---

def _line(metric, d, a, dtype, data, workload_name):
    """Contract line from a workload's result dict."""
    head = {"metric": metric, "value": d["value"], "unit": d["unit"], "n_gpus": 1, "steps": d["steps"],
            "warmup": d.get("warmup", a.warmup), "ms_per_step": d["ms_per_step"], "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": dtype, "data": data,
            "config": {"workload": d["workload"], "dims": d.get("dims"), "parallelism": "single GPU"}}
    rest = {k: v for k, v in d.items() if k not in head and k not in ("workload", "dims") and not k.startswith("_")}
    head.update(rest)
    return head


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="full",
                    choices=["full", "small", "crossmodal", "atlas", "soft", "linear", "dice"])
    ap.add_argument("--concurrency", type=int, default=16,
                    help="pairs per level-synchronous group for the multi-atlas batch")
    ap.add_argument("--no-per-config", action="store_true",
                    help="headline workload only (no other configs, no kernel rooflines)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=int(os.environ.get("OMP_NUM_THREADS", "0") or 0))
    ap.add_argument("--cpu-budget", type=float, default=240.0)
    ap.add_argument("--serial-data", action="store_true",
                    help="generate the atlas pairs in this process (under a profiler, whose signal handlers "
                         "the worker processes cannot inherit)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="after the timed steps, copy a 1 GiB device buffer once (a known-size kernel that "
                         "calibrates HBM byte counters in a profiler run)")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the registration path runs on a single device "
              "(multi-GPU sharding is a design note, not part of this build).", file=sys.stderr)
        raise SystemExit(2)
    cpu = not a.no_cpu_baseline
    sub_steps = max(1, min(a.steps, 5))
    if a.workload in ("linear", "dice"):
        print(json.dumps((linear_bench if a.workload == "linear" else dice_bench)(a)))
        return
    if a.workload == "atlas":
        d = atlas_bench(a, sub_steps, cpu)
        print(json.dumps(_line("multi-atlas registrations/s (16 pairs, dims U[128,224]xU[128,192]xU[96,176], "
                               "5 levels)", d, a, "u64 Hamming / f32",
                               "synthetic (seeded cross-modal phantom pairs)", "atlas")))
        return
    if a.workload == "soft":
        d = soft_bench(a, a.steps, a.warmup)
        print(json.dumps(_line("soft level forward+backward/s @160^3 (grid 6, radius 6, quant 2, T 0.5)", d, a,
                               "f32", "synthetic (seeded cross-modal phantom pair, 160^3 float32)", "soft")))
        return

    d = plain_bench(a, a.workload, a.steps, a.warmup, cpu)
    if a.copy_calibration:
        import torch
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")  # 1 GiB
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    metric = ("symmetric 5-level registrations/s @256x256x128; dataCost G Hamming evals/s" if a.workload == "full"
              else f"registrations/s ({a.workload})")
    line = _line(metric, d, a, "u64 Hamming / f32", "synthetic (seeded CT-like phantom pair, B-spline warp)",
                 a.workload)
    line["host_cpus"] = len(os.sched_getaffinity(0))
    from deedsbcv_amd import api as _api
    line["autotuner"] = _api.tuning_report()
    if a.workload == "full" and not a.no_per_config:
        t0 = time.perf_counter()
        line["kernel_rooflines"] = kernel_rooflines(a.steps)
        per = {}
        s0 = plain_bench(a, "small", max(a.steps, 10), a.warmup, cpu)
        s0.pop("_timings", None)
        per["config0_small_96x96x80"] = s0
        s2 = plain_bench(a, "crossmodal", sub_steps, min(a.warmup, 2), cpu)
        s2.pop("_timings", None)
        per["config2_crossmodal_192x160x192"] = s2
        per["config3_atlas_16_pairs"] = atlas_bench(a, sub_steps, cpu)
        per["config4_soft_160"] = soft_bench(a, sub_steps, 1)
        line["per_config"] = per
        log(f"per-config workloads and kernel rooflines: {time.perf_counter() - t0:.1f} s")
    print(json.dumps(line))


if __name__ == "__main__":
    main()
