"""End-to-end parity at the benchmark sizes: the HIP registration (through the
C ABI) against the unmodified reference deedsBCV binary run on the same NIfTI
files, compared bit for bit on its two deliverables, _displacements.dat and
_deformed.nii.gz (deedsBCV0.cpp:339-356).

* 256x256x128 CT-like pair, default five-level schedule (grid 8..4, radius
  8..4 -- up to 4913 labels per node -- quantisation 5..1, alpha 1.6);
* 192x160x192 cross-modal pair (non-monotone intensity transfer, Rician-like
  noise), the same schedule;
* the 16-pair multi-atlas batch at its drawn sizes (U[128,224] x U[128,192] x
  U[96,176]): the level-synchronous batch equals one-at-a-time registrations
  of every pair (flows and deformed images), and two pairs equal the
  reference binary.

The reference runs in the background on the host cores while the GPU works."""
import concurrent.futures as cf
import multiprocessing as mp

import numpy as np
import pytest

import refrun
from deedsbcv_amd import api, phantom

pytestmark = pytest.mark.gpu

needs_ref = pytest.mark.skipif(not refrun.available(), reason="reference deedsBCV not built (oracle/_ref)")


def _check_against_reference(f, mv, sched, tmp_path, what):
    run = refrun.RefRun(f, mv, sched, tmp_path)
    flow, warped, t = api.register(f, mv, sched["grid"], sched["radius"], sched["quant"], sched["alpha"])
    print(f"{what}: GPU registration {1e3 * t['total']:.1f} ms, waiting for the reference", flush=True)
    rflow, rwarped, _, out = run.result()
    refrun.assert_same(flow, rflow, f"{what}: flow vs reference _displacements.dat")
    refrun.assert_same(warped, rwarped, f"{what}: deformed image vs reference _deformed.nii.gz")
    return t


@needs_ref
def test_full_256x256x128_five_levels_equals_reference(tmp_path):
    print("case: 256x256x128 five-level registration vs reference binary", flush=True)
    f, mv = phantom.workload_full(1)
    t = _check_against_reference(f, mv, phantom.SCHEDULE_FULL, tmp_path, "256x256x128")
    assert t["hamming_evals"] == pytest.approx(67.052e9, rel=1e-3)


@needs_ref
def test_crossmodal_192x160x192_equals_reference(tmp_path):
    print("case: 192x160x192 cross-modal registration vs reference binary", flush=True)
    f, mv = phantom.workload_crossmodal(2)
    _check_against_reference(f, mv, phantom.SCHEDULE_FULL, tmp_path, "192x160x192")


def test_atlas_batch_16_pairs_equals_sequential_and_reference(tmp_path):
    print("case: 16-pair multi-atlas batch at drawn sizes", flush=True)
    dims = phantom.atlas_dims(100, 16)
    # workers forked from a fresh server process, not from this one (its GPU runtime has threads)
    with cf.ProcessPoolExecutor(max_workers=refrun.host_threads(), mp_context=mp.get_context("forkserver")) as ex:
        pairs = list(ex.map(phantom.workload_crossmodal, [100 + i for i in range(16)], dims))
    s = phantom.SCHEDULE_FULL
    # the two smallest pairs also go to the reference binary, in the background
    small = sorted(range(16), key=lambda i: dims[i][0] * dims[i][1] * dims[i][2])[:2]
    threads = max(1, refrun.host_threads() // 2)
    runs = {i: refrun.RefRun(*pairs[i], s, tmp_path / f"p{i}", threads=threads) for i in small} \
        if refrun.available() else {}
    kw = dict(grid=s["grid"], radius=s["radius"], quant=s["quant"], alpha=s["alpha"])
    got = api.register_batch(pairs, concurrency=16, **kw)
    print("atlas: batch done, now one at a time", flush=True)
    for i, ((f, mv), (flow, warped, t)) in enumerate(zip(pairs, got)):
        f1, w1, _ = api.register(f, mv, **kw)
        refrun.assert_same(flow, f1, f"pair {i} {dims[i]}: batch flow vs alone")
        refrun.assert_same(warped, w1, f"pair {i} {dims[i]}: batch deformed image vs alone")
        assert t["hamming_evals"] > 0
    for i, run in runs.items():
        rflow, rwarped, _, _ = run.result()
        refrun.assert_same(got[i][0], rflow, f"pair {i} {dims[i]}: flow vs reference")
        refrun.assert_same(got[i][1], rwarped, f"pair {i} {dims[i]}: deformed image vs reference")
