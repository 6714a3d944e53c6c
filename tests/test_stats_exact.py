"""The product's parallel image statistics (deeds_image_stats) against the
oracle's sequential loops (primsMST.h:101-110 restated), bit for bit, on the
CPU: the chunked ulp-unit emulation must reproduce every rounding of the
sequential float accumulation, including exact halfway cases, binade
crossings, negative running sums, sign changes and tiny inputs."""
import ctypes
import time

import numpy as np
import pytest

from deedsbcv_amd import _lib


def _stats(im):
    im = np.ascontiguousarray(im, np.float32)
    mean = ctypes.c_float()
    sd = ctypes.c_float()
    _lib.call("deeds_image_stats", im.ctypes.data_as(ctypes.c_void_p), ctypes.c_longlong(im.size),
              ctypes.byref(mean), ctypes.byref(sd))
    return mean.value, sd.value


def _cases():
    rng = np.random.default_rng(7)
    yield "ct_integers_8M", rng.integers(0, 4096, 256 * 256 * 128).astype(np.float32)
    yield "ct_like", (np.clip(rng.normal(1000, 300, 3_000_000), -200, 4000)).round().astype(np.float32)
    yield "fractional", (rng.random(2_000_000) * 3000).astype(np.float32)
    yield "signed", rng.normal(0, 1000, 1_000_000).astype(np.float32)
    yield "negative_mean", rng.normal(-700, 50, 500_000).astype(np.float32)
    # raw HU (no +1024): mostly air, so the running sum is negative throughout
    yield "ct_raw_hu", np.where(rng.random(3_000_000) < 0.6, -1000,
                                rng.integers(-200, 300, 3_000_000)).astype(np.float32)
    yield "negative_fractional", (-rng.random(1_500_000) * 2500 - 3.25).astype(np.float32)
    yield "negative_halves", (-(rng.integers(0, 8, 1_000_000) * 512 + 512)).astype(np.float32)
    yield "sign_change", np.concatenate([rng.normal(-900, 30, 400_000), rng.normal(1500, 30, 600_000)]).astype(
        np.float32)
    yield "halves_512", (rng.integers(0, 8, 2_000_000) * 512 + 512).astype(np.float32)
    yield "powers_of_two", np.ldexp(1.0, rng.integers(-4, 14, 1_000_000)).astype(np.float32)
    yield "constant", np.full(1_000_003, 1024.0, np.float32)
    yield "large", (rng.random(800_000) * 1e7).astype(np.float32)
    yield "tiny", (rng.random(400_000) * 1e-3).astype(np.float32)
    yield "mixed_scale", np.concatenate([rng.random(300_000) * 1e-2, rng.random(300_000) * 1e5,
                                         -rng.random(300_000) * 1e3]).astype(np.float32)
    for n in (1, 2, 7, 8191, 8192, 8193, 65537):
        yield f"n{n}", rng.integers(0, 3000, n).astype(np.float32)


@pytest.mark.parametrize("name,im", list(_cases()), ids=[c[0] for c in _cases()])
def test_stats_bit_identical_to_sequential(oracle, name, im):
    t0 = time.perf_counter()
    got = _stats(im)
    t1 = time.perf_counter()
    want = oracle.image_stats(im)
    t2 = time.perf_counter()
    assert np.float32(got[0]).view(np.uint32) == np.float32(want[0]).view(np.uint32), (got, want)
    assert np.float32(got[1]).view(np.uint32) == np.float32(want[1]).view(np.uint32), (got, want)
    print(f"{name}: parallel {1e3 * (t1 - t0):.1f} ms, sequential {1e3 * (t2 - t1):.1f} ms")


def test_stats_of_phantom_images(oracle):
    from deedsbcv_amd import phantom
    f, m = phantom.make_pair((96, 96, 80), 3)
    for im in (f.astype(np.float32) + 1024, m.astype(np.float32) + 1024, f.astype(np.float32)):
        got = _stats(im)
        want = oracle.image_stats(im)
        assert got == want


@pytest.mark.gpu
def test_device_stats_bit_identical_to_sequential(oracle):
    """deeds_ctx_image_stats: the chunk decompositions computed on the GPU."""
    torch = pytest.importorskip("torch")
    from deedsbcv_amd import api
    ctx = api.Context(0)
    for name, im in _cases():
        print("case: device stats", name, flush=True)
        d = torch.from_numpy(np.ascontiguousarray(im, np.float32)).cuda()
        mean = ctypes.c_float()
        sd = ctypes.c_float()
        _lib.ctx_call(ctx.h, "deeds_ctx_image_stats", ctypes.c_void_p(d.data_ptr()), ctypes.c_longlong(im.size),
                      ctypes.byref(mean), ctypes.byref(sd))
        want = oracle.image_stats(im)
        assert (mean.value, sd.value) == want, (name, mean.value, sd.value, want)
