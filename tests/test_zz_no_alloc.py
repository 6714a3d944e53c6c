"""GPU test of the allocation-free Dice entry point (deeds_ctx_dice_noalloc,
called through the C ABI): its device outputs equal deeds_ctx_dice's on two
calls in a row, the second with a smaller input (the shared scratch array keeps
the first call's contents, which the second must clear), and sizes beyond its
limits are rejected with a status."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pair(shape, seed, k, offset):
    rng = np.random.default_rng(seed)
    a = (rng.integers(0, k, shape) + offset).astype(np.int16)
    b = a.copy()
    flip = rng.random(shape) < 0.25
    b[flip] = (rng.integers(0, k, int(flip.sum())) + offset).astype(np.int16)
    return a, b


def test_dice_noalloc_matches_allocating_form():
    torch = pytest.importorskip("torch")
    from deedsbcv_amd import _lib, api
    lib = _lib.load()
    ctx = api.Context(0)
    V = ctypes.c_void_p
    cases = [((96, 80, 72), 1, 5000, -2000),  # more labels than the LDS counters hold
             ((40, 36, 30), 2, 9, 3)]         # then a smaller volume with few labels
    for shape, seed, k, off in cases:
        print(f"case: dice noalloc {shape} labels<={k}", flush=True)
        a, b = _pair(shape, seed, k, off)
        da, db = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
        torch.cuda.synchronize()
        # the allocating form
        cnt = ctypes.c_int()
        _lib.ctx_call(ctx.h, "deeds_ctx_dice", V(da.data_ptr()), V(db.data_ptr()), a.size, 0, None, None,
                      ctypes.byref(cnt))
        want_l = np.empty(cnt.value, np.int16)
        want_d = np.empty(cnt.value, np.float32)
        _lib.ctx_call(ctx.h, "deeds_ctx_dice", V(da.data_ptr()), V(db.data_ptr()), a.size, cnt.value,
                      want_l.ctypes.data_as(V), want_d.ctypes.data_as(V), ctypes.byref(cnt))
        # the allocation-free form, outputs in device memory (the tensors are made before the call)
        cap = cnt.value + 3
        lab = torch.full((cap,), -1, dtype=torch.int16, device="cuda")
        dic = torch.full((cap,), -1.0, dtype=torch.float32, device="cuda")
        dcount = torch.full((1,), -1, dtype=torch.int32, device="cuda")
        torch.cuda.synchronize()
        rc = lib.deeds_ctx_dice_noalloc(ctx.h, V(da.data_ptr()), V(db.data_ptr()), a.size, cap, V(lab.data_ptr()),
                                        V(dic.data_ptr()), V(dcount.data_ptr()))
        assert rc == 0, lib.deeds_ctx_last_error(ctx.h)
        _lib.ctx_call(ctx.h, "deeds_ctx_synchronize")
        n = int(dcount.cpu()[0])
        assert n == cnt.value
        assert np.array_equal(lab.cpu().numpy()[:n], want_l)
        assert np.array_equal(dic.cpu().numpy()[:n].view(np.uint32), want_d.view(np.uint32))
        assert np.all(lab.cpu().numpy()[n:] == -1)  # nothing written beyond the labels
    # a capacity smaller than the label count: the first labels only
    lab2 = torch.full((2,), -1, dtype=torch.int16, device="cuda")
    dic2 = torch.full((2,), -1.0, dtype=torch.float32, device="cuda")
    rc = lib.deeds_ctx_dice_noalloc(ctx.h, V(da.data_ptr()), V(db.data_ptr()), a.size, 2, V(lab2.data_ptr()),
                                    V(dic2.data_ptr()), V(dcount.data_ptr()))
    assert rc == 0
    _lib.ctx_call(ctx.h, "deeds_ctx_synchronize")
    assert int(dcount.cpu()[0]) == cnt.value
    assert np.array_equal(lab2.cpu().numpy(), want_l[:2])
    assert np.array_equal(dic2.cpu().numpy().view(np.uint32), want_d[:2].view(np.uint32))


def test_dice_noalloc_rejects_sizes_beyond_its_limits():
    torch = pytest.importorskip("torch")
    from deedsbcv_amd import _lib, api
    lib = _lib.load()
    ctx = api.Context(0)
    V = ctypes.c_void_p
    x = torch.zeros(8, dtype=torch.int16, device="cuda")
    out = torch.zeros(8, device="cuda")
    c = torch.zeros(1, dtype=torch.int32, device="cuda")
    assert lib.deeds_ctx_dice_noalloc(ctx.h, V(x.data_ptr()), V(x.data_ptr()), 1 << 32, 0, None, None,
                                      V(c.data_ptr())) == 1
    assert b"32-bit" in lib.deeds_ctx_last_error(ctx.h)
    assert lib.deeds_ctx_dice_noalloc(ctx.h, V(x.data_ptr()), V(x.data_ptr()), 8, 70000, V(x.data_ptr()),
                                      V(out.data_ptr()), V(c.data_ptr())) == 1
