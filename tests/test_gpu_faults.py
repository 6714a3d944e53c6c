"""Failure handling: allocation and kernel-launch failures injected from the
host (deeds_inject_fault) come back as statuses -- DEEDS_ERR_ALLOC /
DEEDS_ERR_HIP with a message -- at every point of a registration tried, the
failed call leaves nothing running, and the library then works normally again
(the next call's results equal those of a call that never saw a fault).  No
kernel is disturbed: an injected launch failure is reported after the kernel
was queued normally, an injected allocation failure before hipMalloc."""
import numpy as np
import pytest

from deedsbcv_amd import _lib, api, phantom

pytestmark = pytest.mark.gpu

SCHED = dict(grid=(6, 5, 4), radius=(3, 2, 2), quant=(2, 2, 1), alpha=1.6)
ALLOC, LAUNCH = 1, 2


@pytest.fixture(scope="module")
def pair_and_result():
    f, m = phantom.make_pair((48, 44, 40), 91)
    flow, warped, _ = api.register(f, m, **SCHED)
    return f, m, flow, warped


def _disarm():
    lib = _lib.load()
    lib.deeds_inject_fault(ALLOC, -1)
    lib.deeds_inject_fault(LAUNCH, -1)


@pytest.mark.parametrize("kind,after", [(ALLOC, 0), (ALLOC, 4), (ALLOC, 12), (LAUNCH, 0), (LAUNCH, 37),
                                        (LAUNCH, 120)])
def test_injected_fault_is_a_status_then_recovery(pair_and_result, kind, after):
    torch = pytest.importorskip("torch")
    f, m, flow0, warped0 = pair_and_result
    print(f"case: injected {'alloc' if kind == ALLOC else 'launch'} fault after {after} events", flush=True)
    lib = _lib.load()
    ctx = api.Context(0)  # a fresh context: its workspace is allocated during the call
    fx = torch.from_numpy(f.astype(np.float32)).cuda()
    mv = torch.from_numpy(m.astype(np.float32)).cuda()
    o, n, mm = f.shape
    g = SCHED["grid"][-1]
    flow = torch.zeros(3 * (mm // g) * (n // g) * (o // g), device="cuda")
    warped = torch.zeros_like(fx)
    torch.cuda.synchronize()
    assert lib.deeds_inject_fault(kind, after) == 0
    try:
        with pytest.raises(api.DeedsError) as e:
            ctx.register(fx, mv, flow, warped, **SCHED)
    finally:
        _disarm()
    assert e.value.code == (3 if kind == ALLOC else 2)
    assert "injected" in str(e.value)
    # the same context and the host-pointer form work normally afterwards
    ctx.register(fx, mv, flow, warped, **SCHED)
    torch.cuda.synchronize()
    assert np.array_equal(flow.cpu().numpy(), flow0)
    assert np.array_equal(warped.cpu().numpy(), warped0)
    f1, w1, _ = api.register(f, m, **SCHED)
    assert np.array_equal(f1, flow0) and np.array_equal(w1, warped0)


def test_injected_fault_in_host_form_then_recovery(pair_and_result):
    f, m, flow0, _ = pair_and_result
    lib = _lib.load()
    assert lib.deeds_inject_fault(LAUNCH, 3) == 0
    try:
        with pytest.raises(api.DeedsError):
            api.register(f, m, **SCHED)
    finally:
        _disarm()
    assert np.array_equal(api.register(f, m, **SCHED)[0], flow0)
    assert lib.deeds_inject_fault(7, 0) != 0  # unknown kind
