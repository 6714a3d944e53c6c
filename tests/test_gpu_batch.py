"""GPU tests of the multi-atlas batch: pairs of different sizes registered
concurrently give exactly the results of registering each pair alone."""
import numpy as np
import pytest

from deedsbcv_amd import api, phantom

pytestmark = pytest.mark.gpu

SCHED = dict(grid=(6, 5, 4), radius=(3, 2, 2), quant=(2, 2, 1), alpha=1.6)


def _pairs():
    shapes = [(48, 44, 40), (52, 40, 36), (40, 48, 44), (56, 44, 38), (44, 44, 44)]
    out = []
    for i, s in enumerate(shapes):
        f, m = phantom.make_pair(s, 300 + i)
        if i == 2:
            X = np.eye(4, dtype=np.float32)
            X[0, 3], X[1, 1] = 0.75, 1.01
            out.append((f, m, X.T.ravel().copy()))
        else:
            out.append((f, m))
    return out


def test_batch_equals_sequential():
    print("case: batch of 5 pairs vs sequential", flush=True)
    pairs = _pairs()
    got = api.register_batch(pairs, concurrency=3, **SCHED)
    for p, (flow, warped, t) in zip(pairs, got):
        X = p[2] if len(p) > 2 else None
        f1, w1, _ = api.register(p[0], p[1], X=X, **SCHED)
        assert np.array_equal(flow, f1)
        assert np.array_equal(warped, w1)
        assert t["hamming_evals"] > 0


def test_batch_device_equals_host():
    torch = pytest.importorskip("torch")
    print("case: device batch", flush=True)
    pairs = _pairs()[:3]
    host = api.register_batch(pairs, concurrency=2, **SCHED)
    dev = []
    for p in pairs:
        f = torch.from_numpy(p[0].astype(np.float32)).cuda()
        m = torch.from_numpy(p[1].astype(np.float32)).cuda()
        o, n, mm = f.shape
        g = SCHED["grid"][-1]
        dev.append(dict(fixed=f, moving=m, X=p[2] if len(p) > 2 else None,
                        flow=torch.empty(3 * (mm // g) * (n // g) * (o // g), device="cuda"),
                        warped=torch.empty_like(f)))
    api.register_batch_device(dev, concurrency=3, **SCHED)
    torch.cuda.synchronize()
    for d, (flow, warped, _) in zip(dev, host):
        assert np.array_equal(d["flow"].cpu().numpy(), flow)
        assert np.array_equal(d["warped"].cpu().numpy(), warped)


def test_batch_rejects_bad_pair():
    f, m = phantom.make_pair((40, 40, 40), 1)
    with pytest.raises(api.DeedsError):
        api.register_batch([(f, m), (f[:2], m[:2])], concurrency=2, **SCHED)


def test_batch_graph_replays_equal_sequential():
    """Device batches of the same pairs call after call: the first runs its
    level phases on the streams, the second captures them into HIP graphs, the
    later ones replay the graphs.  Every call's flows and deformed images equal
    one-at-a-time registrations (eager work and host synchronisations between
    the calls)."""
    torch = pytest.importorskip("torch")
    print("case: batch graph capture and replays", flush=True)
    pairs = _pairs()
    seq = [api.register(p[0], p[1], X=p[2] if len(p) > 2 else None, **SCHED)[:2] for p in pairs]
    dev = []
    g = SCHED["grid"][-1]
    for p in pairs:
        f = torch.from_numpy(p[0].astype(np.float32)).cuda()
        m = torch.from_numpy(p[1].astype(np.float32)).cuda()
        o, n, mm = f.shape
        dev.append(dict(fixed=f, moving=m, X=p[2] if len(p) > 2 else None,
                        flow=torch.empty(3 * (mm // g) * (n // g) * (o // g), device="cuda"),
                        warped=torch.empty_like(f)))
    before = api.tuning_report()["graphs"]
    for call in range(5):
        for d in dev:
            d["flow"].fill_(-1.0)
            d["warped"].fill_(-1.0)
        torch.cuda.synchronize()
        api.register_batch_device(dev, concurrency=5, **SCHED)
        torch.cuda.synchronize()
        for i, (d, (flow, warped)) in enumerate(zip(dev, seq)):
            assert np.array_equal(d["flow"].cpu().numpy(), flow), (call, i)
            assert np.array_equal(d["warped"].cpu().numpy(), warped), (call, i)
        _ = api.register(pairs[0][0], pairs[0][1], **SCHED)  # eager work between the calls
    after = api.tuning_report()["graphs"]
    assert after["batch_graph_captures"] >= before["batch_graph_captures"] + 1
    assert after["batch_graph_replays"] >= before["batch_graph_replays"] + 3
