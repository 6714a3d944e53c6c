"""GPU parity of the differentiable soft variant (deedsbcv_amd.soft over the
deeds_hip C ABI) against the dense float64 checker oracle/soft_dense.py,
exhaustive enumeration and central finite differences.

Tolerances: float64 paths agree with the checker to 1e-9 relative (the
operations differ only in summation order); float32 paths to the stated
looser bounds; gradients of the full level step on a 48^3 crop (grid 6,
radius 6, quantisation 2, T = 0.5) agree with float64 central differences
(h = 1e-4) to 1e-4 relative.
"""
import sys
from pathlib import Path

import numpy as np
import pytest

torch = pytest.importorskip("torch")

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import soft_dense as sd  # noqa: E402

from deedsbcv_amd import api, phantom, soft  # noqa: E402

pytestmark = pytest.mark.gpu

F64 = torch.float64


def _cuda(a, dtype=F64):
    return torch.as_tensor(np.ascontiguousarray(a)).to("cuda", dtype)


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


@pytest.mark.parametrize("qs", [1, 2])
def test_features_match_dense(qs):
    print(f"case: soft features qs={qs}", flush=True)
    f, _ = phantom.make_pair((26, 22, 20), 40 + qs)
    want = sd.features(torch.from_numpy(f.astype(np.float64)), qs).numpy()
    got64 = soft.soft_features(_cuda(f), qs).cpu().numpy()
    assert _rel(got64, want) < 1e-9
    got32 = soft.soft_features(_cuda(f, torch.float32), qs).cpu().numpy()
    assert np.abs(got32 - want).max() < 2e-3


def test_features_threshold_to_plain_codes():
    """phi is the quantity descriptor() thresholds: quantising the float32 GPU
    features reproduces the plain GPU MIND codes almost everywhere."""
    print("case: soft features vs plain codes", flush=True)
    from test_soft_oracle import quantize
    f, _ = phantom.make_pair((40, 36, 32), 7)
    phi = soft.soft_features(_cuda(f, torch.float32), 2).cpu().numpy()
    codes = api.descriptor(f, 2)
    assert (quantize(phi) == codes).mean() > 0.999


def test_cost_matches_dense():
    print("case: soft cost", flush=True)
    rng = np.random.default_rng(5)
    o, n, m = 16, 17, 18
    for step, hw, q in [(5, 1, 2), (4, 2, 1), (6, 2, 2)]:
        F = rng.random((o, n, m, 12))
        M = rng.random((o, n, m, 12))
        want = sd.cost(torch.from_numpy(F), torch.from_numpy(M), step, hw, q).numpy()
        got = soft.soft_cost(_cuda(F), _cuda(M), step, hw, q).cpu().numpy()
        assert _rel(got, want) < 1e-12
        got32 = soft.soft_cost(_cuda(F, torch.float32), _cuda(M, torch.float32), step, hw, q).cpu().numpy()
        assert _rel(got32, want) < 1e-5


# (o, n, m), step, hw, quant: every shape with row kernels for the adjoint
# (step 6 / radius 6 / quantisation 2 is the soft level's), one with two
# i-tiles (m + 2 hw q above 256 lanes), and one that takes the per-node and
# pull kernels instead.
@pytest.mark.parametrize("dims,step,hw,q", [((26, 27, 28), 6, 6, 2), ((20, 19, 21), 6, 2, 2), ((13, 14, 17), 4, 2, 2),
                                            ((16, 17, 18), 5, 1, 2), ((10, 9, 262), 4, 2, 2), ((15, 16, 17), 4, 2, 1)])
def test_cost_adjoint_matches_dense_autograd(dims, step, hw, q):
    print(f"case: soft cost adjoint {dims} step {step} hw {hw} q {q}", flush=True)
    rng = np.random.default_rng(11)
    o, n, m = dims
    F = rng.random((o, n, m, 12))
    M = rng.random((o, n, m, 12))
    N = (m // step) * (n // step) * (o // step)
    gC = rng.normal(size=(N, (2 * hw + 1) ** 3))
    Fd = torch.from_numpy(F).requires_grad_(True)
    Md = torch.from_numpy(M).requires_grad_(True)
    sd.cost(Fd, Md, step, hw, q).backward(torch.from_numpy(gC))
    for dtype, tol in [(F64, 1e-11), (torch.float32, 2e-5)]:
        Fg = _cuda(F, dtype).requires_grad_(True)
        Mg = _cuda(M, dtype).requires_grad_(True)
        soft.soft_cost(Fg, Mg, step, hw, q).backward(_cuda(gC, dtype))
        assert _rel(Fg.grad.cpu(), Fd.grad) < tol, dtype
        assert _rel(Mg.grad.cpu(), Md.grad) < tol, dtype


@pytest.mark.parametrize("tree", [([-1, 0, 1], [0, 1, 2]), ([1, -1, 1, 1], [1, 0, 2, 3])])
def test_bp_matches_enumeration(tree):
    print(f"case: soft BP enumeration {tree}", flush=True)
    parents, ordered = tree
    N, hw, quant, T = len(parents), 1, 2, 0.5
    rng = np.random.default_rng(N)
    C = rng.random((N, 27)) * 4
    u0 = rng.normal(size=(3, N)) * 1.5
    w = rng.uniform(0.3, 1.0, N).astype(np.float32)
    want = sd.brute_force_bp(C, u0, parents, w, hw, quant, T)
    tr = (np.array(ordered, np.int32), np.array(parents, np.int32), w)
    got = soft.soft_bp(_cuda(C), _cuda(u0), tr, (1, 1, N), 1, hw, quant, T).cpu().numpy()
    assert np.abs(got - want).max() < 1e-12


def _mst_case(shape=(24, 20, 20), step=4, hw=2, quant=2, seed=3):
    f, _ = phantom.make_pair(shape, seed)
    tree = api.primsGraph(f, step)
    N = tree[0].size
    rng = np.random.default_rng(seed)
    C = rng.random((N, (2 * hw + 1) ** 3)) * 10
    u0 = rng.normal(size=(3, N)) * 2
    return f, tree, C, u0


@pytest.mark.parametrize("T", [0.5, 3.0])
def test_bp_matches_dense(T):
    print(f"case: soft BP dense T={T}", flush=True)
    shape, step, hw, quant = (24, 20, 20), 4, 2, 2
    f, tree, C, u0 = _mst_case(shape, step, hw, quant)
    want = sd.bp(torch.from_numpy(C), torch.from_numpy(u0), tree[0], tree[1], tree[2], hw, quant, T).numpy()
    got = soft.soft_bp(_cuda(C), _cuda(u0), tree, shape, step, hw, quant, T).cpu().numpy()
    assert _rel(got - u0, want - u0) < 1e-9
    got32 = soft.soft_bp(_cuda(C, torch.float32), _cuda(u0, torch.float32), tree, shape, step, hw, quant,
                         T).cpu().numpy()
    assert np.abs(got32 - want).max() < 1e-3


@pytest.mark.parametrize("shape,step,hw,quant", [((20, 18, 21), 4, 2, 2), ((19, 18, 20), 6, 1, 2)])
def test_vjps_match_dense_autograd(shape, step, hw, quant):
    print(f"case: soft VJPs vs dense autograd {shape} step {step}", flush=True)
    rng = np.random.default_rng(8)
    T, qs = 0.5, 2
    f, mv = phantom.make_pair(shape, 12)
    tree = api.primsGraph(f, step)
    N = tree[0].size
    u0 = rng.normal(size=(3, N))
    wgt = rng.normal(size=(3, N))

    # dense float64 reference on the CPU
    a = torch.from_numpy(f.astype(np.float64)).requires_grad_(True)
    b = torch.from_numpy(mv.astype(np.float64)).requires_grad_(True)
    Fd, Md = sd.features(a, qs), sd.features(b, qs)
    Cd = sd.cost(Fd, Md, step, hw, quant)
    ud = sd.bp(Cd, torch.from_numpy(u0), tree[0], tree[1], tree[2], hw, quant, T)
    Fd.retain_grad()
    Md.retain_grad()
    Cd.retain_grad()
    (ud * torch.from_numpy(wgt)).sum().backward()

    # HIP float64 through the autograd Functions
    x = _cuda(f).requires_grad_(True)
    y = _cuda(mv).requires_grad_(True)
    Fg, Mg = soft.soft_features(x, qs), soft.soft_features(y, qs)
    Cg = soft.soft_cost(Fg, Mg, step, hw, quant)
    ug = soft.soft_bp(Cg, _cuda(u0), tree, shape, step, hw, quant, T)
    Fg.retain_grad()
    Mg.retain_grad()
    Cg.retain_grad()
    (ug * _cuda(wgt)).sum().backward()

    assert _rel(ug.detach().cpu(), ud.detach()) < 1e-9
    assert _rel(Cg.grad.cpu(), Cd.grad) < 1e-9
    assert _rel(Fg.grad.cpu(), Fd.grad) < 1e-9
    assert _rel(Mg.grad.cpu(), Md.grad) < 1e-9
    assert _rel(x.grad.cpu(), a.grad) < 1e-9
    assert _rel(y.grad.cpu(), b.grad) < 1e-9


def test_gradient_matches_finite_differences_48():
    """The soft level step at grid 6, radius 6 (2197 labels), quantisation 2,
    T = 0.5 on a 48^3 crop: directional derivatives of the float64 gradient
    along unit-norm random directions against float64 central differences
    with h = 1e-4; the float32 gradient against the float64 one.  (The
    minimum over the twelve SSC distances makes the features piecewise
    smooth; unit-norm directions keep each voxel's step near 1e-7, so a
    central difference almost never straddles a switch of that minimum.)"""
    print("case: soft FD 48^3", flush=True)
    f, mv = phantom.workload_soft(crop=48)
    grid, hw, quant, T = 6, 6, 2, 0.5
    tree = soft.tree_of(torch.from_numpy(f), grid)
    rng = np.random.default_rng(2)
    wgt = _cuda(rng.normal(size=(3, 8, 8, 8)))

    def loss(a, b):
        return (soft.soft_register(a, b, grid, hw, quant, T=T, tree=tree) * wgt).sum()

    x = _cuda(f).requires_grad_(True)
    y = _cuda(mv).requires_grad_(True)
    L = loss(x, y)
    gx, gy = torch.autograd.grad(L, (x, y))
    h = 1e-4
    errs = []
    for trial in range(3):
        dx = _cuda(rng.normal(size=f.shape))
        dy = _cuda(rng.normal(size=f.shape))
        norm = float(torch.sqrt((dx * dx).sum() + (dy * dy).sum()))
        dx, dy = dx / norm, dy / norm
        with torch.no_grad():
            fd = float((loss(x + h * dx, y + h * dy) - loss(x - h * dx, y - h * dy)) / (2 * h))
        an = float((gx * dx).sum() + (gy * dy).sum())
        errs.append(abs(fd - an) / abs(an))
    print("fd relative errors", errs, flush=True)
    assert max(errs) < 1e-4

    x32 = x.detach().float().requires_grad_(True)
    y32 = y.detach().float().requires_grad_(True)
    L32 = (soft.soft_register(x32, y32, grid, hw, quant, T=T, tree=tree) * wgt.float()).sum()
    g32x, g32y = torch.autograd.grad(L32, (x32, y32))
    rel32 = max(_rel(g32x.double().cpu(), gx.cpu()), _rel(g32y.double().cpu(), gy.cpu()))
    print("float32 gradient relative error", rel32, flush=True)
    assert rel32 < 1e-2


def test_cost_bp_gradient_matches_finite_differences_48():
    """The smooth part (features -> cost -> BP -> expected displacement) at the
    same 48^3 geometry: central differences with h = 1e-4 along directions
    with a standard normal entry per feature value."""
    print("case: soft FD features 48^3", flush=True)
    f, mv = phantom.workload_soft(crop=48)
    grid, hw, quant, T = 6, 6, 2, 0.5
    tree = soft.tree_of(torch.from_numpy(f), grid)
    rng = np.random.default_rng(3)
    F = soft.soft_features(_cuda(f), 2).detach()
    M = soft.soft_features(_cuda(mv), 2).detach()
    u0 = torch.zeros((3, 512), dtype=F64, device="cuda")
    wgt = _cuda(rng.normal(size=(3, 512)))

    def loss(a, b):
        return (soft.soft_bp(soft.soft_cost(a, b, grid, hw, quant), u0, tree, f.shape, grid, hw, quant, T)
                * wgt).sum()

    a = F.clone().requires_grad_(True)
    b = M.clone().requires_grad_(True)
    ga, gb = torch.autograd.grad(loss(a, b), (a, b))
    h = 1e-4
    errs = []
    for trial in range(3):
        da = _cuda(rng.normal(size=F.shape))
        db = _cuda(rng.normal(size=M.shape))
        with torch.no_grad():
            fd = float((loss(F + h * da, M + h * db) - loss(F - h * da, M - h * db)) / (2 * h))
        an = float((ga * da).sum() + (gb * db).sum())
        errs.append(abs(fd - an) / abs(an))
    print("fd relative errors", errs, flush=True)
    assert max(errs) < 1e-4


def test_soft_errors():
    x = _cuda(np.zeros((8, 8, 8)), torch.float16)
    with pytest.raises(TypeError):
        soft.soft_features(x, 1)
    C = _cuda(np.zeros((8, 27)))
    u0 = _cuda(np.zeros((3, 8))).requires_grad_(True)
    tree = (np.arange(8, dtype=np.int32), np.array([-1, 0, 0, 1, 2, 3, 4, 5], np.int32), np.ones(8, np.float32))
    with pytest.raises(ValueError):
        soft.soft_bp(C, u0, tree, (8, 8, 8), 4, 1, 1)
    bad = (np.arange(8, dtype=np.int32), np.array([-1, 0, 0, 1, 2, 3, 7, 5], np.int32), np.ones(8, np.float32))
    with pytest.raises(api.DeedsError):  # parent after child in BFS order
        soft.soft_bp(C, u0.detach(), bad, (8, 8, 8), 4, 1, 1)
    with pytest.raises(api.DeedsError):  # six 17^3 double cubes exceed the LDS
        soft.soft_bp(_cuda(np.zeros((8, 17 ** 3))), u0.detach(), tree, (8, 8, 8), 4, 8, 1)
