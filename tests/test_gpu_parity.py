"""GPU parity: the HIP library (through its C ABI) against the oracle and the
golden vectors of the unmodified reference.  Bar: bit-exact for codes, costs,
trees, BP labels and every warp (the kernels reproduce the reference's float /
double evaluation exactly), so every comparison below is array_equal."""
import numpy as np
import pytest

from deedsbcv_amd import api, phantom

pytestmark = pytest.mark.gpu


def _eq(a, b, what):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, (what, a.shape, b.shape)
    if not np.array_equal(a, b):
        bad = np.flatnonzero(a.ravel() != b.ravel())
        raise AssertionError(f"{what}: {bad.size} of {a.size} differ; first at {bad[:5]}: "
                             f"{a.ravel()[bad[:5]]} vs {b.ravel()[bad[:5]]}")


# --------------------------------------------------------------- MIND -------

def test_mind_golden(golden):
    print("case: mind golden 40x36x32", flush=True)
    for qs in (1, 2):
        _eq(api.descriptor(golden["im"], qs), golden[f"mind_q{qs}"], f"mind qs={qs}")
    _eq(api.descriptor(golden["im2"], 2), golden["mind2_q2"], "mind im2")


@pytest.mark.parametrize("shape,seed", [((37, 33, 29), 21), ((64, 48, 40), 22), ((23, 70, 31), 23)])
def test_mind_vs_oracle(oracle, shape, seed):
    print(f"case: mind vs oracle {shape}", flush=True)
    f, _ = phantom.make_pair(shape, seed)
    im = f.astype(np.float32) + 1024
    for qs in (1, 2, 3):
        _eq(api.descriptor(im, qs), oracle.descriptor(im, qs), f"mind {shape} qs={qs}")


# ---------------------------------------------------------- data cost -------

@pytest.mark.parametrize("case", [(5, 2, 2), (4, 2, 1), (6, 1, 3)])
def test_cost_golden(golden, case):
    step, hw, q = case
    print(f"case: datacost golden {case}", flush=True)
    _eq(api.dataCostCL(golden["mind_q2"], golden["mind2_q2"], step, hw, q),
        golden[f"cost_s{step}h{hw}q{q}"], f"cost {case}")


@pytest.mark.parametrize("case", [(8, 3, 5, 1), (7, 2, 4, 1), (5, 4, 2, 1), (4, 4, 1, 1), (3, 2, 1, 1),
                                  (8, 2, 3, 2), (5, 2, 2, 0), (4, 1, 1, 2)])
def test_cost_vs_oracle(oracle, case):
    step, hw, q, randnum = case
    print(f"case: datacost vs oracle {case}", flush=True)
    f, mv = phantom.make_pair((45, 41, 38), 31)
    F = oracle.descriptor(f.astype(np.float32) + 1024, 2)
    M = oracle.descriptor(mv.astype(np.float32) + 1024, 2)
    _eq(api.dataCostCL(F, M, step, hw, q, 1.6, randnum), oracle.datacost(F, M, step, hw, q, 1.6, randnum),
        f"cost {case}")


# The five levels of the default schedule have dedicated row kernels; the
# 256-wide shape gives rows as long as the 256x256x128 workload's (one wave
# of position runs nearly full), the 45-wide one short rows.
@pytest.mark.parametrize("case", [(8, 8, 5), (7, 7, 4), (6, 6, 3), (5, 5, 2), (4, 4, 1)])
@pytest.mark.parametrize("shape", [(45, 41, 38), (256, 20, 18)])
def test_cost_schedule_levels_vs_oracle(oracle, case, shape):
    step, hw, q = case
    print(f"case: datacost schedule level {case} {shape}", flush=True)
    f, mv = phantom.make_pair(shape, 32)
    F = oracle.descriptor(f.astype(np.float32) + 1024, 2)
    M = oracle.descriptor(mv.astype(np.float32) + 1024, 2)
    _eq(api.dataCostCL(F, M, step, hw, q, 1.6, 1), oracle.datacost(F, M, step, hw, q, 1.6, 1),
        f"cost {case} {shape}")


# ---------------------------------------------------------------- MST -------

@pytest.mark.parametrize("case", [(5, 2, 2), (4, 2, 1), (6, 1, 3)])
def test_prims_golden(golden, case):
    step = case[0]
    tag = "s%dh%dq%d" % case
    print(f"case: prims golden {case}", flush=True)
    got = api.primsGraph(golden["im2"] if step == 4 else golden["im"], step)
    for a, name in zip(got, ("ordered", "parents", "edgemst")):
        _eq(a, golden[f"{name}_{tag}"], name)


@pytest.mark.parametrize("step", [3, 4, 5, 7, 8])
def test_prims_vs_oracle(oracle, step):
    print(f"case: prims vs oracle step={step}", flush=True)
    f, mv = phantom.make_pair((57, 50, 44), 41)
    for im in (f.astype(np.float32) + 1024, mv.astype(np.float32) * 0.73 + 1000.25):  # non-integer image too
        got = api.primsGraph(im, step)
        want = oracle.prims(im, step)
        for a, b, name in zip(got, want, ("ordered", "parents", "edgemst")):
            _eq(a, b, f"{name} step={step}")


def test_prims_many_distinct_costs_vs_oracle(oracle):
    """Integer image whose edge costs take more than 65536 distinct values:
    Prim's keys fall back from 16-bit ranks to floats, same tree."""
    print("case: prims with >65536 distinct integer costs", flush=True)
    rng = np.random.default_rng(7)
    im = rng.integers(0, 60000, size=(96, 160, 160)).astype(np.float32)
    got = api.primsGraph(im, 4)
    want = oracle.prims(im, 4)
    for a, b, name in zip(got, want, ("ordered", "parents", "edgemst")):
        _eq(a, b, f"{name} many distinct costs")


@pytest.mark.parametrize("step", [4, 6, 7])
def test_prims_device_image_vs_oracle(oracle, step):
    """deeds_ctx_prims_graph (device image, statistics on the GPU) builds the oracle's tree."""
    torch = pytest.importorskip("torch")
    print(f"case: device-image prims vs oracle step={step}", flush=True)
    ctx = api.Context(0)
    f, mv = phantom.make_pair((61, 54, 47), 43)
    for im in (f.astype(np.float32) + 1024, mv.astype(np.float32) * 0.61 + 999.5):
        t = torch.from_numpy(im).cuda()
        torch.cuda.synchronize()
        got = ctx.prims_graph(t, step)
        want = oracle.prims(im, step)
        for a, b, name in zip(got, want, ("ordered", "parents", "edgemst")):
            _eq(a, b, f"{name} step={step}")


# ----------------------------------------------------------------- BP -------

@pytest.mark.parametrize("case", [(5, 2, 2), (4, 2, 1), (6, 1, 3)])
def test_bp_golden(golden, case):
    step, hw, q = case
    tag = f"s{step}h{hw}q{q}"
    print(f"case: bp golden {case}", flush=True)
    u0 = golden[f"u0_{tag}"]
    got = api.regularisationCL(golden[f"cost_{tag}"], u0[0], u0[1], u0[2], hw, step, q, golden[f"ordered_{tag}"],
                               golden[f"parents_{tag}"], golden[f"edgemst_{tag}"], (32, 36, 40))
    _eq(np.stack(got), golden[f"u1_{tag}"], f"u1 {case}")


@pytest.mark.parametrize("case", [(8, 8, 5), (6, 6, 3), (4, 4, 1), (5, 3, 2), (3, 1, 1)])
def test_bp_vs_oracle(oracle, case):
    step, hw, q = case
    print(f"case: bp vs oracle {case}", flush=True)
    rng = np.random.default_rng(step * 100 + hw)
    f, mv = phantom.make_pair((48, 44, 40), 51)
    im = f.astype(np.float32) + 1024
    F = oracle.descriptor(im, 2)
    M = oracle.descriptor(mv.astype(np.float32) + 1024, 2)
    cost = oracle.datacost(F, M, step, hw, q)
    tree = oracle.prims(im, step)
    o1, n1, m1 = 40 // step, 44 // step, 48 // step
    u0 = (rng.normal(size=(3, o1, n1, m1)) * 2).astype(np.float32)
    got = api.regularisationCL(cost, u0[0], u0[1], u0[2], hw, step, q, *tree, (40, 44, 48))
    want = oracle.regularisation(cost, u0[0], u0[1], u0[2], hw, step, q, *tree, (48, 44, 40))
    _eq(np.stack(got), np.stack(want), f"bp {case}")


def _bfs_tree(par, rng):
    """A tree given by parent links (par[0] = -1) as (ordered, parents) with
    shuffled node ids, `ordered` sorted by (depth, id) as primsGraph returns it."""
    n = len(par)
    depth = np.zeros(n, np.int64)
    for v in range(1, n):
        depth[v] = depth[par[v]] + 1
    lab = np.concatenate([[0], 1 + rng.permutation(n - 1)])
    parents = np.full(n, -1, np.int32)
    parents[lab[1:]] = lab[np.asarray(par[1:])]
    nd = np.empty(n, np.int64)
    nd[lab] = depth
    ordered = np.lexsort((np.arange(n), nd)).astype(np.int32)
    return ordered, parents


@pytest.mark.parametrize("kind", ["star", "bushy", "chain", "binary"])
def test_bp_caller_trees_vs_oracle(oracle, kind):
    """regularisationCL on trees a caller hands in (not grid MSTs): nodes with
    up to 40 children (more than a grid node has), long single-child chains
    and bushy trees exercise every fold path of the segment programs."""
    print(f"case: bp caller tree {kind}", flush=True)
    rng = np.random.default_rng(17)
    step, hw, q = 4, 2, 2
    o1, n1, m1 = 5, 6, 7
    n = o1 * n1 * m1
    if kind == "star":
        par = [-1] + [0] * 40 + [int(rng.integers(1, 41)) for _ in range(n - 41)]
        par = par[:41] + [int(rng.integers(0, v)) if v > 120 else par[v] for v in range(41, n)]
    elif kind == "bushy":
        par = [-1] + [int(rng.integers(0, v)) for v in range(1, n)]
    elif kind == "chain":
        par = [-1] + [v - 1 if v % 37 else int(rng.integers(0, v)) for v in range(1, n)]
    else:
        par = [-1] + [(v - 1) // 2 for v in range(1, n)]
    ordered, parents = _bfs_tree(par, rng)
    edgemst = rng.uniform(0.1, 1.0, n).astype(np.float32)
    len3 = (2 * hw + 1) ** 3
    cost = rng.uniform(0.0, 50.0, (n, len3)).astype(np.float32)
    u0 = (rng.normal(size=(3, o1, n1, m1)) * 2).astype(np.float32)
    shape = (o1 * step, n1 * step, m1 * step)
    got = api.regularisationCL(cost, u0[0], u0[1], u0[2], hw, step, q, ordered, parents, edgemst, shape)
    want = oracle.regularisation(cost, u0[0], u0[1], u0[2], hw, step, q, ordered, parents, edgemst, shape[::-1])
    _eq(np.stack(got), np.stack(want), f"bp caller tree {kind}")


# ---------------------------------------------------------- transforms ------

def test_transforms_golden(golden):
    print("case: transforms golden", flush=True)
    g = golden["grid6"]
    _eq(np.stack(api.upsampleDeformationsCL(g[0], g[1], g[2], (32, 36, 40))), golden["up"], "upsample")
    _eq(np.stack(api.consistentMappingCL(*g, 5)), golden["cons"], "consistent")
    ux, vx, wx = golden["up"]
    w, ssd = api.warpAffine(golden["im2"], golden["im"], golden["X"], ux, vx, wx)
    _eq(w, golden["warp_affine"], "warpAffine")
    np.testing.assert_allclose(ssd, golden["warp_affine_ssd"], rtol=1e-4)  # logging-only figure
    w, ssd = api.warpImageCL(golden["im2"], golden["im"], ux, vx, wx)
    _eq(w, golden["warp_image"], "warpImageCL")
    np.testing.assert_allclose(ssd, golden["warp_image_ssd"], rtol=1e-4)
    _eq(api.warpAffineS(golden["seg"], golden["X"], ux, vx, wx), golden["warp_seg"], "warpAffineS")


def test_transforms_vs_oracle(oracle):
    print("case: transforms vs oracle", flush=True)
    rng = np.random.default_rng(8)
    for (m2, n2, o2), (m, n, o) in (((7, 6, 5), (50, 45, 41)), ((12, 9, 10), (12, 9, 10)), ((3, 4, 5), (31, 29, 23))):
        g = (rng.normal(size=(6, o2, n2, m2)) * 3).astype(np.float32)
        _eq(np.stack(api.upsampleDeformationsCL(g[0], g[1], g[2], (o, n, m))),
            np.stack(oracle.upsample(g[0], g[1], g[2], (m, n, o))), "upsample")
        for factor in (4, 5, 8):
            _eq(np.stack(api.consistentMappingCL(*g, factor)), np.stack(oracle.consistent(*g, factor)), "consistent")


def test_jacobian_logging(oracle):
    rng = np.random.default_rng(9)
    g = (rng.normal(size=(3, 9, 10, 11)) * 2).astype(np.float32)
    s, neg = api.jacobian(g[0], g[1], g[2], 4)
    s0, neg0 = oracle.jacobian(g[0], g[1], g[2], 4)
    assert abs(s - float(s0)) <= 1e-4 * max(1.0, float(s0))
    assert abs(neg - float(neg0)) < 1e-6
    # the partial sums are added in a fixed order: the figures repeat exactly
    assert api.jacobian(g[0], g[1], g[2], 4) == (s, neg)


def test_logging_figures_repeat_exactly():
    print("case: logging figures of two identical registrations", flush=True)
    f, mv = phantom.make_pair((48, 44, 40), 77)
    sched = dict(grid=(6, 5, 4), radius=(4, 3, 2), quant=(3, 2, 1))
    _, _, t1 = api.register(f, mv, sched["grid"], sched["radius"], sched["quant"], 1.6)
    _, _, t2 = api.register(f, mv, sched["grid"], sched["radius"], sched["quant"], 1.6)
    for k in ("ssd_before", "ssd_after", "jac_std", "jac_neg"):
        assert t1[k] == t2[k], (k, t1[k], t2[k])


def test_apply_matches_reference_applyBCV(golden):
    print("case: applyBCV golden", flush=True)
    s = golden["e2e_schedule"]
    seg = golden["e2e_seg"]
    _eq(api.apply_seg(seg, golden["e2e_flow"], int(s[0][-1])), golden["e2e_apply_seg"], "applyBCV")


# --------------------------------------------------------- end to end -------

def test_register_golden(golden):
    print("case: register golden 48x40x36", flush=True)
    s = golden["e2e_schedule"]
    flow, warped, t = api.register(golden["e2e_fixed"], golden["e2e_moving"], s[0], s[1], s[2])
    _eq(flow, golden["e2e_flow"], "flow")
    _eq(warped, golden["e2e_warped"], "warped")
    seg = api.apply_seg(golden["e2e_seg"], flow, int(s[0][-1]))
    _eq(seg, golden["e2e_warped_seg"], "warped seg")


def test_register_small_workload_vs_oracle(oracle):
    """The 96x96x80 three-level workload: GPU == oracle (the oracle == the reference
    binary on this workload is tests/test_oracle.py::test_oracle_register_small_workload_vs_ref_binary)."""
    print("case: register small workload vs oracle", flush=True)
    f, mv = phantom.workload_small(0)
    sch = phantom.SCHEDULE_SMALL
    flow, warped, t = api.register(f, mv, sch["grid"], sch["radius"], sch["quant"], sch["alpha"])
    oflow, owarped = oracle.register(f, mv, sch["grid"], sch["radius"], sch["quant"], sch["alpha"])
    _eq(flow, oflow, "flow")
    _eq(warped, owarped, "warped")
    assert t["hamming_evals"] > 0 and t["total"] > 0


def test_register_deterministic_and_affine():
    print("case: register determinism with affine", flush=True)
    f, mv = phantom.make_pair((64, 56, 48), 61)
    X = np.eye(4, dtype=np.float32)
    X[0, 3], X[1, 1], X[2, 0] = 1.25, 1.02, 0.01
    Xc = X.T.ravel().copy()
    a = api.register(f, mv, (6, 5, 4), (4, 3, 2), (3, 2, 1), 1.6, X=Xc)
    b = api.register(f, mv, (6, 5, 4), (4, 3, 2), (3, 2, 1), 1.6, X=Xc)
    _eq(a[0], b[0], "flow determinism")
    _eq(a[1], b[1], "warped determinism")


def test_register_with_affine_vs_oracle(oracle):
    print("case: register with affine vs oracle", flush=True)
    f, mv = phantom.make_pair((52, 48, 44), 71)
    X = np.eye(4, dtype=np.float32)
    X[0, 3], X[1, 1], X[2, 0], X[1, 2] = -1.5, 0.97, 0.015, -0.02
    Xc = X.T.ravel().copy()
    flow, warped, _ = api.register(f, mv, (6, 4), (3, 2), (2, 1), 1.6, X=Xc)
    oflow, owarped = oracle.register(f, mv, (6, 4), (3, 2), (2, 1), 1.6, X=Xc)
    _eq(flow, oflow, "flow")
    _eq(warped, owarped, "warped")


# ------------------------------------------------------------ errors --------

def test_errors_are_statuses_not_faults():
    im = np.zeros((10, 10, 10), np.float32)
    with pytest.raises(api.DeedsError):
        api.descriptor(im, 6)  # box window larger than the volume
    with pytest.raises(api.DeedsError):
        api.dataCostCL(np.zeros((10, 10, 10), np.uint64), np.zeros((10, 10, 10), np.uint64), 20, 1, 1)
    with pytest.raises(api.DeedsError):
        api.regularisationCL(np.zeros((8, 27), np.float32), *np.zeros((3, 2, 2, 2), np.float32), 1, 5, 1,
                             np.array([0, 0, 1, 2, 3, 4, 5, 6], np.int32), np.zeros(8, np.int32),
                             np.ones(8, np.float32), (10, 10, 10))
    # the library still works afterwards
    assert api.descriptor(im, 1).shape == im.shape


# ----------------------------------------------------- argmin tables --------

def _tree_case(oracle, shape, step, hw, q, seed):
    rng = np.random.default_rng(seed)
    f, mv = phantom.make_pair(shape, seed)
    im = f.astype(np.float32) + 1024
    F = oracle.descriptor(im, 2)
    M = oracle.descriptor(mv.astype(np.float32) + 1024, 2)
    cost = oracle.datacost(F, M, step, hw, q)
    tree = oracle.prims(im, step)
    m, n, o = shape
    u0 = (rng.normal(size=(3, o // step, n // step, m // step)) * 2).astype(np.float32)
    return cost, tree, u0


@pytest.mark.parametrize("case", [(8, 8, 5), (6, 4, 3), (4, 2, 1), (5, 3, 2)])
def test_bp_argmin_tables_vs_oracle(oracle, case):
    """deeds_ctx_regularisation's d_allinds equals the oracle's messageDT argmin
    tables for every node and label (not only the selected labels), the root's
    row being the identity (regularisation.h:273-274).  The Hamming costs are
    multiples of one float scale, so equal sums tie exactly and the lowest-index
    rule of regularisation.h:45-51 decides many entries."""
    torch = pytest.importorskip("torch")
    step, hw, q = case
    print(f"case: bp argmin tables {case}", flush=True)
    shape = (48, 44, 40)
    cost, tree, u0 = _tree_case(oracle, shape, step, hw, q, 55 + step)
    want = oracle.regularisation(cost, u0[0], u0[1], u0[2], hw, step, q, *tree, shape, want_inds=True)
    ctx = api.Context(0)
    dcost = torch.from_numpy(cost).cuda()
    du0 = torch.from_numpy(u0.reshape(3, -1)).cuda()
    inds = torch.full(cost.shape, -1, dtype=torch.int16, device="cuda")
    u1 = ctx.regularisation(dcost, du0, hw, step, q, tree, shape[::-1], allinds=inds)
    torch.cuda.synchronize()
    _eq(u1.cpu().numpy(), np.stack(want[:3]).reshape(3, -1), f"u1 {case}")
    _eq(inds.cpu().numpy(), want[3], f"allinds {case}")
    root = int(tree[0][0])
    _eq(inds[root].cpu().numpy(), np.arange(cost.shape[1], dtype=np.int16), "root row")


def test_bp_wide_depth_vs_oracle(oracle):
    """A caller tree with 25 000 nodes at one depth (more label rows than LDS
    holds): the selection keeps them in the tree's global scratch and gives the
    oracle's labels and tables."""
    torch = pytest.importorskip("torch")
    print("case: bp wide depth (25000 nodes at depth 2)", flush=True)
    rng = np.random.default_rng(23)
    step, hw, q = 1, 1, 1
    o1, n1, m1 = 9, 50, 57  # 25 650 nodes: root, 200 children, 25 449 grandchildren
    n = o1 * n1 * m1
    par = [-1] + [0] * 200 + [1 + (v - 201) % 200 for v in range(201, n)]
    ordered, parents = _bfs_tree(par, rng)
    width = np.bincount(np.r_[0, np.ones(200, int), 2 * np.ones(n - 201, int)]).max()
    assert width > 20000
    edgemst = rng.uniform(0.1, 1.0, n).astype(np.float32)
    cost = rng.integers(0, 40, (n, 27)).astype(np.float32) * np.float32(0.25)
    u0 = (rng.normal(size=(3, o1, n1, m1)) * 2).astype(np.float32)
    shape = (o1 * step, n1 * step, m1 * step)
    want = oracle.regularisation(cost, u0[0], u0[1], u0[2], hw, step, q, ordered, parents, edgemst, shape[::-1],
                                 want_inds=True)
    ctx = api.Context(0)
    inds = torch.full(cost.shape, -1, dtype=torch.int16, device="cuda")
    u1 = ctx.regularisation(torch.from_numpy(cost).cuda(), torch.from_numpy(u0.reshape(3, -1)).cuda(), hw, step, q,
                            (ordered, parents, edgemst), shape, allinds=inds)
    torch.cuda.synchronize()
    _eq(u1.cpu().numpy(), np.stack(want[:3]).reshape(3, -1), "u1 wide")
    _eq(inds.cpu().numpy(), want[3], "allinds wide")


# ------------------------------------------------ device field forms --------

@pytest.mark.parametrize("on", ["default stream", "side stream"])
def test_device_field_forms_vs_oracle(oracle, on):
    """deeds_ctx_upsample_deformations / consistent_mapping / warp_affine /
    warp_image / warp_affine_s / apply_float / apply_seg / jacobian on device
    tensors give the oracle's arrays bit for bit (the logging figures to 1e-4),
    with torch's work on its default (null) stream and on a stream of its own."""
    torch = pytest.importorskip("torch")
    print(f"case: device field forms vs oracle ({on})", flush=True)
    if on == "side stream":
        with torch.cuda.stream(torch.cuda.Stream()):
            _device_field_forms(oracle, torch)
    else:
        _device_field_forms(oracle, torch)


def _device_field_forms(oracle, torch):
    rng = np.random.default_rng(81)
    ctx = api.Context(0)
    m2, n2, o2, m, n, o = 9, 8, 7, 45, 40, 36
    g = (rng.normal(size=(6, o2, n2, m2)) * 2.5).astype(np.float32)
    up = ctx.upsample_deformations(torch.from_numpy(g[:3].copy()).cuda(), (o, n, m))
    want_up = np.stack(oracle.upsample(g[0], g[1], g[2], (m, n, o)))
    _eq(up.cpu().numpy(), want_up, "device upsample")
    cm = torch.from_numpy(g.copy()).cuda()
    ctx.consistent_mapping(cm, 5)
    _eq(cm.cpu().numpy(), np.stack(oracle.consistent(*g, 5)), "device consistent mapping")
    f, mv = phantom.make_pair((m, n, o), 82)
    im = mv.astype(np.float32) + 1024
    im1b = f.astype(np.float32) + 1024
    X = np.eye(4, dtype=np.float32)
    X[0, 1], X[1, 3], X[2, 0], X[0, 3] = 0.02, 1.3, -0.015, -0.4
    Xc = X.T.ravel().copy()
    d_im, d_b = torch.from_numpy(im).cuda(), torch.from_numpy(im1b).cuda()
    w, ssd = ctx.warp_affine(d_im, Xc, up, im1b=d_b)
    wo, ssdo = oracle.warp_affine(im, im1b, Xc, *want_up)
    _eq(w.cpu().numpy(), wo, "device warpAffine")
    np.testing.assert_allclose(ssd, ssdo, rtol=1e-4)
    w2, _ = ctx.warp_image(d_im, up)
    _eq(w2.cpu().numpy(), oracle.warp_image(im, None, *want_up)[0], "device warpImageCL")
    seg = ((f > 60).astype(np.int16) + (f > 160)).astype(np.int16)
    ws = ctx.warp_affine_s(torch.from_numpy(seg).cuda(), Xc, up)
    _eq(ws.cpu().numpy(), oracle.warp_affine_s(seg, Xc, *want_up), "device warpAffineS")
    s, neg = ctx.jacobian(torch.from_numpy(g[:3].copy()).cuda(), 4)
    s0, neg0 = oracle.jacobian(g[0], g[1], g[2], 4)
    assert abs(s - float(s0)) <= 1e-4 * max(1.0, float(s0)) and abs(neg - float(neg0)) < 1e-6
    # applyBCVfloat / applyBCV on device data: the grid field upsampled and warped in one call
    step = 5
    flow = (rng.normal(size=3 * (m // step) * (n // step) * (o // step)) * 2).astype(np.float32)
    d_flow = torch.from_numpy(flow).cuda()
    _eq(ctx.apply_float(d_im, d_flow, step, Xc).cpu().numpy(), api.apply_float(im, flow, step, Xc),
        "device applyBCVfloat vs host form")
    sz1 = flow.size // 3
    fu = [flow[i * sz1:(i + 1) * sz1].reshape(o // step, n // step, m // step) for i in range(3)]
    full = oracle.upsample(fu[0], fu[1], fu[2], (m, n, o))
    _eq(ctx.apply_float(d_im, d_flow, step, Xc).cpu().numpy(), oracle.warp_affine(im, im1b, Xc, *full)[0],
        "device applyBCVfloat vs oracle")
    _eq(ctx.apply_seg(torch.from_numpy(seg).cuda(), d_flow, step, Xc).cpu().numpy(),
        oracle.warp_affine_s(seg, Xc, *full), "device applyBCV vs oracle")
    torch.cuda.synchronize()


def test_context_rejects_bad_tensors():
    """The Python mirror checks dtype, device, contiguity and size before any launch."""
    torch = pytest.importorskip("torch")
    ctx = api.Context(0)
    f = torch.zeros((20, 24, 28), device="cuda")
    flow = torch.zeros(3 * 5 * 6 * 7, device="cuda")
    with pytest.raises(ValueError):
        ctx.register(f.to(torch.int16), f.to(torch.int16), flow, grid=(4,), radius=(1,), quant=(1,))
    with pytest.raises(ValueError):
        ctx.register(f, f.transpose(0, 1).contiguous().transpose(0, 1), flow, grid=(4,), radius=(1,), quant=(1,))
    with pytest.raises(ValueError):
        ctx.register(f, f, flow[:10], grid=(4,), radius=(1,), quant=(1,))
    with pytest.raises(ValueError):
        ctx.descriptor(f.cpu(), 1)


@pytest.mark.parametrize("case", [(8, 8, 5), (7, 7, 4), (6, 6, 3), (5, 5, 2), (4, 4, 1)])
def test_cost_every_tile_shape_vs_oracle(oracle, case):
    """The autotuner's candidates (label planes per workgroup; pipelined or
    plain inner loop; sample rows kept from plane to plane or restaged), each
    pinned in turn, give the oracle's cost volume."""
    from deedsbcv_amd import _lib
    step, hw, q = case
    print(f"case: datacost tile shapes {case}", flush=True)
    f, mv = phantom.make_pair((96, 40, 36), 33)
    F = oracle.descriptor(f.astype(np.float32) + 1024, 2)
    M = oracle.descriptor(mv.astype(np.float32) + 1024, 2)
    want = oracle.datacost(F, M, step, hw, q, 1.6, 1)
    len_ = 2 * hw + 1
    lib = _lib.load()
    try:
        for planes in ((len_ + 1) // 2, len_, 2 * len_, 3 * len_, len_ * len_):
            for pipe in (1, 0, 3, 2):
                assert lib.deeds_set_datacost_tiles(planes, pipe) == 0
                _eq(api.dataCostCL(F, M, step, hw, q, 1.6, 1), want, f"cost {case} planes={planes} pipe={pipe}")
    finally:
        lib.deeds_set_datacost_tiles(0, 1)
    assert api.tuning_report()["datacost_rows"] is not None
