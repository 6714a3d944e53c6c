"""CPU tests of the oracle (the C restatement in oracle/) -- no GPU.

The oracle is pinned three ways:
  1. bit-for-bit against the golden vectors written by the unmodified
     reference (tests/golden/make_golden.py);
  2. bit-for-bit against the reference compiled from its own sources
     (oracle/_ref), on fresh seeded inputs, where that build exists;
  3. against what the mathematics fixes (brute-force window sums, popcount of
     thermometer codes, Kruskal's MST weight, brute-force min-convolution and
     brute-force MAP labelings on tiny trees), so that a dropped term, a wrong
     sign or a transposed axis fails even where 1 and 2 cannot see it.
"""
from pathlib import Path

import numpy as np
import pytest
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import minimum_spanning_tree

from deedsbcv_amd import phantom

# ---------------------------------------------------------------- golden ----


def test_golden_mind(oracle, golden):
    for qs in (1, 2):
        assert np.array_equal(oracle.descriptor(golden["im"], qs), golden[f"mind_q{qs}"])
    assert np.array_equal(oracle.descriptor(golden["im2"], 2), golden["mind2_q2"])


@pytest.mark.parametrize("case", [(5, 2, 2), (4, 2, 1), (6, 1, 3)])
def test_golden_cost_tree_bp(oracle, golden, case):
    step, hw, q = case
    tag = f"s{step}h{hw}q{q}"
    cost = oracle.datacost(golden["mind_q2"], golden["mind2_q2"], step, hw, q)
    assert np.array_equal(cost, golden[f"cost_{tag}"])
    tree = oracle.prims(golden["im2"] if step == 4 else golden["im"], step)
    for a, name in zip(tree, ("ordered", "parents", "edgemst")):
        assert np.array_equal(a, golden[f"{name}_{tag}"]), name
    u0 = golden[f"u0_{tag}"]
    u1 = oracle.regularisation(cost, u0[0], u0[1], u0[2], hw, step, q, *tree, (40, 36, 32))
    assert np.array_equal(np.stack(u1), golden[f"u1_{tag}"])


@pytest.mark.parametrize("ln", [3, 5, 9])
def test_golden_message_dt(oracle, golden, ln):
    off = golden[f"dt_off_{ln}"]
    msg, inds = oracle.message_dt(golden[f"dt_in_{ln}"], off[0], off[1], off[2])
    assert np.array_equal(msg, golden[f"dt_msg_{ln}"])
    assert np.array_equal(inds, golden[f"dt_inds_{ln}"])


def test_golden_transforms(oracle, golden):
    g = golden["grid6"]
    assert np.array_equal(np.stack(oracle.upsample(g[0], g[1], g[2], (40, 36, 32))), golden["up"])
    assert np.array_equal(np.stack(oracle.consistent(*g, 5)), golden["cons"])
    ux, vx, wx = golden["up"]
    w, ssd = oracle.warp_affine(golden["im2"], golden["im"], golden["X"], ux, vx, wx)
    assert np.array_equal(w, golden["warp_affine"])
    assert np.array_equal(ssd, golden["warp_affine_ssd"])
    w, ssd = oracle.warp_image(golden["im2"], golden["im"], ux, vx, wx)
    assert np.array_equal(w, golden["warp_image"])
    assert np.array_equal(ssd, golden["warp_image_ssd"])
    assert np.array_equal(oracle.warp_affine_s(golden["seg"], golden["X"], ux, vx, wx), golden["warp_seg"])


def test_golden_end_to_end(oracle, golden):
    s = golden["e2e_schedule"]
    flow, warped = oracle.register(golden["e2e_fixed"], golden["e2e_moving"], s[0], s[1], s[2])
    assert np.array_equal(flow, golden["e2e_flow"])
    assert np.array_equal(warped, golden["e2e_warped"])


# ------------------------------------------------- against the reference ----


@pytest.mark.parametrize("seed", [21, 22])
def test_ref_mind_cost_tree(oracle, ref, seed):
    f, mv = phantom.make_pair((37, 33, 29), seed)  # odd sizes: partial blocks at the border
    im, im2 = f.astype(np.float32) + 1024, mv.astype(np.float32) + 1024
    for qs in (1, 2, 3):
        assert np.array_equal(oracle.descriptor(im, qs), ref.descriptor(im, qs))
    F, M = oracle.descriptor(im, 2), oracle.descriptor(im2, 2)
    for step, hw, q in ((5, 2, 2), (4, 3, 1), (7, 1, 4), (8, 1, 5)):
        assert np.array_equal(oracle.datacost(F, M, step, hw, q), ref.datacost(F, M, step, hw, q))
    for step in (3, 4, 5, 6):
        for a, b in zip(oracle.prims(im, step), ref.prims(im, step)):
            assert np.array_equal(a, b)


def test_ref_randnum_variants(oracle, ref):
    f, mv = phantom.make_pair((32, 32, 32), 3)
    F = oracle.descriptor(f.astype(np.float32) + 1024, 1)
    M = oracle.descriptor(mv.astype(np.float32) + 1024, 1)
    for randnum in (0, 2):
        for step in (4, 5, 8):
            assert np.array_equal(oracle.datacost(F, M, step, 1, 2, 1.6, randnum),
                                  ref.datacost(F, M, step, 1, 2, 1.6, randnum))


def test_ref_bp_random_trees(oracle, ref):
    rng = np.random.default_rng(7)
    f, _ = phantom.make_pair((30, 30, 25), 9)
    im = f.astype(np.float32) + 1024
    for step, hw, q in ((5, 2, 2), (3, 1, 1), (6, 3, 1)):
        tree = oracle.prims(im, step)
        m1, n1, o1 = 30 // step, 30 // step, 25 // step
        cost = (rng.random(size=(m1 * n1 * o1, (2 * hw + 1) ** 3)) * 20).astype(np.float32)
        u0 = (rng.normal(size=(3, o1, n1, m1)) * 2).astype(np.float32)
        a = oracle.regularisation(cost, u0[0], u0[1], u0[2], hw, step, q, *tree, (30, 30, 25))
        b = ref.regularisation(cost, u0[0], u0[1], u0[2], hw, step, q, *tree, (30, 30, 25))
        for x, y in zip(a, b):
            assert np.array_equal(x, y)


def test_ref_jacobian(oracle, ref):
    rng = np.random.default_rng(3)
    g = (rng.normal(size=(3, 9, 10, 11)) * 2).astype(np.float32)
    assert oracle.jacobian(g[0], g[1], g[2], 4)[0] == ref.jacobian(g[0], g[1], g[2], 4)


# ------------------------------------------------ mathematical pins ---------


def _window_sum(x, hw):
    """Brute-force clamped box sum along all three axes (float64)."""
    out = x.astype(np.float64)
    for ax in range(3):
        L = out.shape[ax]
        acc = np.zeros_like(out)
        for t in range(L):
            lo, hi = max(t - hw, 0), min(t + hw, L - 1)
            sl = [slice(None)] * 3
            sl[ax] = slice(lo, hi + 1)
            tgt = [slice(None)] * 3
            tgt[ax] = t
            acc[tuple(tgt)] = out[tuple(sl)].sum(axis=ax)
        out = acc
    return out


def test_boxfilter_is_window_sum(oracle):
    rng = np.random.default_rng(0)
    x = rng.integers(0, 50, size=(9, 11, 13)).astype(np.float32)  # small integers: float sums exact
    for hw in (1, 2, 3):
        assert np.array_equal(oracle.boxfilter(x, hw), _window_sum(x, hw).astype(np.float32))


def test_mind_code_structure(oracle):
    f, _ = phantom.make_pair((24, 22, 20), 4)
    codes = oracle.descriptor(f.astype(np.float32) + 1024, 2)
    assert not np.any(codes >> np.uint64(60))
    for l in range(12):
        field = (codes >> np.uint64(5 * l)) & np.uint64(31)
        assert set(np.unique(field).tolist()) <= {0, 1, 3, 7, 15, 31}
    # the smallest of the 12 normalised values is 0 -> its field is full (31)
    full = np.zeros(codes.shape, bool)
    for l in range(12):
        full |= ((codes >> np.uint64(5 * l)) & np.uint64(31)) == 31
    assert full.all()


def test_mind_thresholds(oracle):
    """-log((c+1.5f)/6): float argument, log correctly rounded to float."""
    import math
    t = oracle.mind_thresholds()
    args = [np.float32(np.float32(c + 1.5) / np.float32(6)) for c in range(5)]
    want = np.array([np.float32(-math.log(float(a))) for a in args], np.float32)
    assert np.array_equal(t, want)


def test_datacost_brute_force(oracle):
    """popcount of two thermometer codes == sum over fields of |level difference|."""
    rng = np.random.default_rng(1)
    m, n, o = 12, 11, 10
    levels_f = rng.integers(0, 6, size=(o, n, m, 12))
    levels_m = rng.integers(0, 6, size=(o, n, m, 12))
    pack = lambda L: sum(((np.uint64(1) << L[..., l].astype(np.uint64)) - np.uint64(1)) << np.uint64(5 * l)
                         for l in range(12)).astype(np.uint64)
    F, M = pack(levels_f), pack(levels_m)
    step, hw, q = 3, 1, 2
    cost = oracle.datacost(F, M, step, hw, q, 1.6, 1)
    scale, skip = oracle.datacost_scale(step, q, 1.6, 1)
    len_ = 2 * hw + 1
    for node in range(cost.shape[0]):
        y, x, z = node % (m // step), (node // (m // step)) % (n // step), node // ((m // step) * (n // step))
        for l in range(len_ ** 3):
            ls, lx, lz = l % len_, (l // len_) % len_, l // len_ ** 2
            s = 0
            for k in range(0, step, skip[2]):
                for j in range(0, step, skip[1]):
                    for i in range(0, step, skip[0]):
                        fi, fj, fk = y * step + i, x * step + j, z * step + k
                        mi = min(max(fi + (ls - hw) * q, 0), m - 1)
                        mj = min(max(fj + (lx - hw) * q, 0), n - 1)
                        mk = min(max(fk + (lz - hw) * q, 0), o - 1)
                        s += int(np.abs(levels_f[fk, fj, fi] - levels_m[mk, mj, mi]).sum())
            assert cost[node, l] == np.float32(s) * scale


def test_prims_is_a_minimum_spanning_tree(oracle):
    f, _ = phantom.make_pair((40, 36, 32), 5)
    im = f.astype(np.float32) + 1024
    step = 4
    m1, n1, o1 = 10, 9, 8
    N = m1 * n1 * o1
    ordered, parents, edgemst = oracle.prims(im, step)
    ec, nb = oracle.edge_sad(im, step)
    _, sd = oracle.image_stats(im)
    w = -np.exp(-(ec / np.float32(step ** 3)).astype(np.float64) / (2.0 * float(sd)))
    # spanning tree: root first, every other node's parent precedes it
    root = m1 // 2 + (n1 // 2) * m1 + (o1 // 2) * m1 * n1
    assert ordered[0] == root and parents[root] == -1
    pos = np.empty(N, int)
    pos[ordered] = np.arange(N)
    assert sorted(ordered.tolist()) == list(range(N))
    assert all(pos[parents[c]] < pos[c] for c in range(N) if c != root)
    # BFS depth order, ties by vertex id
    depth = np.zeros(N, int)
    for c in ordered[1:]:
        depth[c] = depth[parents[c]] + 1
    keys = list(zip(depth[ordered], ordered))
    assert keys == sorted(keys)
    # total weight equals Kruskal's (shift weights positive for scipy)
    rows, cols, vals = [], [], []
    for e in range(6 * N):
        if nb[e] >= 0:
            rows.append(e % N), cols.append(nb[e]), vals.append(w[e] + 2.0)
    G = coo_matrix((vals, (rows, cols)), shape=(N, N)).tocsr()
    kruskal = minimum_spanning_tree(G).sum() - 2.0 * (N - 1)
    tree_w = 0.0
    for c in range(N):
        if c == root:
            continue
        p = parents[c]
        e = [k for k in range(6) if nb[c + k * N] == p]
        assert len(e) == 1
        tree_w += w[c + e[0] * N]
        # float expf vs double exp: equal to within one float ulp
        assert abs(float(edgemst[c]) + w[c + e[0] * N]) <= 1.2e-7 * abs(w[c + e[0] * N])
    assert abs(tree_w - kruskal) < 1e-6 * N


def test_message_dt_brute_force(oracle):
    rng = np.random.default_rng(2)
    for ln in (3, 5):
        cube = (rng.normal(size=ln ** 3) * 3).astype(np.float32)
        ox, oy, oz = rng.uniform(-1.5, 1.5, size=3).astype(np.float32)
        msg, inds = oracle.message_dt(cube, ox, oy, oz)
        v = cube.reshape(ln, ln, ln)  # [k, j, i]
        t = np.arange(ln)
        zy = ((t[:, None] - t[None, :]).astype(np.float32) + oy) ** 2  # [i_out, i_in]
        zx = ((t[:, None] - t[None, :]).astype(np.float32) + ox) ** 2
        zz = ((t[:, None] - t[None, :]).astype(np.float32) + oz) ** 2
        for l in range(ln ** 3):
            io, jo, ko = l % ln, (l // ln) % ln, l // ln ** 2
            # same association as the separable passes: ((v + zy) + zx) + zz
            a = (v + zy[io][None, None, :]).astype(np.float32)
            a = (a + zx[jo][None, :, None]).astype(np.float32)
            a = (a + zz[ko][:, None, None]).astype(np.float32)
            assert msg[l] == a.min()
            src = inds[l]
            assert a.reshape(-1)[src] == a.min()


def test_regularisation_is_map_on_chains(oracle):
    """On a 3-node star the BP labeling minimises the tree energy
    D_root(l_r) + sum_c [ w_c D_c(l_c) + |l_r - l_c + (u0_r - u0_c)/q|^2 ]
    (the child's cost is scaled by its edge weight before the distance
    transform, the pairwise term is not)."""
    rng = np.random.default_rng(4)
    hw, step, q = 1, 1, 1.0
    ln = 2 * hw + 1
    L = ln ** 3
    # a 3x1x1 grid; root = node 1 (m/2), children 0 and 2
    m, n, o = 3, 1, 1
    ordered = np.array([1, 0, 2], np.int32)
    parents = np.array([1, -1, 1], np.int32)
    for trial in range(5):
        cost = (rng.random(size=(3, L)) * 6).astype(np.float32)
        edgemst = np.array([rng.uniform(0.3, 1), 0, rng.uniform(0.3, 1)], np.float32)
        u0 = (rng.normal(size=(3, 1, 1, 3)) * 0.7).astype(np.float32)
        u1, v1, w1, inds, sel = oracle.regularisation(cost, u0[0], u0[1], u0[2], hw, step, q, ordered, parents,
                                                      edgemst, (m, n, o), want_inds=True)
        lab = np.arange(L)
        li, lj, lk = lab % ln, (lab // ln) % ln, lab // ln ** 2

        def pair(p, c):  # d(parent label, child label) in label units
            offx = (u0[0].ravel()[p] - u0[0].ravel()[c]) / q
            offy = (u0[1].ravel()[p] - u0[1].ravel()[c]) / q
            offz = (u0[2].ravel()[p] - u0[2].ravel()[c]) / q
            return ((li[:, None] - li[None, :] + offy) ** 2 + (lj[:, None] - lj[None, :] + offx) ** 2 +
                    (lk[:, None] - lk[None, :] + offz) ** 2)

        E = (cost[1][:, None, None] + (edgemst[0] * cost[0][None, :, None] + pair(1, 0)[:, :, None]) +
             (edgemst[2] * cost[2][None, None, :] + pair(1, 2)[:, None, :]))
        s = sel
        assert abs(E[s[1], s[0], s[2]] - E.min()) < 1e-4


# ------------------------------------------------------------ diceMulti -----

def _label_pair(shape, seed, n_labels=6, offset=0):
    """Two overlapping label volumes: random boxes of labels on background."""
    rng = np.random.default_rng(seed)
    a = np.full(shape, offset, np.int16)
    b = np.full(shape, offset, np.int16)
    for lab in range(1, n_labels + 1):
        for v in (a, b):
            lo = [rng.integers(0, s // 2) for s in shape]
            hi = [l + rng.integers(2, s // 2) for l, s in zip(lo, shape)]
            v[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] = lab + offset
    return a, b


def test_dice_closed_forms(oracle):
    a, b = _label_pair((20, 18, 16), 5)
    labels, d = oracle.dice(a, a)
    assert labels.tolist() == sorted(set(np.unique(a).tolist()) - {int(a.min())})
    assert np.all(d == 1.0)
    # disjoint label sets: every Dice is 0
    labels, d = oracle.dice(a, (a + 100).astype(np.int16))
    assert np.all(d == 0.0)
    # brute force with numpy counts on a small pair (labels of seg1 but the smallest)
    labels, d = oracle.dice(a, b)
    want = []
    for lab in np.unique(a)[1:]:
        inter = np.float32(np.count_nonzero((a == lab) & (b == lab)))
        c1, c2 = np.float32(np.count_nonzero(a == lab)), np.float32(np.count_nonzero(b == lab))
        want.append(np.float32(2.0) * inter / (c1 + c2))
    assert np.array_equal(d, np.array(want, np.float32))


def _dice_text(labels_dice):
    """The reference's two output lines (diceMulti.cpp:58-64) for (labels, dice)."""
    d = labels_dice[1]
    mean = np.float32(0.0)
    for x in d:
        mean = np.float32(mean + x)
    mean = np.float32(mean / np.float32(len(d)))
    return "Dice (for each label): " + "".join(" %4.3f, " % x for x in d) + "\nAVERAGE DICE: %f\n" % mean


@pytest.mark.parametrize("case", [(7, 6, 0), (8, 12, -3), (9, 3, 200)])
def test_ref_dice_binary(oracle, ref, tmp_path, case):
    """The oracle's Dice, printed as diceMulti prints it, equals the unmodified
    reference binary's output on the same files."""
    import subprocess
    from deedsbcv_amd import nifti
    seed, k, offset = case
    a, b = _label_pair((33, 29, 25), seed, k, offset)
    nifti.write_nifti(tmp_path / "a.nii.gz", a)
    nifti.write_nifti(tmp_path / "b.nii.gz", b)
    exe = Path(__file__).resolve().parents[1] / "oracle" / "_ref" / "diceMulti"
    if not exe.exists():
        pytest.skip("reference diceMulti not built")
    r = subprocess.run([str(exe), str(tmp_path / "a.nii.gz"), str(tmp_path / "b.nii.gz")], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 0
    assert r.stdout == _dice_text(oracle.dice(a, b))


# ------------------------------------------------------------- linearBCV ----

def _affine_X(seed, scale=0.03, shift=2.0):
    """A column-wise 4x4 affine near identity (the layout linearBCV keeps)."""
    rng = np.random.default_rng(seed)
    A = np.eye(4, dtype=np.float32)
    A[:3, :3] += rng.normal(0, scale, (3, 3)).astype(np.float32)
    A[:3, 3] = rng.normal(0, shift, 3).astype(np.float32)
    return A.T.ravel().copy()  # X[i + 4*j] = A[i, j]


def _points(seed, n, X, noise=0.3, outliers=0.2):
    """pts1 (4 x n, homogeneous) and pts2 = X applied to pts1 plus noise and outliers."""
    rng = np.random.default_rng(seed)
    p1 = np.ones((4, n), np.float32)
    p1[:3] = rng.uniform(0, 60, (3, n)).astype(np.float32)
    A = np.asarray(X, np.float32).reshape(4, 4).T
    p2 = (A @ p1).astype(np.float32)
    p2[:3] += rng.normal(0, noise, (3, n)).astype(np.float32)
    bad = rng.random(n) < outliers
    p2[:3, bad] += rng.uniform(-15, 15, (3, int(bad.sum()))).astype(np.float32)
    p2[3] = 1
    return p1.ravel().copy(), p2.ravel().copy()


def test_ref_lin_warp(oracle, ref_linear):
    f, _ = phantom.make_pair((37, 31, 27), 8)
    im = f.astype(np.float32) + 1024.5
    for seed in (1, 2, 3):
        X = _affine_X(seed)
        assert np.array_equal(oracle.lin_warp(im, X), ref_linear.lin_warp(im, X))


def test_ref_qrsolve_and_robust_fit(oracle, ref_linear):
    rng = np.random.default_rng(4)
    A = rng.normal(0, 1, (4, 4)).astype(np.float32).ravel()
    b = np.eye(4, dtype=np.float32).ravel()
    inv = oracle.qrsolve(A, b, 4, 4)
    assert np.array_equal(inv, ref_linear.qrsolve(A, b, 4, 4))
    # closed form: the QR solve of A X = I inverts A
    assert np.abs(A.reshape(4, 4).T @ inv.reshape(4, 4).T - np.eye(4)).max() < 1e-4
    for seed, n in ((5, 400), (6, 1001), (7, 57)):
        X = _affine_X(seed)
        p1, p2 = _points(seed, n, X)
        for rigid in (0, 1):
            got = oracle.affine_robust(p1, p2, n, rigid)
            assert np.array_equal(got, ref_linear.affine_robust(p1, p2, n, rigid)), (seed, rigid)
        # noise-free points: the trimmed fit recovers the affine map (to float QR
        # accuracy over coordinates up to 60; a transposed or dropped term is off by O(1))
        q1, q2 = _points(seed, n, X, noise=0.0, outliers=0.0)
        assert np.abs(oracle.affine_robust(q1, q2, n, 0) - np.asarray(X).reshape(4, 4).T.ravel()).max() < 1e-2


def test_ref_rigid_recovers_rotation(oracle):
    """Closed form: points rotated and translated exactly give that rigid map."""
    rng = np.random.default_rng(12)
    th = 0.1
    R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]], np.float64)
    p1 = np.ones((4, 300), np.float32)
    p1[:3] = rng.uniform(0, 50, (3, 300))
    p2 = p1.copy()
    p2[:3] = (R @ p1[:3].astype(np.float64) + np.array([[1.5], [-2.0], [0.5]])).astype(np.float32)
    RT = oracle.affine_robust(p1.ravel(), p2.ravel(), 300, 1).reshape(4, 4)
    assert np.abs(RT[:3, :3] - R).max() < 1e-3


def test_ref_estimate_affine(oracle, ref_linear):
    f, mv = phantom.make_pair((48, 44, 40), 13)
    im1, im2 = f.astype(np.float32) + 1024, mv.astype(np.float32) + 1024
    F, M = oracle.descriptor(im1, 2), oracle.descriptor(im2, 2)
    for step, hw, q in ((5, 3, 2), (4, 2, 1)):
        c1 = oracle.datacost(F, M, step, hw, q, 1.0)
        c2 = oracle.datacost(M, F, step, hw, q, 1.0)
        Xprev = _affine_X(step, 0.01, 0.5)
        for rigid in (0, 1):
            got = oracle.estimate_affine(Xprev, im1, im2, c1, c2, step, q, hw, rigid)
            want = ref_linear.estimate_affine(Xprev, im1, im2, c1, c2, step, q, hw, rigid)
            for a, b in zip(got, want):
                assert np.array_equal(a, b), (step, hw, q, rigid)


def _cpp_float(x):
    """C++ ostream default formatting of a float (6 significant digits, %g)."""
    return "%g" % np.float32(x)


def test_ref_linear_binary(oracle, tmp_path):
    """orc_linear_register reproduces the reference linearBCV's _matrix.txt."""
    import subprocess
    from deedsbcv_amd import nifti
    exe = Path(__file__).resolve().parents[1] / "oracle" / "_ref" / "linearBCV"
    if not exe.exists():
        pytest.skip("reference linearBCV not built")
    f, mv = phantom.make_pair((64, 56, 48), 17)
    nifti.write_nifti(tmp_path / "f.nii.gz", f)
    nifti.write_nifti(tmp_path / "m.nii.gz", mv)
    for rigid in (0, 1):
        args = [str(exe), "-F", str(tmp_path / "f.nii.gz"), "-M", str(tmp_path / "m.nii.gz"), "-O",
                str(tmp_path / f"o{rigid}"), "-G", "6x5x4", "-L", "3x2x2", "-Q", "2x2x1"]
        if rigid:
            args += ["-R", "1"]
        r = subprocess.run(args, capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stdout + r.stderr
        X = oracle.linear_register(f.astype(np.float32), mv.astype(np.float32), (6, 5, 4), (3, 2, 2), (2, 2, 1),
                                   rigid)
        want = "".join("  ".join(_cpp_float(X[i + 4 * c]) for c in range(4)) + "\n" for i in range(4))
        assert (tmp_path / f"o{rigid}_matrix.txt").read_text() == want


def test_oracle_register_small_workload_vs_ref_binary(oracle, tmp_path):
    """End to end on the 96x96x80 three-level workload (grid 6,5,4; radius 6,5,4;
    quantisation 3,2,1; alpha 1.6, symmetric): the oracle's flow and deformed
    image equal the unmodified reference deedsBCV binary's _displacements.dat
    and _deformed.nii.gz bit for bit (deedsBCV0.cpp:339-356)."""
    import refrun
    if not refrun.available():
        pytest.skip("reference deedsBCV not built")
    f, mv = phantom.workload_small(0)
    s = phantom.SCHEDULE_SMALL
    run = refrun.RefRun(f, mv, s, tmp_path)
    flow, warped = oracle.register(f, mv, s["grid"], s["radius"], s["quant"], s["alpha"])
    rflow, rwarped, _, _ = run.result()
    refrun.assert_same(flow, rflow, "flow vs reference _displacements.dat")
    refrun.assert_same(warped, rwarped, "warped vs reference _deformed.nii.gz")
