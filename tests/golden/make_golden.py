"""Writes tests/golden/golden_small.npz from the UNMODIFIED reference.

Runs the reference's own functions (oracle/_ref/libdeedsref.so, built by
`make -C oracle ref` from the reference sources where they lie) and its
deedsBCV / applyBCV binaries on small seeded inputs, and stores inputs and
outputs as a fixture.  The fixture holds data only; the tests compare the C
restatement (oracle/) and the HIP library against it.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(HERE.parent))

from cbind import load_ref  # noqa: E402
from deedsbcv_amd import nifti, phantom  # noqa: E402

# (step, hw, quant) data-cost / regularisation cases
COST_CASES = [(5, 2, 2), (4, 2, 1), (6, 1, 3)]
E2E_SCHEDULE = dict(grid=(6, 5, 4), radius=(3, 2, 2), quant=(2, 2, 1))


def main():
    R = load_ref()
    if R is None:
        raise SystemExit("oracle/_ref/libdeedsref.so is missing: run `make -C oracle ref` first")
    out = {}
    f, mv = phantom.make_pair((40, 36, 32), 11)
    im = f.astype(np.float32) + 1024
    im2 = mv.astype(np.float32) + 1024
    out["im"], out["im2"] = im, im2
    for qs in (1, 2):
        out[f"mind_q{qs}"] = R.descriptor(im, qs)
    out["mind2_q2"] = R.descriptor(im2, 2)
    F, M = out["mind_q2"], out["mind2_q2"]
    rng = np.random.default_rng(5)
    for step, hw, q in COST_CASES:
        tag = f"s{step}h{hw}q{q}"
        cost = R.datacost(F, M, step, hw, q)
        out[f"cost_{tag}"] = cost
        ordered, parents, edgemst = R.prims(im2 if step == 4 else im, step)
        out[f"ordered_{tag}"], out[f"parents_{tag}"], out[f"edgemst_{tag}"] = ordered, parents, edgemst
        m1, n1, o1 = 40 // step, 36 // step, 32 // step
        u0 = (rng.normal(size=(3, o1, n1, m1)) * 1.5).astype(np.float32)
        out[f"u0_{tag}"] = u0
        u1, v1, w1 = R.regularisation(cost, u0[0], u0[1], u0[2], hw, step, q, ordered, parents, edgemst, (40, 36, 32))
        out[f"u1_{tag}"] = np.stack([u1, v1, w1])
    for ln in (3, 5, 9):
        cube = (rng.normal(size=ln ** 3) * 4).astype(np.float32)
        off = rng.uniform(-2, 2, size=3).astype(np.float32)
        msg, inds = R.message_dt(cube, off[0], off[1], off[2])
        out[f"dt_in_{ln}"], out[f"dt_off_{ln}"], out[f"dt_msg_{ln}"], out[f"dt_inds_{ln}"] = cube, off, msg, inds
    # transforms
    g = (rng.normal(size=(6, 6, 7, 8)) * 2.5).astype(np.float32)  # six fields on an 8x7x6 grid
    out["grid6"] = g
    out["up"] = np.stack(R.upsample(g[0], g[1], g[2], (40, 36, 32)))
    out["cons"] = np.stack(R.consistent(*g, 5))
    X = np.eye(4, dtype=np.float32)
    X[0, 1], X[1, 3], X[2, 0], X[0, 3] = 0.03, 1.7, -0.02, -0.6
    Xc = X.T.ravel().copy()
    out["X"] = Xc
    ux, vx, wx = out["up"]
    out["warp_affine"], out["warp_affine_ssd"] = R.warp_affine(im2, im, Xc, ux, vx, wx)
    out["warp_image"], out["warp_image_ssd"] = R.warp_image(im2, im, ux, vx, wx)
    seg = ((f > 60).astype(np.int16) + (f > 160) + (f > 240)).astype(np.int16)
    out["seg"] = seg
    out["warp_seg"] = R.warp_affine_s(seg, Xc, ux, vx, wx)
    # end to end through the reference's command line
    fb, mb = phantom.make_pair((48, 40, 36), 12)
    out["e2e_fixed"], out["e2e_moving"] = fb, mb
    segb = ((fb > 60).astype(np.int16) + (fb > 160)).astype(np.int16)
    out["e2e_seg"] = segb
    with tempfile.TemporaryDirectory() as d:
        nifti.write_nifti(f"{d}/f.nii.gz", fb)
        nifti.write_nifti(f"{d}/m.nii.gz", mb)
        nifti.write_nifti(f"{d}/s.nii.gz", segb)
        s = E2E_SCHEDULE
        args = [str(ROOT / "oracle/_ref/deedsBCV"), "-F", f"{d}/f.nii.gz", "-M", f"{d}/m.nii.gz", "-O", f"{d}/o",
                "-G", "x".join(map(str, s["grid"])), "-L", "x".join(map(str, s["radius"])),
                "-Q", "x".join(map(str, s["quant"])), "-S", f"{d}/s.nii.gz"]
        subprocess.run(args, check=True, capture_output=True)
        out["e2e_flow"] = nifti.read_displacements(f"{d}/o_displacements.dat")
        out["e2e_warped"] = nifti.read_nifti(f"{d}/o_deformed.nii.gz")[0]
        out["e2e_warped_seg"] = nifti.read_nifti(f"{d}/o_deformed_seg.nii.gz")[0]
        import gzip
        for name in ("deformed", "deformed_seg"):
            with gzip.open(f"{d}/o_{name}.nii.gz", "rb") as fh:
                out[f"e2e_file_{name}"] = np.frombuffer(fh.read(), np.uint8).copy()
        subprocess.run([str(ROOT / "oracle/_ref/applyBCV"), "-M", f"{d}/s.nii.gz", "-O", f"{d}/o", "-D",
                        f"{d}/a.nii.gz"], check=True, capture_output=True)
        out["e2e_apply_seg"] = nifti.read_nifti(f"{d}/a.nii.gz")[0]
    out["e2e_schedule"] = np.array([s["grid"], s["radius"], s["quant"]], np.int32)
    np.savez_compressed(HERE / "golden_small.npz", **out)
    print("wrote", HERE / "golden_small.npz", sum(v.nbytes for v in out.values()), "raw bytes")


if __name__ == "__main__":
    main()
