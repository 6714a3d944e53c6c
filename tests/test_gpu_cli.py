"""GPU tests of the drop-in command lines (deedsBCV, applyBCV, applyBCVfloat):
same flags and output files as the reference binaries, byte-identical
decompressed outputs on the golden pair."""
import gzip
import subprocess
from pathlib import Path

import numpy as np
import pytest

from deedsbcv_amd import nifti

pytestmark = pytest.mark.gpu
BIN = Path(__file__).resolve().parents[1] / "deedsbcv_amd" / "bin"


def _run(args):
    r = subprocess.run([str(a) for a in args], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def test_deedsBCV_cli_matches_reference_files(golden, tmp_path):
    print("case: deedsBCV cli golden", flush=True)
    if not (BIN / "deedsBCV").exists():
        pytest.skip("tools not built")
    nifti.write_nifti(tmp_path / "f.nii.gz", golden["e2e_fixed"])
    nifti.write_nifti(tmp_path / "m.nii.gz", golden["e2e_moving"])
    nifti.write_nifti(tmp_path / "s.nii.gz", golden["e2e_seg"])
    s = golden["e2e_schedule"]
    out = _run([BIN / "deedsBCV", "-F", tmp_path / "f.nii.gz", "-M", tmp_path / "m.nii.gz", "-O", tmp_path / "o",
                "-G", "x".join(map(str, s[0])), "-L", "x".join(map(str, s[1])), "-Q", "x".join(map(str, s[2])),
                "-S", tmp_path / "s.nii.gz"])
    assert "Finished. Total time" in out
    flow = np.fromfile(tmp_path / "o_displacements.dat", np.float32)
    assert np.array_equal(flow, golden["e2e_flow"])
    for name in ("deformed", "deformed_seg"):
        with gzip.open(tmp_path / f"o_{name}.nii.gz", "rb") as fh:
            got = np.frombuffer(fh.read(), np.uint8)
        assert np.array_equal(got, golden[f"e2e_file_{name}"]), name
    # applyBCV on the same field reproduces the reference's applyBCV output
    _run([BIN / "applyBCV", "-M", tmp_path / "s.nii.gz", "-O", tmp_path / "o", "-D", tmp_path / "a.nii.gz"])
    assert np.array_equal(nifti.read_nifti(tmp_path / "a.nii.gz")[0], golden["e2e_apply_seg"])
    # applyBCVfloat on the moving image reproduces the deformed image up to the
    # +-1024 offset round trip of deedsBCV0.cpp:139-144 / :352-354
    _run([BIN / "applyBCVfloat", "-M", tmp_path / "m.nii.gz", "-O", tmp_path / "o", "-D", tmp_path / "b.nii.gz"])
    b = nifti.read_nifti(tmp_path / "b.nii.gz")[0]
    assert np.abs(b - golden["e2e_warped"]).max() < 2e-3


def test_cli_error_paths(tmp_path):
    if not (BIN / "deedsBCV").exists():
        pytest.skip("tools not built")
    r = subprocess.run([str(BIN / "deedsBCV"), "-F", "a.nii", "-M", "b.nii", "-O", "x"], capture_output=True,
                       text=True)
    assert r.returncode != 0 and "nii.gz" in r.stdout
    r = subprocess.run([str(BIN / "deedsBCV"), "-F", str(tmp_path / "missing.nii.gz"), "-M",
                        str(tmp_path / "missing2.nii.gz"), "-O", str(tmp_path / "x")], capture_output=True, text=True)
    assert r.returncode != 0 and "gzopen" in r.stdout
