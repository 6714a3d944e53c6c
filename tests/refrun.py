"""Runs the unmodified reference deedsBCV binary (oracle/_ref/deedsBCV, built
from the reference sources by `make -C oracle ref`) on a pair of volumes and
reads back its _displacements.dat and _deformed.nii.gz.  Test infrastructure
only: the reference is the checker here, never the thing measured.

The run can start in the background (the reference's OpenMP threads on the
host cores) while the caller does its GPU work, and be joined later."""
from __future__ import annotations

import os
import subprocess
from pathlib import Path

import numpy as np

from deedsbcv_amd import nifti

ROOT = Path(__file__).resolve().parents[1]
REF_DEEDS = ROOT / "oracle" / "_ref" / "deedsBCV"


def host_threads():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


class RefRun:
    def __init__(self, fixed, moving, sched, workdir, seg=None, threads=None, tag="o"):
        workdir = Path(workdir)
        workdir.mkdir(parents=True, exist_ok=True)
        self.dir, self.tag = workdir, tag
        nifti.write_nifti(workdir / f"{tag}_f.nii.gz", fixed)
        nifti.write_nifti(workdir / f"{tag}_m.nii.gz", moving)
        cmd = [str(REF_DEEDS), "-F", str(workdir / f"{tag}_f.nii.gz"), "-M", str(workdir / f"{tag}_m.nii.gz"),
               "-O", str(workdir / tag), "-G", "x".join(map(str, sched["grid"])),
               "-L", "x".join(map(str, sched["radius"])), "-Q", "x".join(map(str, sched["quant"])),
               "-a", str(sched["alpha"])]
        if seg is not None:
            nifti.write_nifti(workdir / f"{tag}_s.nii.gz", seg)
            cmd += ["-S", str(workdir / f"{tag}_s.nii.gz")]
        env = dict(os.environ, OMP_NUM_THREADS=str(threads or host_threads()))
        self.proc = subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, env=env)

    def result(self, timeout=600):
        """(flow, deformed volume, deformed segmentation or None, stdout) once the run has ended."""
        out, _ = self.proc.communicate(timeout=timeout)
        if self.proc.returncode != 0:
            raise RuntimeError(f"reference deedsBCV failed ({self.proc.returncode}):\n{out[-2000:]}")
        d, t = self.dir, self.tag
        flow = nifti.read_displacements(d / f"{t}_displacements.dat")
        warped = nifti.read_nifti(d / f"{t}_deformed.nii.gz")[0]
        segp = d / f"{t}_deformed_seg.nii.gz"
        seg = nifti.read_nifti(segp)[0] if segp.exists() else None
        return flow, warped, seg, out


def available():
    return REF_DEEDS.exists()


def assert_same(a, b, what):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, (what, a.shape, b.shape)
    if not np.array_equal(a, b):
        bad = np.flatnonzero(a.ravel() != b.ravel())
        raise AssertionError(f"{what}: {bad.size} of {a.size} differ; first at {bad[:5]}: "
                             f"{a.ravel()[bad[:5]]} vs {b.ravel()[bad[:5]]}; max |diff| "
                             f"{np.abs(a.astype(np.float64) - b.astype(np.float64)).max()}")
