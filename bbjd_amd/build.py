"""Build step: compiles libbbjd.so (HIP kernels for gfx950 + C host driver)
and the oracle's C restatement (test infrastructure) with make."""
from __future__ import annotations

import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _make(path, jobs=8, force=True):
    subprocess.run(["make", "-C", path, f"-j{jobs}"] + (["-B"] if force else []), check=True)


def build(jobs=8, oracle=True, force=True):
    """force: recompile every source (make -B), also when the built
    libraries are up to date"""
    _make(os.path.join(ROOT, "bbjd_amd", "csrc"), jobs, force)
    if oracle:
        _make(os.path.join(ROOT, "oracle"), jobs, force)


if __name__ == "__main__":
    build()
