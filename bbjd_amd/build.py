"""Build step: compiles libbbjd.so (HIP kernels for gfx950 + C host driver)
and the oracle's C restatement (test infrastructure) with make."""
from __future__ import annotations

import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _make(path, jobs=8):
    subprocess.run(["make", "-C", path, f"-j{jobs}"], check=True)


def build(jobs=8, oracle=True):
    _make(os.path.join(ROOT, "bbjd_amd", "csrc"), jobs)
    if oracle:
        _make(os.path.join(ROOT, "oracle"), jobs)


if __name__ == "__main__":
    build()
