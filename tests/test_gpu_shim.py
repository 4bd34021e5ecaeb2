"""The C++ mirror of shmem/tree.h (include/gz_shmem.hpp) run as a C++ program: build, the walks,
the potential, the direct sum, the tidal tensor and the error paths against the oracle, in
tests/cpp/test_shmem_shim.cc (tolerances stated there)."""
from __future__ import annotations

import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "test_shmem_shim")


def test_shmem_shim_program():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    assert os.path.exists(BIN), "tests/cpp/test_shmem_shim is not built (run build())"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=110)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "shim ok" in r.stdout
