"""The C ABI library loads without a GPU and exports every function include/gz_grav.h declares."""
from __future__ import annotations

import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names.update(re.findall(r"^\s*(?:const\s+)?[a-z_]+\s*\**\s*(gz_[a-z0-9_]+)\s*\(", text,
                                re.M))
    return sorted(names)


def test_header_declares_the_surface():
    names = _declared()
    for must in ("gz_tree_build", "gz_tree_accel_grouped", "gz_tree_accel", "gz_direct_sum",
                 "gz_order_targets", "gz_kick"):
        assert must in names


def _param_counts():
    out = {}
    for h in os.listdir(os.path.join(ROOT, "include")):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"(gz_[a-z0-9_]+)\s*\(([^)]*)\)\s*;", text):
            params = m.group(2).strip()
            out[m.group(1)] = 0 if params in ("", "void") else params.count(",") + 1
    return out


def test_binding_prototypes_match_header_arity():
    from gizmo_shared_memory_amd import _lib
    counts = _param_counts()
    for name, (_, args) in _lib._PROTOS.items():
        assert counts[name] == len(args), (name, counts[name], len(args))


def test_library_exports_every_declared_symbol():
    from gizmo_shared_memory_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libgzgrav.so is not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding's prototypes cover the same surface
    assert set(_lib.exported_names()) == set(_declared())
    bound = _lib.load()
    assert bound.gz_status_string(0) == b"GZ_OK"


def test_cpp_mirror_of_the_shmem_calls_compiles_and_links():
    """include/gz_shmem.hpp (shmem/tree.h's build and walk calls in C++ over the C ABI) builds
    with g++ against the library; its behaviour is tested on the GPU (tests/test_gpu_shim.py)."""
    import shutil
    import subprocess
    if not shutil.which("g++") or not os.path.isdir("/opt/rocm/include"):
        pytest.skip("g++ or the ROCm headers are missing")
    src = os.path.join(ROOT, "tests", "cpp", "test_shmem_shim.cc")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-Werror",
                        "-D__HIP_PLATFORM_AMD__", "-I", os.path.join(ROOT, "include"), "-I",
                        os.path.join(ROOT, "oracle"), "-I", "/opt/rocm/include", src],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert os.path.exists(os.path.join(ROOT, "tests", "cpp", "test_shmem_shim")), \
        "tests/cpp/test_shmem_shim is not built (run __graft_entry__.build())"
