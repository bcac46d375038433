"""libmbgpu.so loads and exports every entry point include/mbgpu.h declares (no GPU needed)."""
import ctypes
import os
import re

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "mbgpu.h")
LIB = os.path.join(REPO, "dlpoly2-mbpol-heat_amd", "libmbgpu.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b([a-z_][a-z0-9_]*)\s*\(", text, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "return")))


def test_header_declares_the_reference_leaves():
    names = declared_functions()
    for n in ("mbpol_2b_poly_", "mbpol_3b_poly_", "mbpol_2b_cutoff_", "mbpol_3b_cutoff_",
              "mbgpu_mbpol_forces", "mbgpu_heat_flux"):
        assert n in names


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "build first: __graft_entry__.build()"
    lib = ctypes.CDLL(LIB)
    for n in declared_functions():
        assert hasattr(lib, n), f"{n} not exported"


def test_python_binding_covers_the_header(pkg):
    from importlib import import_module
    mb = import_module("dlpoly2-mbpol-heat_amd.mbgpu")
    assert set(declared_functions()) == set(mb._SIGS)


def test_cutoffs_without_gpu(pkg):
    # the cutoff leaves touch no device
    assert pkg.mbpol_2b_cutoff() == 6.5
    assert pkg.mbpol_3b_cutoff() == 4.5


def test_create_without_gpu_fails_cleanly(pkg):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(pkg.MbgpuError):
        pkg.Context(64, 256)
