"""The C-ABI library: builds, loads without a GPU, exports every symbol the
public header declares, and fails loudly (no CPU fallback) without a GPU."""
import os
import re
import subprocess

import pytest

from tests.conftest import ROOT

HEADER = os.path.join(ROOT, "include", "rockable_amd.h")


def declared_symbols():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\*)\s+(rk_\w+)\s*\(", txt, re.M)))


def test_header_declares_the_abi():
    syms = declared_symbols()
    assert "rk_step" in syms and "rk_load_conf" in syms and len(syms) >= 18


def test_library_exports_every_declared_symbol():
    from rockable_amd.engine import EXPORTED, LIB_PATH, load_library
    assert os.path.exists(LIB_PATH), "build first: python -m rockable_amd.build"
    out = subprocess.run(["nm", "-D", "--defined-only", LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r" T (rk_\w+)$", out, re.M))
    missing = [s for s in declared_symbols() if s not in exported]
    assert not missing, missing
    assert set(EXPORTED) == set(declared_symbols())
    load_library()


def test_no_cpu_fallback_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from rockable_amd.engine import Engine
    with pytest.raises(RuntimeError):
        Engine(None)


def test_product_never_links_the_oracle():
    """The product path must not import, link or call the oracle."""
    for dirpath, _, files in os.walk(os.path.join(ROOT, "rockable_amd")):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".hpp", ".h")):
                txt = open(os.path.join(dirpath, f)).read()
                assert "rk_oracle" not in txt and "from oracle" not in txt and "import oracle" not in txt, f
    out = subprocess.run(["ldd", os.path.join(ROOT, "rockable_amd", "librockable_amd.so")], capture_output=True,
                         text=True).stdout
    assert "rk_oracle" not in out


def test_null_engine_is_rejected_without_a_device_call():
    """Every entry point that takes an engine returns -1 for a null handle,
    before any HIP call (so this runs without a GPU)."""
    import ctypes
    from rockable_amd.engine import load_library
    L = load_library()
    assert L.rk_last_error(None) == b"null engine"
    assert L.rk_load_conf(None, b"x") == -1
    assert L.rk_load_driving_system(None, b"x") == -1
    assert L.rk_set_force_law(None, b"Avalanche") == -1
    assert L.rk_set_option(None, b"fuse_law", 1.0) == -1
    assert L.rk_update_nl(None) == -1
    assert L.rk_accelerations(None) == -1
    assert L.rk_step(None, 1) == -1
    assert L.rk_integrate(None, None, 1, None) == -1
    assert L.rk_save_conf(None, None) == -1
    assert L.rk_synchronize(None) == -1
    assert L.rk_get_counts(None, None) == -1
    assert L.rk_reset_kernel_time(None) == -1
    v = ctypes.c_double()
    assert L.rk_get_option(None, b"fuse_law", ctypes.byref(v)) == -1


def test_cxx_caller_builds_and_fails_cleanly_without_gpu(tmp_path):
    """The C++ example caller links against the library through the header
    alone; without a device it reports the failure and exits with status 1
    (no abort, no CPU fallback)."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    exe = os.path.join(ROOT, "rockable_amd", "rk_run")
    assert os.path.exists(exe), "build first: python -m rockable_amd.build"
    r = subprocess.run([exe, "nothing.conf", str(tmp_path)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "no usable device" in r.stderr
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr
