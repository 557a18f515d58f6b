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
