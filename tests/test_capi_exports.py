"""The C-ABI library loads and exports every function include/*.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "lsdtt3_amd", "lib", "liblsdtt3_gpu.so")


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(lsdtt_gpu_\w+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    assert len(declared_functions()) >= 10


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_every_declared_symbol_is_exported():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_package_fails_loudly_without_library(monkeypatch, tmp_path):
    import importlib
    import lsdtt3_amd._native as nat
    monkeypatch.setattr(nat, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(nat, "_lib", None)
    with pytest.raises(ImportError):
        nat.lib()
    importlib.reload(nat)
