"""The C ABI library loads and exports every symbol include/hipsw.h declares
(no compute call: this runs without a GPU)."""
import ctypes
import os
import re

from swiftbwa_amd import native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "hipsw.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(hipsw_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_header():
    lib = ctypes.CDLL(native.LIB_PATH)
    syms = header_symbols()
    assert len(syms) >= 10
    for s in syms:
        assert hasattr(lib, s), s


def test_binding_covers_header():
    assert set(native.declared_symbols()) == set(header_symbols())


def test_load_binds_signatures():
    lib = native.load()
    assert lib.hipsw_min_batch_size() == 32
    # engine creation without a device returns NULL (the "no GPU" signal)
    import subprocess, sys
    code = ("import sys; sys.path.insert(0, %r); from swiftbwa_amd import native; "
            "import os; lib = native.load(); print(bool(lib.hipsw_engine_create(99)))" % ROOT)
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert out.stdout.strip() == "False"
