"""The C-ABI library loads and exports every symbol include/hbp.h declares (no GPU calls)."""
import os
import re
import subprocess

import pytest

from tnbp_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "hbp.h")).read()
    return sorted(set(re.findall(r"\b(hbp_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = _declared()
    assert "hbp_contract" in names and "hbp_peps_bp_run" in names
    assert len(names) >= 40


@pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason="libhbp.so not built")
def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (hbp_[a-z_0-9]+)", out))
    missing = [n for n in _declared() if n not in exported]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason="libhbp.so not built")
def test_ctypes_binding_covers_header():
    lib = _lib.load()
    for n in _declared():
        assert n in _lib.SIGNATURES, n
        getattr(lib, n)
    assert lib.hbp_version().startswith(b"hbp")


def test_missing_library_fails_loudly(tmp_path):
    saved = _lib._lib
    _lib._lib = None
    try:
        with pytest.raises(ImportError):
            _lib.load(str(tmp_path / "nope.so"))
    finally:
        _lib._lib = saved


def test_host_classes_keep_their_protocols():
    """Context is a context manager (smoke() and the drivers use `with`), and
    the module-level helpers stay functions (no method displaced into them)."""
    from tnbp_amd import tci
    assert hasattr(tci.Context, "__enter__") and hasattr(tci.Context, "__exit__")
    for name in ("close", "synchronize", "memory", "trim"):
        assert callable(getattr(tci.Context, name)), name
    import types
    for fn in (tci.device_memory, tci.autotune_report):
        assert not any(isinstance(c, types.CodeType) for c in fn.__code__.co_consts), fn
