"""The C-ABI library loads and exports every function the headers declare;
no compute call is made (these run without a GPU)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = []
    for h in ("shoccs_hip.h", "shoccs.h"):
        src = open(os.path.join(ROOT, "include", h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?[\w]+\s*\*?\s*(sh(?:x|occs)_\w+)\s*\(", src, flags=re.M)
    return sorted(set(names))


def test_headers_declare_entry_points():
    names = _declared()
    assert "shx_derivative_apply" in names
    assert "shx_rk4_step" in names
    assert "shoccs_problem_create" in names
    assert len(names) > 40


def test_library_exports_every_declared_symbol():
    from shoccs_amd import _lib

    L = _lib.lib()
    missing = [n for n in _declared() if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_covers_every_symbol():
    from shoccs_amd import _lib

    assert set(_declared()) == set(_lib.EXPORTED)


def test_context_fails_loudly_without_device():
    import shoccs_amd as S

    if os.path.exists("/dev/kfd"):
        pytest.skip("a device is present")
    with pytest.raises(S.ShoccsError):
        S.Context(0)


def test_no_cpu_fallback_in_product():
    """The package never loads the oracle: grep the product sources."""
    for dirpath, _, files in os.walk(os.path.join(ROOT, "shoccs_amd")):
        for f in files:
            if f.endswith((".py", ".cpp", ".hpp", ".hip", ".h")):
                src = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in src.lower(), f
