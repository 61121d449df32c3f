"""CPU-side checks of the C-ABI library: it loads, exports every symbol the
public headers declare, and the headers declare what the binding expects.
No compute call is made (no GPU here)."""
import ctypes
import os
import re

import pytest

from spme_amd import lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in ("spme.h", "spme_ext.h"):
        src = open(os.path.join(ROOT, "include", h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"\b([a-z_0-9]+)\s*\([^;{]*\)\s*;", src):
            names.add(m.group(1))
    return names


def test_headers_declare_the_binding():
    assert _declared() == set(lib.EXPORTED)


@pytest.mark.skipif(not os.path.exists(lib.LIB_PATH), reason="library not built")
def test_library_exports_every_symbol():
    L = ctypes.CDLL(lib.LIB_PATH)
    for name in lib.EXPORTED:
        assert hasattr(L, name), name


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(lib, "LIB_PATH", str(tmp_path / "nope.so"))
    monkeypatch.setattr(lib, "_lib", None)
    with pytest.raises(lib.SpmeError, match="no CPU fallback"):
        lib.load()


@pytest.mark.skipif(not os.path.exists(lib.LIB_PATH), reason="library not built")
def test_cutoff2xi_host_only():
    """cutoff2xi runs on the host: the real-space tensor at rmax equals eps."""
    from oracle import oracle as O
    from spme_amd.engine import cutoff2xi
    for rc, eps in [(4.0, 1e-4), (5.0, 1e-6), (8.0, 1e-9)]:
        xi = cutoff2xi(rc, eps, 100.0)
        F, G = O.real_coeffs(rc, xi)
        assert abs(max(abs(F), abs(F + G)) - eps) < 1e-6 * eps
