"""The C-ABI library loads (no GPU needed) and exports every entry point include/gaukuk.h declares."""
import ctypes
import os
import re

from gaukuk_amd import hip

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "gaukuk.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gk_[a-z_0-9]+)\s*\(", text)))


def test_header_matches_binding_list():
    assert declared_symbols() == sorted(hip.SYMBOLS)


def test_library_exports_all_symbols():
    lib = ctypes.CDLL(hip.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_version_string_without_gpu():
    assert "gfx950" in hip.version()


def test_create_rejects_bad_grid_without_gpu():
    from gaukuk_amd import problems as P
    prob = P.kelvin_helmholtz(8, seed=0)
    g = hip.make_grid(prob)
    g.nz, g.ny = 2, 1   # a 3D grid of one row: rejected before any device work
    h = ctypes.c_void_p()
    st = hip.lib().gk_create(ctypes.byref(g), ctypes.byref(hip.make_phys(prob)), ctypes.byref(h))
    assert st == 1 and b"3D grids need NY >= 2" in hip.lib().gk_last_error(None)
    g.nz, g.ny, g.nghost = 1, 8, 1
    st = hip.lib().gk_create(ctypes.byref(g), ctypes.byref(hip.make_phys(prob)), ctypes.byref(h))
    assert st == 1 and b"NGhost" in hip.lib().gk_last_error(None)
