"""C ABI of libgdt.so without a GPU: the library loads, exports every entry
point include/gdt.h declares, and fails cleanly where a device is needed."""
import ctypes as C
import os
import re

import pytest

import gpu_determ_amd as G
import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "gdt.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gdt_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    names = declared()
    assert len(names) >= 30
    lib = C.CDLL(G.LIB_PATH)
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(G.ABI_SYMBOLS)


def test_library_is_gfx950_code():
    data = open(G.LIB_PATH, "rb").read()
    assert b"gfx950" in data


@pytest.mark.skipif(G.device_count() > 0, reason="checks the no-device path")
def test_no_device_errors_cleanly():
    with pytest.raises(G.GdtError, match="no HIP device"):
        G.Context((8, 8, 8), (-1, -1, -1), (1, 1, 1))
    assert G.lib().gdt_destroy(None) == 0
    assert G.lib().gdt_build_tables(None, 0) == -1
    assert b"null gdt handle" in G.lib().gdt_last_error()


def test_bad_grid_rejected():
    g = G.GridDesc()
    for a in range(3):
        g.n[a] = 8
        g.vmin[a] = 1.0
        g.vmax[a] = -1.0
    h = C.c_void_p()
    assert G.lib().gdt_create(C.byref(h), C.byref(g), 0) == -1
    assert b"vmin" in G.lib().gdt_last_error()


@pytest.mark.parametrize("M,seed", [(1 << 12, 1), (1 << 16, 7), (1 << 18, 2026)])
def test_generator_choice_matches_oracle(M, seed):
    """The host-side generator choice of the library and of the oracle are
    the same rule, summed in the same order: same a, same P_2 bits."""
    a1, p1 = G.korobov_pick(M, seed, 8)
    a2, p2 = O.korobov_pick(M, seed, 8)
    assert a1 == a2 and p1 == p2
