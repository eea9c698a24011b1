"""The C-ABI library loads and exports every entry point include/chimera_hip.h
declares; host-only queries and argument checks behave (no GPU needed)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "chimera_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(
        r"(?:\b(?:int|int64_t|size_t)\s+|\bconst\s+uint32_t\s*\*\s*)(chm_\w+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    from chimeracl_amd import _lib
    names = _declared()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(_lib.lib, n)]
    assert not missing, missing
    # and the ctypes table binds exactly the declared set
    assert set(_lib.exported_symbols()) == set(names)


def test_struct_layouts_match_header():
    from chimeracl_amd import _lib
    # chm_spectral_t: 6 int32 + double + 5 pointers + 15 arrays of 5 pointers
    assert ctypes.sizeof(_lib.Spectral) == 6 * 4 + 8 + 5 * 8 + 15 * 5 * 8
    assert ctypes.sizeof(_lib.Grid) == 16 + 32
    assert ctypes.sizeof(_lib.Planes) == 8 * 5
    assert ctypes.sizeof(_lib.TablesOut) == 6 * 8 + 9 * 5 * 8


def test_workspace_queries():
    from chimeracl_amd import _lib
    L = _lib.lib
    assert L.chm_sort_workspace_bytes(1000, 500) >= (500 + 4) * 4 * 2
    # deposit staging: 4 corners x components x cells, doubles
    assert L.chm_deposit_workspace_bytes(512, 128, 1, 1) >= 4 * 9 * 511 * 127 * 8
    assert L.chm_transform_workspace_bytes(512, 128, 1, 4, 8) >= 4 * 2 * 127 * 512 * 16
    assert L.chm_tables_workspace_bytes(512, 128, 1) >= 127 * 254 * 8


def test_argument_errors_without_gpu():
    """Entry points reject bad arguments before touching the device."""
    from chimeracl_amd import _lib
    L = _lib.lib
    g = _lib.Grid(Nx=2, Nr=2, M=1)  # too small
    rc = L.chm_index_sort_f64(None, None, None, 10, ctypes.byref(g), None, None, None, None, None, 0, None)
    assert rc == -1
    g = _lib.Grid(Nx=64, Nr=24, M=9)  # mode out of range
    rc = L.chm_deposit_scalar_f64(None, None, None, None, None, None, -1, ctypes.byref(g), None, None, 0, None)
    assert rc == -1
    sp = _lib.Spectral(Nx=100, Nr=24, M=1)  # not a power of two
    pl = (_lib.Planes * 1)()
    fb = (_lib.Fb * 1)()
    rc = L.chm_transform_f64(ctypes.byref(sp), pl, fb, 1, 0, 1, 0, None, 0, None)
    assert rc == -3
    with pytest.raises(RuntimeError):
        _lib.call("chm_push_x_f64", None, None, None, None, None, None, None, 0.1, -5, None)


def test_align_multi_argument_errors_without_gpu():
    """chm_align_multi checks its array table before it launches: array
    counts outside 1..16, element sizes other than 4 and 8, a destination that
    is also a source or another destination are refused (CHM_EINVAL); n = 0
    launches nothing."""
    from chimeracl_amd import _lib
    L = _lib.lib
    P = ctypes.c_void_p
    src = (P * 2)(0x1000, 0x2000)
    dst = (P * 2)(0x3000, 0x4000)
    eb = (ctypes.c_int32 * 2)(8, 4)
    a = lambda arr: ctypes.addressof(arr)
    srt = 0x5000
    assert L.chm_align_multi(srt, 10, 0, a(src), a(dst), a(eb), None) == -1
    assert L.chm_align_multi(srt, 10, 17, a(src), a(dst), a(eb), None) == -1
    assert L.chm_align_multi(srt, 10, 2, a(src), a(dst), a((ctypes.c_int32 * 2)(8, 3)), None) == -1
    assert L.chm_align_multi(srt, 10, 2, a(src), a((P * 2)(0x3000, 0x1000)), a(eb), None) == -1
    assert L.chm_align_multi(srt, 10, 2, a(src), a((P * 2)(0x3000, 0x3000)), a(eb), None) == -1
    assert L.chm_align_multi(None, 10, 2, a(src), a(dst), a(eb), None) == -1
    assert L.chm_align_multi(None, 0, 2, a(src), a(dst), a(eb), None) == 0


def test_noalloc_sort_rejects_oversize_without_gpu():
    """The allocation-free sort checks the size against its fixed scratch
    before it touches the device: 2^22 particles on ~2^22 cells exceed it
    (CHM_ESCRATCH), up to 2^21 particles on 2^22 cells fit."""
    from chimeracl_amd import _lib
    L = _lib.lib
    g = _lib.Grid(Nx=4096, Nr=1025, M=1)
    nbins = 4095 * 1024 + 1
    assert L.chm_sort_workspace_bytes(1 << 21, nbins) <= 48 << 20
    assert L.chm_sort_workspace_bytes(1 << 22, nbins) > 48 << 20
    rc = L.chm_index_sort_f64_noalloc(None, None, None, 1 << 22, ctypes.byref(g), None, None,
                                      None, None, None)
    assert rc == -4


def test_dht_timing_record_empty_without_launches():
    from chimeracl_amd import _lib
    L = _lib.lib
    assert L.chm_dht_timing(1) == 0
    assert L.chm_dht_timing_collect(None, None, None, 10) == 0
    assert L.chm_dht_timing(0) == 0


def test_sort_sorted_keys_address():
    """chm_sort_sorted_keys is host-side address arithmetic: for a workspace
    of chm_sort_workspace_bytes(Np, nbins) bytes it names Np keys inside it
    (4-byte aligned), and NULL for a missing workspace or grid, or Np <= 0.
    Nothing is launched, so no GPU is needed."""
    import ctypes
    from chimeracl_amd import _lib
    base = 1 << 32
    for Nx, Nr, Np in ((64, 24, 1000), (4096, 512, 66_797_760), (33, 5, 1)):
        g = _lib.Grid(Nx, Nr, 1, 0, 0.0, 1.0, 0.0, 1.0)
        nbins = (Nx - 1) * (Nr - 1) + 1
        size = _lib.lib.chm_sort_workspace_bytes(Np, nbins)
        a = _lib.lib.chm_sort_sorted_keys(base, Np, ctypes.byref(g))
        assert a is not None and a % 4 == 0
        assert base <= a and a + 4 * Np <= base + size
    g = _lib.Grid(64, 24, 1, 0, 0.0, 1.0, 0.0, 1.0)
    assert _lib.lib.chm_sort_sorted_keys(None, 10, ctypes.byref(g)) is None
    assert _lib.lib.chm_sort_sorted_keys(base, 0, ctypes.byref(g)) is None
    assert _lib.lib.chm_sort_sorted_keys(base, 10, None) is None
