"""chm_index_sort_f64_noalloc (the sort with its scratch in a fixed device
array) against chm_index_sort_f64 with a caller-allocated workspace, through
the C ABI: bit-identical keys, cell offsets, permutation and Np_stay on two
calls in a row (the second smaller), and at the stated maximum size."""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


def _grid(Nx, Nr, dx=0.02, dr=0.1):
    from chimeracl_amd import _lib
    xmin = -(Nx - 1) * dx
    return _lib.Grid(Nx, Nr, 1, 0, xmin, 1.0 / dx, -0.5 * dr, 1.0 / dr), xmin, dr * (Nr - 1.5)


def _positions(n, xmin, rmax, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(xmin - 0.05, 0.05, n)          # a few outside in x
    r = rng.uniform(0.0, 1.02 * rmax, n)            # a few outside in r
    th = rng.uniform(0, 2 * np.pi, n)
    return x, r * np.sin(th), r * np.cos(th)


def _sort(name, g, x, y, z, with_ws):
    from chimeracl_amd import _lib
    dev = torch.device("cuda")
    n = x.size
    nbins = (g.Nx - 1) * (g.Nr - 1) + 1
    X = [torch.from_numpy(v).to(dev) for v in (x, y, z)]
    key = torch.full((n,), -1, dtype=torch.int32, device=dev)
    off = torch.full((nbins + 1,), -1, dtype=torch.int32, device=dev)
    srt = torch.full((n,), -1, dtype=torch.int32, device=dev)
    nps = torch.full((1,), -1, dtype=torch.int32, device=dev)
    st = torch.cuda.current_stream().cuda_stream
    args = [X[0].data_ptr(), X[1].data_ptr(), X[2].data_ptr(), n, ctypes.byref(g),
            key.data_ptr(), off.data_ptr(), srt.data_ptr(), nps.data_ptr()]
    if with_ws:
        ws = torch.empty(int(_lib.lib.chm_sort_workspace_bytes(n, nbins)), dtype=torch.uint8,
                         device=dev)
        _lib.call(name, *args, ws.data_ptr(), ws.numel(), st)
    else:
        _lib.call(name, *args, st)
    torch.cuda.synchronize()
    return [t.cpu().numpy().view(np.uint32) for t in (key, off, srt, nps)]


def test_noalloc_sort_matches_workspace_sort():
    cases = [(512, 128, 1_500_000, 1), (512, 128, 300_000, 2),   # two in a row, the second smaller
             (4096, 1025, 1 << 21, 3)]                           # the stated maximum: 2^21, ~2^22 cells
    for Nx, Nr, n, seed in cases:
        print(f"noalloc sort case Nx={Nx} Nr={Nr} Np={n}", flush=True)
        g, xmin, rmax = _grid(Nx, Nr)
        x, y, z = _positions(n, xmin, rmax, seed)
        ref = _sort("chm_index_sort_f64", g, x, y, z, True)
        got = _sort("chm_index_sort_f64_noalloc", g, x, y, z, False)
        for a, b, what in zip(got, ref, ("keys", "cell_offset", "sort_indx", "np_stay")):
            assert np.array_equal(a, b), (what, Nx, Nr, n)
        key = ref[0]
        assert np.array_equal(ref[2], np.argsort(key, kind="stable").astype(np.uint32))
        assert 0 < int(ref[3][0]) < n   # some particles outside the valid cells


def test_noalloc_sort_rejects_oversize():
    from chimeracl_amd import _lib
    g, _, _ = _grid(4096, 1025)
    rc = _lib.lib.chm_index_sort_f64_noalloc(None, None, None, (1 << 22), ctypes.byref(g),
                                             None, None, None, None, None)
    assert rc == -4
