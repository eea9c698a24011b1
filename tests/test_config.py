"""The workload definitions and the seeded plasma loader (host logic)."""
import numpy as np
import pytest

from chimeracl_amd import config as C


@pytest.mark.parametrize("name,count", [("C1", 1_026_144), ("C4", 2_058_336)])
def test_particle_counts(name, count):
    c = C.CONFIGS[name]
    assert C.n_particles(c["Nx"], c["Nr"], c["ppc"]) == count


def test_grid_config_maps_dz_dr():
    g = C.grid_config("C2")
    dx = (g["Xmax"] - g["Xmin"]) / (g["Nx"] - 1)
    dr = g["Rmax"] / (g["Nr"] - 1.5)
    assert dx == pytest.approx(C.DZ) and dr == pytest.approx(C.DR) and g["dt"] == C.DZ
    assert g["DampCells"] == 50


def test_loader_is_seeded_and_in_cells():
    g = C.grid_config("C1", Nx=40, Nr=12)
    a = C.load_plasma(g, (2, 2, 4), 0.002, seed=0)
    b = C.load_plasma(g, (2, 2, 4), 0.002, seed=0)
    for k in a:
        assert np.array_equal(a[k], b[k])
    n = C.n_particles(40, 12, (2, 2, 4))
    assert a["x"].size == n
    dx, dr = C.DZ, C.DR
    cell = np.arange(n) // 16
    ix = 1 + cell % 37
    ir = cell // 37
    xg = g["Xmin"] + dx * ix
    assert np.all(a["x"] >= xg - 1e-12) and np.all(a["x"] < xg + dx + 1e-12)
    r = np.sqrt(a["y"] ** 2 + a["z"] ** 2)
    assert np.all(r >= ir * dr - 1e-12) and np.all(r < (ir + 1) * dr + 1e-12)
    w0 = 2 * np.pi * dx * dr * 0.002 / 16
    assert np.allclose(a["w"], w0 * r)
    assert np.all(a["g_inv"] == 1) and np.all(a["px"] == 0)


def test_density_profiles():
    g = C.grid_config("C3")
    x, f = C.density_profile("C3", g)
    assert x[0] < g["Xmin"] and f[:3] == [0.0, 0.0, 1.0]
    x, f = C.density_profile("C4", C.grid_config("C4"))
    assert len(x) == 32 and np.all(np.abs(np.array(f) - 1) <= 0.2)
    assert C.density_profile("C1", C.grid_config("C1")) is None


def test_cell_order_is_a_stable_permutation_into_key_order():
    g = C.grid_config("C1", Nx=40, Nr=12)
    a = C.load_plasma(g, (2, 2, 4), 0.002, seed=0)
    b = C.load_plasma(g, (2, 2, 4), 0.002, seed=0, order="cell")
    ka = C.cell_keys(g, a["x"], a["y"], a["z"])
    kb = C.cell_keys(g, b["x"], b["y"], b["z"])
    assert np.all(np.diff(kb) >= 0)
    perm = np.argsort(ka, kind="stable")
    for k in a:
        assert np.array_equal(a[k][perm], b[k])
    # every load cell straddles two grid rows: the load order is not key order
    assert np.any(np.diff(ka) < 0)
