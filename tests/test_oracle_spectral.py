"""Pins of the CPU oracle's grid, tables, transforms and spectral operators
against closed forms, mpmath fixtures, analytic derivatives and the
reference's known-answer scripts.  CPU only."""
import json
import os

import numpy as np
import pytest
from scipy.special import jv

from chimeracl_amd import config as C

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _grid(Nx=128, Nr=64, M=1, dx=0.1, dr=0.1, damp=None):
    g = {"Xmin": -(Nx - 1) * dx, "Xmax": 0.0, "Nx": Nx, "Rmax": dr * (Nr - 1.5), "Nr": Nr, "M": M,
         "dt": dx}
    if damp:
        g["DampCells"] = damp
    return g


def test_geometry_closed_forms(oracle_mod):
    O = oracle_mod
    a = O.grid_args(_grid(Nx=100, Nr=40, dx=0.05, dr=0.2))
    assert a["dx"] == pytest.approx(0.05, rel=1e-15)
    assert a["dr"] == pytest.approx(0.2, rel=1e-15)
    assert a["Rgrid"][0] == pytest.approx(-0.1) and a["Rgrid"][1] == pytest.approx(0.1)
    assert a["dV_inv"][0] == 0
    assert a["dV_inv"][5] == pytest.approx(1 / (2 * np.pi * 0.05 * 0.2 * 0.9))
    sp = O.spectral_axes(a)
    assert np.allclose(sp["kx"][:50], 2 * np.pi * np.arange(50) / (100 * 0.05))
    R = 0.2 * 39
    assert np.allclose(jv(1, sp["kr"][1] * R), 0, atol=1e-13)
    w = sp["w"][0]
    assert np.allclose(w ** 2, sp["kx"][None, :] ** 2 + sp["kr"][0][:, None] ** 2)


def test_bessel_roots_vs_mpmath(oracle_mod):
    roots = json.load(open(os.path.join(GOLD, "bessel_roots.json")))
    for Nr in (16, 32, 128, 512):
        a = oracle_mod.grid_args(_grid(Nr=Nr, M=3))
        sp = oracle_mod.spectral_axes(a)
        R = a["dr"] * (Nr - 1)
        for m in range(5):
            ref = np.array(roots[str(m)][: Nr - 1])
            assert np.allclose(sp["kr"][m] * R, ref, rtol=2e-15, atol=0), (Nr, m)


def test_bessel_values_fixture():
    vals = json.load(open(os.path.join(GOLD, "bessel_values.json")))
    for m, pts in vals.items():
        for x, v in pts:
            assert abs(jv(int(m), x) - v) <= 1e-15 + 1e-13 * abs(v)


def test_dht_is_inverse(oracle_mod):
    for M in (0, 1, 2):
        a = oracle_mod.grid_args(_grid(Nr=96, M=M))
        d = oracle_mod.dht_matrices(a)
        for m in range(M + 1):
            assert np.abs(d["DHT"][m] @ d["DHT_inv"][m] - np.eye(95)).max() < 1e-11
            assert np.linalg.cond(d["DHT_inv"][m]) < 300


@pytest.mark.parametrize("m", [0, 1])
def test_dht_gaussian_hankel_pair(oracle_mod, m):
    """The forward DHT of r^m exp(-r^2/a^2) sampled on the grid gives the
    Fourier-Bessel series coefficients of that function on [0, R]:
    2/(R^2 J_{m+1}(j_k)^2) * a^(2m+2) k^m / 2^(m+1) * exp(-k^2 a^2 / 4)."""
    a = oracle_mod.grid_args(_grid(Nr=128, M=1))
    d = oracle_mod.dht_matrices(a)
    sp = oracle_mod.spectral_axes(a)
    r = a["Rgrid"][1:]
    aa = 1.2
    R = a["dr"] * (a["Nr"] - 1)
    f = r ** m * np.exp(-r ** 2 / aa ** 2)
    Fk = d["DHT"][m] @ f
    k = sp["kr"][m]
    jk = k * R
    ana = 2 / (R ** 2 * jv(m + 1, jk) ** 2) * aa ** (2 * m + 2) * k ** m / 2 ** (m + 1) * np.exp(-k ** 2 * aa ** 2 / 4)
    sel = slice(0, 20)
    err = np.abs(Fk[sel] - ana[sel]).max() / np.abs(ana[sel]).max()
    assert err < 5e-3, err


def _modes(fun, x, r, M, nphi=96):
    """F_m^ref(x, r) = i^m (1/2pi) oint F e^{-i m phi} dphi, y = r sin phi,
    z = r cos phi (the deposit/gather convention of grid_deposit_m1.cl)."""
    phi = 2 * np.pi * np.arange(nphi) / nphi
    X, Rr, P = np.meshgrid(x, r, phi, indexing="xy")
    X, Rr, P = (np.moveaxis(v, 2, 0) for v in (X, Rr, P))
    F = fun(X, Rr * np.sin(P), Rr * np.cos(P))
    out = []
    for m in range(M + 1):
        c = (1j ** m) * (F * np.exp(-1j * m * P)).mean(0)
        out.append(c.real if m == 0 else c)
    return out


def _blob(x0=-6.0, y0=0.5, z0=0.3, sx=0.9, sr=0.8):
    def f(X, Y, Z):
        return np.exp(-(X - x0) ** 2 / sx ** 2 - ((Y - y0) ** 2 + (Z - z0) ** 2) / sr ** 2)

    def d(X, Y, Z, comp):
        e = f(X, Y, Z)
        if comp == 0:
            return -2 * (X - x0) / sx ** 2 * e
        if comp == 1:
            return -2 * (Y - y0) / sr ** 2 * e
        return -2 * (Z - z0) / sr ** 2 * e
    return f, d


def _solver_with(oracle_mod, M):
    s = oracle_mod.OracleSolver(_grid(Nx=128, Nr=64, M=M))
    x, r = s.Args["Xgrid"], s.Args["Rgrid"][1:]
    return s, x, r


def _set_real(s, name, modes):
    for m, v in enumerate(modes):
        if m > s.M:
            break
        s.f[f"{name}_m{m}"][1:] = v


def test_grad_matches_analytic_derivative(oracle_mod):
    """field_grad (transformer_methods_cl.py:87-133) applied to the modes of
    an off-axis Gaussian equals the modes of its analytic gradient, for
    modes 0 and 1 (M=3 keeps the neighbours of m=1 and m=2)."""
    s, x, r = _solver_with(oracle_mod, 3)
    f, d = _blob()
    _set_real(s, "rho", _modes(f, x, r, 3))
    s.fb_transform(scals=["rho"], dir=0)
    s.field_grad("rho", "G")
    s.fb_transform(vects=["G"], dir=1)
    for k, c in enumerate("xyz"):
        ana = _modes(lambda X, Y, Z: d(X, Y, Z, k), x, r, 3)
        for m in range(2):
            got = s.f[f"G{c}_m{m}"][1:]
            err = np.abs(got - ana[m]).max() / np.abs(ana[m]).max()
            assert err < 2e-3, (c, m, err)


def test_rot_matches_analytic_curl(oracle_mod):
    """field_rot (transformer_methods_cl.py:185-263) of G = (f, 0, g) equals
    the modes of (0, d_z f - d_x g, d_y g ... ) computed analytically."""
    s, x, r = _solver_with(oracle_mod, 3)
    f, df = _blob()
    g, dg = _blob(x0=-5.5, y0=-0.4, z0=0.6, sx=1.0, sr=0.9)
    _set_real(s, "Gx", _modes(f, x, r, 3))
    _set_real(s, "Gz", _modes(g, x, r, 3))
    s.fb_transform(vects=["G"], dir=0)
    s.field_rot("G", "B")
    s.fb_transform(vects=["B"], dir=1)
    # curl (Gx, 0, Gz): (dy Gz - dz 0, dz Gx - dx Gz, dx 0 - dy Gx)
    comps = {"x": lambda X, Y, Z: dg(X, Y, Z, 1),
             "y": lambda X, Y, Z: df(X, Y, Z, 2) - dg(X, Y, Z, 0),
             "z": lambda X, Y, Z: -df(X, Y, Z, 1)}
    for c, fun in comps.items():
        ana = _modes(fun, x, r, 3)
        for m in range(2):
            got = s.f[f"B{c}_m{m}"][1:]
            err = np.abs(got - ana[m]).max() / np.abs(ana[m]).max()
            assert err < 2e-3, (c, m, err)


def test_div_matches_analytic(oracle_mod):
    s, x, r = _solver_with(oracle_mod, 3)
    f, df = _blob()
    g, dg = _blob(x0=-5.5, y0=-0.4, z0=0.6, sx=1.0, sr=0.9)
    h, dh = _blob(x0=-6.5, y0=0.2, z0=-0.5, sx=0.8, sr=1.0)
    for c, fun in zip("xyz", (f, g, h)):
        _set_real(s, "J" + c, _modes(fun, x, r, 3))
    s.fb_transform(vects=["J"], dir=0)
    div = s.field_div("J")
    ana = _modes(lambda X, Y, Z: df(X, Y, Z, 0) + dg(X, Y, Z, 1) + dh(X, Y, Z, 2), x, r, 3)
    phs = s.phase(1)[None, :]
    for m in range(2):
        got = np.fft.ifft(div[m] * phs, axis=1)
        got = s.dht["DHT_inv"][m] @ (got.real if m == 0 else got)
        err = np.abs(got - ana[m]).max() / np.abs(ana[m]).max()
        assert err < 2e-3, (m, err)


def test_round_trip_identity(oracle_mod):
    """examples/test_transformer.py:35-43: backward(forward(F)) = F, rows 1.."""
    for M in (0, 1, 2):
        s = oracle_mod.OracleSolver(_grid(Nx=96, Nr=40, M=M))
        rng = np.random.default_rng(M)
        ref = {}
        for m in range(M + 1):
            v = rng.normal(size=(39, 96)) + (1j * rng.normal(size=(39, 96)) if m else 0)
            s.f[f"rho_m{m}"][1:] = v
            ref[m] = v
        s.fb_transform(scals=["rho"], dir=0)
        for m in range(M + 1):
            s.f[f"rho_m{m}"][:] = 0
        s.fb_transform(scals=["rho"], dir=1)
        for m in range(M + 1):
            assert np.abs(s.f[f"rho_m{m}"][1:] - ref[m]).max() < 1e-12


def test_psatd_vacuum_exact_rotation(oracle_mod):
    """advance_e_g_m with J = dN = 0 is the exact harmonic evolution:
    w^2 |E|^2 + |G|^2 is conserved and n steps equal one step of n*dt."""
    s = oracle_mod.OracleSolver(_grid(Nx=64, Nr=24, M=1))
    rng = np.random.default_rng(3)
    for f in ("E", "G"):
        for c in "xyz":
            for m in range(2):
                s.f[f"{f}{c}_fb_m{m}"][:] = rng.normal(size=(23, 64)) + 1j * rng.normal(size=(23, 64))
    w = s.sp["w"]
    e0 = {m: s.f[f"Ex_fb_m{m}"].copy() for m in range(2)}
    g0 = {m: s.f[f"Gx_fb_m{m}"].copy() for m in range(2)}
    inv0 = sum((w[m] ** 2 * abs(s.f[f"Ex_fb_m{m}"]) ** 2 + abs(s.f[f"Gx_fb_m{m}"]) ** 2).sum() for m in range(2))
    n = 7
    for _ in range(n):
        s.push_fields()
    inv1 = sum((w[m] ** 2 * abs(s.f[f"Ex_fb_m{m}"]) ** 2 + abs(s.f[f"Gx_fb_m{m}"]) ** 2).sum() for m in range(2))
    assert abs(inv1 - inv0) / inv0 < 1e-12
    T = n * s.Args["dt"]
    for m in range(2):
        ana = np.cos(w[m] * T) * e0[m] + np.sin(w[m] * T) / w[m] * g0[m]
        assert np.abs(s.f[f"Ex_fb_m{m}"] - ana).max() < 1e-10 * np.abs(ana).max()


def test_damp_profile_closed_form(oracle_mod):
    for N in (4, 50):
        p = oracle_mod.damp_profile(N)
        assert p.size == 2 * N
        assert np.all(p[:N] == 0)
        assert np.all(p[int(np.ceil(4 * N / 3)):] == 1)
        assert np.all(np.diff(p) >= -1e-15)


def test_smoothing_filter_closed_form(oracle_mod):
    a = oracle_mod.grid_args(_grid(Nx=64, Nr=24, M=1))
    sp = oracle_mod.spectral_axes(a)
    f = sp["Smooth"][0]
    kxmax = np.argmax(sp["kx"])
    assert np.allclose(f[:, kxmax], 0, atol=1e-15)
    assert np.allclose(f[-1, :], 0, atol=1e-15)
    assert f[0, 0] == pytest.approx(1 - np.sin(0.5 * np.pi * sp["kr"][0][0] / sp["kr"][0][-1]) ** 2)


def test_laser_group_velocity_kat(oracle_mod):
    """examples/test_laser_veloc.py:9-59 with the reference's parameters:
    the laser centroid deficit 1 - v matches (2 pi R)^-2 (r-weighted) and
    2 (2 pi R)^-2 (on axis).  The reference prints the deviation without a
    threshold; 10% is the bar here."""
    O = oracle_mod
    g = {"Xmin": -50.0, "Xmax": 50.0, "Rmax": 50.0, "dt": 0.05, "M": 0}
    dx, dr = 0.05, 0.3
    g["Nx"] = int((g["Xmax"] - g["Xmin"]) / dx) // 2 * 2
    g["Nr"] = int(g["Rmax"] / dr) // 2 * 2 + 1
    laser = {"k0": 1.0, "a0": 1.0, "x0": 0, "Lx": 12.0, "R": 12.0, "x_foc": 20.0}
    s = O.OracleSolver(g)
    O.add_gausian_pulse(s, laser)
    Xg, Rg = s.Args["Xgrid"], s.Args["Rgrid"]
    c1, c2 = [], []
    for _ in range(100):
        s.push_fields()
        s.fb_transform(scals=["Ez"], dir=1)
        v = s.f["Ez_m0"]
        Px = (Rg[1:, None] * v[1:, :] ** 2).sum(0)
        c1.append((Xg[None, :] * Px).sum() / Px.sum())
        Px = v[1, :] ** 2
        c2.append((Xg[None, :] * Px).sum() / Px.sum())
    th = (2 * np.pi * laser["R"]) ** -2
    e1 = np.abs(1 - np.diff(c1) / g["dt"] - th).mean() / th
    e2 = np.abs(1 - np.diff(c2) / g["dt"] - 2 * th).mean() / (2 * th)
    assert e1 < 0.1 and e2 < 0.1, (e1, e2)


def test_oracle_tiny_trajectory_fixture(oracle_mod):
    """Regression: 3 steps of the 64 x 24, M=1 set-up reproduce the committed
    oracle state."""
    from tests import simpair as S
    O = oracle_mod
    ref = np.load(os.path.join(GOLD, "oracle_tiny.npz"))
    g = S.tiny_grid(Nx=64, Nr=24, M=1)
    s = O.OracleSolver(dict(g))
    O.add_gausian_pulse(s, S.tiny_laser(g))
    eo, io = C.species_config("C1", s.Args)
    arr = C.load_plasma(g, (2, 2, 4), 0.002, seed=0)
    e, i = O.OracleParticles(eo), O.OracleParticles(io)
    e.add_particles(arr)
    i.add_particles(arr)
    loop = O.OracleLoop(s, [e, i])
    for _ in range(3):
        loop.step()
    for k in S.real_names(1):
        a = s.f[k][1:]
        assert np.linalg.norm(a) == pytest.approx(ref["norm_" + k][0], rel=1e-12, abs=1e-300)
        assert np.allclose(a.ravel()[ref["idx_" + k]], ref["samp_" + k], rtol=1e-10, atol=1e-12 * np.abs(a).max())
    for k in ("px", "py", "pz", "x"):
        assert np.allclose(e.data[k][::97], ref["p_" + k], rtol=1e-10, atol=1e-14)
    assert np.array_equal(e.data["sort_indx"][:512], ref["sort_head"])
    assert e.Args["Np_stay"] == ref["np_stay"][0]
