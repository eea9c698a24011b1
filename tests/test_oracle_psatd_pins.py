"""Pins of the oracle's PSATD source terms (advance_e_g_m,
solver_ms_pic.cl:114-134), of the whole step's field/particle coupling and of
the moving window's phase convention (transformer_generic.cl:28-55), each
against physics that does not come from the update formula itself.  CPU only.

Units (normalised to the laser wavelength, time to lambda0/c, density to the
critical density n_c, fields to m c omega0 / e): with omega0 = 2 pi,
    div E = 2 pi rho,   dE/dt = curl B - 2 pi J,   du/dt = 2 pi (q/m)(E + v x B),
so the cold-plasma frequency is omega_p = 2 pi sqrt(n / n_c).  In spectral
form with G = curl B and N = grad rho (dN in the solver):
    dE/dt = G - 2 pi J,   dG/dt = -w^2 E - 2 pi N,
with J constant over the step and N linear in time from dN0 to dN1.

The checks are functions returning an error figure, so that
test_planted_mistakes_are_caught can run them against oracles with a
planted mistake (a dropped 2 pi, a flipped sign, swapped dN0 / dN1, the
opposite phase convention) and show that each mistake fails one of them.
"""
import numpy as np
import pytest
from scipy.linalg import expm
from scipy.special import jn_zeros, jv

from chimeracl_amd import config as C

TWO_PI = 2 * np.pi


def _grid(Nx=48, Nr=20, M=1, dx=0.1, dr=0.1, damp=None):
    g = {"Xmin": -(Nx - 1) * dx, "Xmax": 0.0, "Nx": Nx, "Rmin": 0.0, "Rmax": dr * (Nr - 1.5),
         "Nr": Nr, "M": M, "dt": dx}
    if damp is not None:
        g["DampCells"] = damp
    return g


def _crand(rng, shape):
    return rng.normal(size=shape) + 1j * rng.normal(size=shape)


def _set(s, fam, vals):
    """vals[(c, m)] -> spectral plane of field family fam."""
    for (c, m), v in vals.items():
        s.f[f"{fam}{c}_fb_m{m}"][:] = v


# ---------------------------------------------------------------- checks
def static_charge_error(O):
    """Electrostatic equilibrium: a charge density rho at rest (dN0 = dN1 =
    grad rho, J = 0) with its Coulomb field E = -2 pi grad rho / w^2 (Gauss's
    law div E = 2 pi rho for a curl-free E) and G = curl B = 0 is a fixed
    point of the update.  Returns the worst relative change of E and G."""
    s = O.OracleSolver(_grid(M=1))
    rng = np.random.default_rng(11)
    for m in range(2):
        s.f[f"rho_fb_m{m}"][:] = _crand(rng, s.f[f"rho_fb_m{m}"].shape)
    s.field_grad("rho", "dN1")
    e0 = {}
    for m in range(2):
        for c in "xyz":
            n = s.f[f"dN1{c}_fb_m{m}"]
            s.f[f"dN0{c}_fb_m{m}"][:] = n
            s.f[f"J{c}_fb_m{m}"][:] = 0
            s.f[f"G{c}_fb_m{m}"][:] = 0
            e0[c, m] = -TWO_PI * n / s.sp["w"][m] ** 2
    _set(s, "E", e0)
    s.push_fields()
    err = 0.0
    for (c, m), e in e0.items():
        scale = np.abs(e).max()
        err = max(err, np.abs(s.f[f"E{c}_fb_m{m}"] - e).max() / scale,
                  np.abs(s.f[f"G{c}_fb_m{m}"] / s.sp["w"][m]).max() / scale)
    return err


def driven_error(O, nsteps=9):
    """A constant current J (no charge: dN = 0) drives dE/dt = G - 2 pi J,
    dG/dt = -w^2 E.  nsteps updates are compared with the exact solution at
    t = nsteps*dt, computed per element as the matrix exponential of that
    linear system on (E, G, J) (scipy.linalg.expm; no closed form used).
    Returns the relative error over a sample of elements."""
    s = O.OracleSolver(_grid(M=1))
    rng = np.random.default_rng(12)
    st = {}
    for fam in ("E", "G", "J"):
        for m in range(2):
            for c in "xyz":
                v = _crand(rng, s.f[f"{fam}{c}_fb_m{m}"].shape)
                if fam == "G":
                    v = v * s.sp["w"][m]
                st[fam, c, m] = v
                s.f[f"{fam}{c}_fb_m{m}"][:] = v
    for fam in ("dN0", "dN1"):
        for m in range(2):
            for c in "xyz":
                s.f[f"{fam}{c}_fb_m{m}"][:] = 0
    for _ in range(nsteps):
        s.push_fields()
    T = nsteps * s.Args["dt"]
    err = 0.0
    shp = s.f["Ex_fb_m0"].shape
    idx = [(int(i), int(j)) for i, j in zip(rng.integers(0, shp[0], 40), rng.integers(0, shp[1], 40))]
    for m in range(2):
        w = s.sp["w"][m]
        for c in "xyz":
            for i, j in idx:
                A = np.array([[0.0, 1.0, -TWO_PI], [-w[i, j] ** 2, 0.0, 0.0], [0.0, 0.0, 0.0]])
                y = expm(A * T) @ np.array([st["E", c, m][i, j], st["G", c, m][i, j],
                                             st["J", c, m][i, j]])
                sc = abs(st["E", c, m][i, j]) + abs(st["J", c, m][i, j]) / w[i, j] + \
                    abs(st["G", c, m][i, j]) / w[i, j]
                err = max(err, abs(s.f[f"E{c}_fb_m{m}"][i, j] - y[0]) / sc,
                          abs(s.f[f"G{c}_fb_m{m}"][i, j] - y[1]) / (w[i, j] * sc))
    return err


def gauss_error(O, nsteps=4):
    """Gauss's law is kept by the update when the sources obey continuity.
    Each spectral element is given a wave vector k = (kx, kr cos a, kr sin a)
    with a random angle a, so |k| = w and the update's per-component
    arithmetic acts on the Cartesian components of 3-vectors.  Start from
    div E = i k.E = 2 pi rho0 (plus a random transverse part) and k.G = 0;
    every step takes rho_n -> rho_(n+1) (random), dN = i k rho (the
    gradient), and J = i k (rho_(n+1) - rho_n) / (dt w^2) plus a random
    transverse current, so that i k.J = -(rho_(n+1) - rho_n)/dt.  After each
    step i k.E must equal 2 pi rho_(n+1) and k.G must vanish.  Returns the
    worst relative violation."""
    s = O.OracleSolver(_grid(M=1))
    a = s.Args
    rng = np.random.default_rng(13)
    dt = a["dt"]
    shp = s.f["Ex_fb_m0"].shape
    K, rho = {}, {}
    for m in range(2):
        kr = s.sp["kr"][m][:, None] * np.ones(shp)
        kx = a["kx"][None, :] * np.ones(shp)
        ang = rng.uniform(0, TWO_PI, shp)
        K[m] = np.stack([kx, kr * np.cos(ang), kr * np.sin(ang)])
        assert np.allclose((K[m] ** 2).sum(0), s.sp["w"][m] ** 2)
        rho[m] = _crand(rng, shp)

    def transverse(m):
        v = _crand(rng, (3,) + shp)
        k = K[m]
        return v - k * (k * v).sum(0) / (k * k).sum(0)

    for m in range(2):
        k, w2 = K[m], s.sp["w"][m] ** 2
        E = -1j * k * TWO_PI * rho[m] / w2 + transverse(m)   # i k.E = 2 pi rho
        G = transverse(m) * np.sqrt(w2)
        for ci, c in enumerate("xyz"):
            s.f[f"E{c}_fb_m{m}"][:] = E[ci]
            s.f[f"G{c}_fb_m{m}"][:] = G[ci]
            s.f[f"dN1{c}_fb_m{m}"][:] = 1j * k[ci] * rho[m]
    err = 0.0
    for _ in range(nsteps):
        for m in range(2):
            k, w2 = K[m], s.sp["w"][m] ** 2
            new = _crand(rng, shp)
            J = 1j * k * (new - rho[m]) / (dt * w2) + transverse(m)
            for ci, c in enumerate("xyz"):
                s.f[f"dN0{c}_fb_m{m}"][:] = s.f[f"dN1{c}_fb_m{m}"]
                s.f[f"dN1{c}_fb_m{m}"][:] = 1j * k[ci] * new
                s.f[f"J{c}_fb_m{m}"][:] = J[ci]
            rho[m] = new
        s.push_fields()
        for m in range(2):
            k = K[m]
            E = np.stack([s.f[f"E{c}_fb_m{m}"] for c in "xyz"])
            G = np.stack([s.f[f"G{c}_fb_m{m}"] for c in "xyz"])
            gauss = (1j * (k * E).sum(0) - TWO_PI * rho[m])
            scale = TWO_PI * np.abs(rho[m]).max()
            err = max(err, np.abs(gauss).max() / scale,
                      np.abs((k * G).sum(0) / s.sp["w"][m] ** 2).max() / scale)
    return err


def langmuir_frequency_ratio(O, nsteps=200, dens=0.05, amp=1e-3):
    """Whole-step known answer: a cold plasma oscillation.  Electrons
    (2x2x4 per cell, immobile ions at the same positions) get the
    irrotational velocity u = amp grad[cos(kx x) J0(kr r)] (kx = two
    wavelengths along the box, kr the first J0 root over the radial
    period), so the motion is purely electrostatic and oscillates at the
    plasma frequency 2 pi sqrt(n), reduced by the step's smoothing of rho
    and J (a factor S(kx, kr) on both, so omega^2 = S omega_p^2; S is close
    to 1 here).  The electrons' projected momentum over nsteps steps is fitted
    with a sinusoid; returns fitted omega / (sqrt(S) omega_p)."""
    Nx, Nr, dx, dr = 32, 32, 0.05, 0.1
    g = _grid(Nx=Nx, Nr=Nr, M=1, dx=dx, dr=dr, damp=0)
    s = O.OracleSolver(dict(g))
    a = s.Args
    eo = {"Nppc": (2, 2, 4), "dx": a["dx"], "dr": a["dr"], "dt": a["dt"], "dens": dens,
          "charge": -1}
    io = dict(eo, charge=1, Immobile=True)
    arr = C.load_plasma(g, (2, 2, 4), dens, seed=0)
    kx = TWO_PI / (Nx * a["dx"])
    kr = jn_zeros(0, 1)[0] / (a["dr"] * (Nr - 1))
    x, y, z = arr["x"], arr["y"], arr["z"]
    r = np.hypot(y, z)
    arr["px"] = -amp * kx * np.sin(kx * x) * jv(0, kr * r)
    ur = -amp * kr * np.cos(kx * x) * jv(1, kr * r)
    arr["py"], arr["pz"] = ur * y / r, ur * z / r
    arr["g_inv"] = 1 / np.sqrt(1 + arr["px"] ** 2 + arr["py"] ** 2 + arr["pz"] ** 2)
    e, i = O.OracleParticles(eo), O.OracleParticles(io)
    e.add_particles(arr)
    i.add_particles(arr)
    loop = O.OracleLoop(s, [e, i])
    P = []
    for _ in range(nsteps):
        loop.step()
        d = e.data
        b = np.sin(kx * d["x"]) * jv(0, kr * np.hypot(d["y"], d["z"]))
        P.append((d["px"] * b).sum() / (b * b).sum())
    P = np.array(P)
    t = a["dt"] * (np.arange(nsteps) + 0.5)    # momenta sit at half steps
    wp = TWO_PI * np.sqrt(dens)
    S = ((1 - np.sin(0.5 * np.pi * kx / s.sp["kx"].max()) ** 2)
         * (1 - np.sin(0.5 * np.pi * kr / s.sp["kr"][0].max()) ** 2))
    ws = np.linspace(0.3 * wp, 3.0 * wp, 5401)
    res = [np.linalg.lstsq(np.stack([np.sin(w * t), np.cos(w * t)], 1), P, rcond=None)[1]
           for w in ws]
    res = [float(v[0]) if len(v) else np.inf for v in res]
    return ws[int(np.argmin(res))] / (np.sqrt(S) * wp)


def window_drift_error(O, nsteps=120, K=10):
    """Moving window (Frame at Velocity 1, a shift of K cells every K steps,
    frame.py:22-35 and the phase e^(-/+ i kx Xmin) of every transform): a
    vacuum laser pulse travels at its group velocity, which falls short of c
    by (2 pi R)^-2 for the r-weighted centroid (the reference's own
    known answer, examples/test_laser_veloc.py:46), so in the coordinates of a
    window moving at c its centroid moves back by T (2 pi R)^-2.  Runs the
    loop (no particles) and returns |centroid drift - expected| in units of
    the drift the window itself makes (T)."""
    Nx, Nr, dx, dr = 256, 32, 0.05, 0.25
    g = _grid(Nx=Nx, Nr=Nr, M=0, dx=dx, dr=dr, damp=8)
    s = O.OracleSolver(dict(g))
    a = s.Args
    R = 2.0
    x0 = a["Xmin"] + 0.5 * (a["Xmax"] - a["Xmin"])
    O.add_gausian_pulse(s, {"k0": 1.0, "a0": 1.0, "x0": x0, "Lx": 1.0, "R": R, "x_foc": x0})
    fr = O.OracleFrame({"Velocity": 1.0, "dt": a["dt"], "Steps": K})
    loop = O.OracleLoop(s, [], frames=[fr])

    def centroid():
        v = s.f["Ez_m0"][1:]
        px = (a["Rgrid"][1:, None] * v ** 2).sum(0)
        return (a["Xgrid"] * px).sum() / px.sum() - a["Xmin"]

    c0 = centroid()
    xmin0 = a["Xmin"]
    for _ in range(nsteps):
        loop.step()
    c1 = centroid()
    T = nsteps * a["dt"]
    moved = a["Xmin"] - xmin0
    assert abs(moved - T) < 1e-9   # the window moved at c
    expected = T - moved - T * (TWO_PI * R) ** -2
    return abs((c1 - c0) - expected) / T


# ---------------------------------------------------------------- tests
def test_psatd_static_charge_fixed_point(oracle_mod):
    assert static_charge_error(oracle_mod) < 1e-12


def test_psatd_constant_current_exact_solution(oracle_mod):
    assert driven_error(oracle_mod) < 1e-11


def test_psatd_keeps_gauss_law(oracle_mod):
    assert gauss_error(oracle_mod) < 1e-11


def test_langmuir_oscillation_frequency(oracle_mod):
    """Cold plasma oscillation at 2 pi sqrt(n): pins the 2 pi of J and rho in
    the field update together with the deposit's normalisation and the
    push's FactorPush = 2 pi dt q/m (observed 1.0015)."""
    assert abs(langmuir_frequency_ratio(oracle_mod) - 1) < 0.01


def test_moving_window_keeps_pulse_in_place(oracle_mod):
    """Observed 4e-4 of the window's own motion (2e-3 cells)."""
    assert window_drift_error(oracle_mod) < 2e-3


# ---------------------------------------------------------------- planted mistakes
def _mutant(kind):
    """A PSATD update with one planted mistake (for the mutation test)."""
    def upd(e0, g0, j, n0, n1, c1, c2, c3, dti):
        s_j = 1.0 if kind == "J without 2 pi" else TWO_PI
        s_n = 1.0 if kind == "dN without 2 pi" else TWO_PI
        j0, a0, a1 = j * s_j, n0 * s_n, n1 * s_n
        if kind == "dN0, dN1 swapped":
            a0, a1 = a1, a0
        sj = -1.0 if kind == "sign of J in E" else 1.0
        e1 = c1 * e0 + c2 * c3 * (g0 - sj * j0) + c3 * (c1 * a0 - a1 - (a0 - a1) * dti * c2 * c3)
        dn = (a1 - a0) if kind == "sign of (n0 - n1) in G" else (a0 - a1)
        jj = 0.0 if kind == "no + J in G" else j0
        nn = a1 if kind == "c2 n1 in G" else a0
        g1 = -c2 * e0 + c1 * (g0 - j0) + jj + c3 * (dti * (1.0 - c1) * dn - c2 * nn)
        return e1, g1
    return upd


MUTANTS = ["J without 2 pi", "dN without 2 pi", "dN0, dN1 swapped", "sign of J in E",
           "sign of (n0 - n1) in G", "no + J in G", "c2 n1 in G"]


@pytest.mark.parametrize("kind", MUTANTS)
def test_planted_mistakes_are_caught(oracle_mod, monkeypatch, kind):
    """Each planted mistake in the oracle's PSATD update makes one of the
    checks above fail (and the unmodified update passes all of them)."""
    O = oracle_mod
    monkeypatch.setattr(O, "psatd_update", _mutant(kind))
    errs = {"static": static_charge_error(O), "driven": driven_error(O), "gauss": gauss_error(O)}
    assert max(errs["static"] / 1e-12, errs["driven"] / 1e-11, errs["gauss"] / 1e-11) > 1e3, errs


def test_planted_mistake_control(oracle_mod, monkeypatch):
    """The mutant factory with no mistake is the oracle's update."""
    O = oracle_mod
    monkeypatch.setattr(O, "psatd_update", _mutant("none"))
    assert static_charge_error(O) < 1e-12
    assert driven_error(O) < 1e-11
    assert gauss_error(O) < 1e-11


def test_planted_phase_convention_is_caught(oracle_mod, monkeypatch):
    """The opposite phase convention (e^(+i kx Xmin) on the forward
    transform) moves the pulse in the moving window."""
    O = oracle_mod
    orig = O.OracleSolver.phase
    monkeypatch.setattr(O.OracleSolver, "phase", lambda self, d: orig(self, 1 - d))
    assert window_drift_error(O) > 0.1
