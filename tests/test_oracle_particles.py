"""Pins of the CPU oracle's particle loops (particles_oracle.c): closed forms,
sum rules, textbook Boris identities, numpy brute force and an independent
pure-Python restatement of the reference kernels on tiny inputs.  CPU only."""
import numpy as np
import pytest


def _grid(oracle_mod, Nx=32, Nr=16, M=1):
    g = {"Xmin": -(Nx - 1) * 0.1, "Xmax": 0.0, "Nx": Nx, "Rmax": 0.1 * (Nr - 1.5), "Nr": Nr, "M": M,
         "dt": 0.1}
    return oracle_mod.OracleSolver(g)


def _parts(oracle_mod, s, arr, charge=-1, immobile=False):
    cfg = {"dx": s.Args["dx"], "dr": s.Args["dr"], "dt": s.Args["dt"], "charge": charge,
           "Nppc": (2, 2, 4), "dens": 0.01}
    if immobile:
        cfg["Immobile"] = True
    p = oracle_mod.OracleParticles(cfg)
    p.add_particles(arr)
    return p


def _rand(s, n, seed, margin=0.0):
    a = s.Args
    rng = np.random.default_rng(seed)
    x = rng.uniform(a["Xmin"] + margin, a["Xmax"] - margin, n)
    r = rng.uniform(0.001, a["Rgrid"][-1], n)
    th = rng.uniform(0, 2 * np.pi, n)
    return {"x": x, "y": r * np.sin(th), "z": r * np.cos(th), "px": rng.normal(0, 0.5, n),
            "py": rng.normal(0, 0.5, n), "pz": rng.normal(0, 0.5, n), "w": rng.uniform(0.5, 1.5, n),
            "g_inv": rng.uniform(0.3, 1, n)}


def test_push_closed_form(oracle_mod):
    s = _grid(oracle_mod)
    arr = {k: np.array([v]) for k, v in dict(x=-1.0, y=0.2, z=0.1, px=1.0, py=0.0, pz=-2.0, w=1.0,
                                             g_inv=1 / np.sqrt(6.0)).items()}
    p = _parts(oracle_mod, s, arr)
    p.push_coords("half")
    h = 0.5 * 0.1 / np.sqrt(6.0)
    assert p.data["x"][0] == pytest.approx(-1.0 + h)
    assert p.data["z"][0] == pytest.approx(0.1 - 2 * h)
    assert p.data["y"][0] == 0.2


def test_cell_keys_and_stable_sort(oracle_mod):
    s = _grid(oracle_mod, Nx=40, Nr=20)
    a = s.Args
    arr = _rand(s, 5000, 1)
    arr["x"][:300] = a["Xmin"] + a["dx"] * np.random.default_rng(2).integers(0, 40, 300)
    p = _parts(oracle_mod, s, arr)
    p.sort_parts(s)
    # numpy brute force of index_and_sum_in_cell (particles_generic.cl:105-124)
    r = np.sqrt(arr["y"] ** 2 + arr["z"] ** 2)
    ix = np.floor((arr["x"] - a["Xmin"]) * a["dx_inv"]).astype(int)
    ir = np.floor((r - a["Rmin"]) * a["dr_inv"]).astype(int)
    ok = (ix > 0) & (ix < 39 - 1) & (ir < 19 - 1) & (ir >= 0)
    key = np.where(ok, ix + ir * 39, 19 * 39).astype(np.uint32)
    assert np.array_equal(p.data["indx_in_cell"], key)
    assert np.array_equal(p.data["sort_indx"], np.argsort(key, kind="stable"))
    counts = np.bincount(key, minlength=19 * 39 + 1)
    assert np.array_equal(p.data["cell_offset"], np.concatenate([[0], np.cumsum(counts)]))
    assert p.Args["Np_stay"] == ok.sum()


def _deposit_raw(oracle_mod, s, p, vector=False):
    """Deposit without the axis post-processing (raw node sums)."""
    saved = s._postproc
    s._postproc = lambda name: None
    try:
        if vector:
            s.depose_currents([p])
        else:
            s.depose_charge([p])
    finally:
        s._postproc = saved


def test_deposit_sum_rule(oracle_mod):
    """Before 1/dV, the nodes of rho_m0 sum to sum(q w) over particles in
    valid cells, J_m0 to sum(q w g_inv u)."""
    s = _grid(oracle_mod)
    arr = _rand(s, 4000, 3)
    p = _parts(oracle_mod, s, arr)
    p.sort_parts(s)
    valid = p.data["sort_indx"][: p.Args["Np_stay"]]
    _deposit_raw(oracle_mod, s, p)
    assert s.f["rho_m0"].sum() == pytest.approx(-arr["w"][valid].sum(), rel=1e-12)
    _deposit_raw(oracle_mod, s, p, vector=True)
    for c in "xyz":
        tot = -(arr["w"] * arr["g_inv"] * arr["p" + c])[valid].sum()
        assert s.f[f"J{c}_m0"].sum() == pytest.approx(tot, rel=1e-11)


def test_deposit_single_particle_closed_form(oracle_mod):
    s = _grid(oracle_mod, M=2)
    a = s.Args
    ix, ir, fx, fr, phi = 7, 3, 0.3, 0.6, 0.7
    x = a["Xmin"] + (ix + fx) * a["dx"]
    r = a["Rmin"] + (ir + fr) * a["dr"]
    arr = {"x": np.array([x]), "y": np.array([r * np.sin(phi)]), "z": np.array([r * np.cos(phi)]),
           "w": np.array([2.0])}
    p = _parts(oracle_mod, s, arr, charge=1, immobile=True)
    p.sort_parts(s)
    _deposit_raw(oracle_mod, s, p)
    wgt = np.outer([1 - fr, fr], [1 - fx, fx]) * 2.0
    blk = s.f["rho_m0"][ir:ir + 2, ix:ix + 2]
    assert np.allclose(blk, wgt, rtol=1e-12)
    assert np.isclose(s.f["rho_m0"].sum(), 2.0)
    # mode m: weight ((y + i z)/r)^m = (i e^{-i phi})^m
    for m in (1, 2):
        ph = (1j * np.exp(-1j * phi)) ** m
        assert np.allclose(s.f[f"rho_m{m}"][ir:ir + 2, ix:ix + 2], wgt * ph, rtol=1e-12, atol=1e-15)


def test_deposit_mode2_ring(oracle_mod):
    """A ring of particles with weights 1 + cos(2 phi) deposits m=1 = 0 and
    m=2 = -(sum w cos 2phi e^{-2i phi}) = -N/2 per unit CIC weight."""
    s = _grid(oracle_mod, M=2)
    a = s.Args
    N = 64
    phi = 2 * np.pi * (np.arange(N) + 0.25) / N
    r = a["Rmin"] + 4.5 * a["dr"]
    x = np.full(N, a["Xmin"] + 10.5 * a["dx"])
    arr = {"x": x, "y": r * np.sin(phi), "z": r * np.cos(phi), "w": 1 + np.cos(2 * phi)}
    p = _parts(oracle_mod, s, arr, charge=1, immobile=True)
    p.sort_parts(s)
    _deposit_raw(oracle_mod, s, p)
    tot0 = s.f["rho_m0"].sum()
    assert tot0 == pytest.approx(N)
    assert np.abs(s.f["rho_m1"]).sum() < 1e-12 * N
    assert s.f["rho_m2"].sum() == pytest.approx(-N / 2, rel=1e-12)


def _set_fields(s, rng, uniform=None):
    for f in ("E", "B"):
        for c in "xyz":
            for m in range(s.M + 1):
                shp = s.f[f"{f}{c}_m{m}"].shape
                if uniform is not None:
                    v = uniform.get((f, c, m), 0.0)
                    s.f[f"{f}{c}_m{m}"][:] = v
                else:
                    s.f[f"{f}{c}_m{m}"][:] = rng.normal(size=shp) + (1j * rng.normal(size=shp) if m else 0)


def test_gather_pure_E_and_pure_B(oracle_mod):
    """Boris with a uniform E: u' = u + FactorPush*E.  With a uniform B:
    |u| is conserved and u turns about B by 2 atan(|t|), t = F/2 B gamma^-1."""
    s = _grid(oracle_mod)
    arr = _rand(s, 500, 4, margin=0.5)
    arr["x"] = np.clip(arr["x"], s.Args["Xmin"] + 0.2, s.Args["Xmax"] - 0.4)
    r = np.sqrt(arr["y"] ** 2 + arr["z"] ** 2)
    keep = r < 0.9
    arr = {k: v[keep] for k, v in arr.items()}
    arr["g_inv"] = 1 / np.sqrt(1 + arr["px"] ** 2 + arr["py"] ** 2 + arr["pz"] ** 2)
    p = _parts(oracle_mod, s, {k: v.copy() for k, v in arr.items()})
    p.sort_parts(s)
    E = (0.3, -0.2, 0.5)
    _set_fields(s, None, {("E", c, 0): E[k] for k, c in enumerate("xyz")})
    s.gather_and_push([p])
    F = p.Args["FactorPush"]
    for k, c in enumerate("xyz"):
        assert np.allclose(p.data["p" + c], arr["p" + c] + F * E[k], rtol=1e-12, atol=1e-12)
    # pure B
    p = _parts(oracle_mod, s, {k: v.copy() for k, v in arr.items()})
    p.sort_parts(s)
    B = np.array([0.4, 0.1, -0.3])
    _set_fields(s, None, {("B", c, 0): B[k] for k, c in enumerate("xyz")})
    s.gather_and_push([p])
    u0 = np.stack([arr["px"], arr["py"], arr["pz"]], 1)
    u1 = np.stack([p.data["px"], p.data["py"], p.data["pz"]], 1)
    assert np.allclose(np.linalg.norm(u1, axis=1), np.linalg.norm(u0, axis=1), rtol=1e-12)
    g = np.sqrt(1 + (u0 ** 2).sum(1))
    t = 0.5 * F * np.linalg.norm(B) / g
    bh = B / np.linalg.norm(B)
    perp0 = u0 - (u0 @ bh)[:, None] * bh
    perp1 = u1 - (u1 @ bh)[:, None] * bh
    cosang = (perp0 * perp1).sum(1) / (np.linalg.norm(perp0, axis=1) * np.linalg.norm(perp1, axis=1))
    assert np.allclose(cosang, np.cos(2 * np.arctan(t)), atol=1e-10)


def test_gather_mode1_reconstruction(oracle_mod):
    """A uniform m=1 field a gathers as 2 Re(a conj((y + i z)/r)) away from
    the axis (for r < dr/2 the ghost row holds -a: the m>=1 axis condition,
    grid_generic.cl:76-86)."""
    s = _grid(oracle_mod)
    arr = _rand(s, 300, 6, margin=0.5)
    r = np.sqrt(arr["y"] ** 2 + arr["z"] ** 2)
    keep = (r < 0.9) & (r > 0.5 * s.Args["dr"]) & (arr["x"] < s.Args["Xmax"] - 0.4) \
        & (arr["x"] > s.Args["Xmin"] + 0.2)
    arr = {k: v[keep] for k, v in arr.items()}
    arr["px"][:] = arr["py"][:] = arr["pz"][:] = 0
    p = _parts(oracle_mod, s, arr)
    p.sort_parts(s)
    a1 = 0.3 - 0.7j
    _set_fields(s, None, {("E", "x", 1): a1})
    s.gather_and_push([p])
    r = np.sqrt(arr["y"] ** 2 + arr["z"] ** 2)
    ex = 2 * (a1 * np.conj((arr["y"] + 1j * arr["z"]) / r)).real
    assert np.allclose(p.data["px"], p.Args["FactorPush"] * ex, rtol=1e-11, atol=1e-13)


def test_gather_skips_index_beyond_np_stay(oracle_mod):
    """Quirk of grid_deposit_m1.cl:367-368: a valid particle whose index is
    >= Np_stay is not pushed."""
    s = _grid(oracle_mod)
    arr = _rand(s, 200, 8, margin=0.5)
    arr["y"] *= 0.5
    arr["z"] *= 0.5  # all radii inside the valid cells
    arr["x"][5] = s.Args["Xmax"] + 5.0  # one particle outside -> Np_stay = Np - 1
    p = _parts(oracle_mod, s, {k: v.copy() for k, v in arr.items()})
    p.sort_parts(s)
    assert p.Args["Np_stay"] == 199
    _set_fields(s, None, {("E", "x", 0): 1.0})
    s.gather_and_push([p])
    assert p.data["px"][199] == arr["px"][199]
    assert p.data["px"][5] == arr["px"][5]
    moved = np.flatnonzero(p.data["px"] != arr["px"])
    assert 199 not in moved and len(moved) > 150


def test_profile_interpolant(oracle_mod):
    s = _grid(oracle_mod)
    arr = _rand(s, 1000, 9)
    p = _parts(oracle_mod, s, {k: v.copy() for k, v in arr.items()})
    xx = [s.Args["Xmin"] - 1, -2.0, -1.0, 1.0]
    ff = [0.0, 0.5, 2.0, 3.0]
    p.dens_profile(xx, ff)
    assert np.allclose(p.data["w"], arr["w"] * np.interp(arr["x"], xx, ff), rtol=1e-13)


# ------------------------------------------------------------------ pure Python restatement
def _py_deposit_scalar(s, arr, srt, off, charge):
    """Independent transcription of depose_scalar (grid_deposit_m1.cl:20-152)
    and its 4-colour launch (grid_methods_cl.py:66-69), loops in Python."""
    a = s.Args
    Nx, Nr = a["Nx"], a["Nr"]
    m0 = np.zeros((Nr, Nx))
    m1 = np.zeros((Nr, Nx), complex)
    for off_c in range(4):
        for g in range((Nr // 2) * (Nx // 2)):
            ir, ix = divmod(g, Nx // 2)
            ix, ir = 2 * ix + (off_c & 1), 2 * ir + (off_c >> 1)
            if not (ix > 0 and ix < Nx - 2 and ir < Nr - 2):
                continue
            c = ix + ir * (Nx - 1)
            if off[c] == off[c + 1]:
                continue
            acc0 = [[0.0, 0.0], [0.0, 0.0]]
            accr = [[0.0, 0.0], [0.0, 0.0]]
            acci = [[0.0, 0.0], [0.0, 0.0]]
            for ip in range(off[c], off[c + 1]):
                p = srt[ip]
                xp, yp, zp = float(arr["x"][p]), float(arr["y"][p]), float(arr["z"][p])
                wp = float(arr["w"][p]) * charge
                rp = np.sqrt(yp * yp + zp * zp)
                rinv = 1.0 / rp
                e0, e1 = yp * rinv, zp * rinv
                sX1 = (xp - a["Xmin"]) * a["dx_inv"] - ix
                sX0 = 1.0 - sX1
                sR1 = (rp - a["Rmin"]) * a["dr_inv"] - ir
                sR0 = 1.0 - sR1
                sX0 *= wp
                sX1 *= wp
                Cc = [[sR0 * sX0, sR0 * sX1], [sR1 * sX0, sR1 * sX1]]
                for i in range(2):
                    for j in range(2):
                        acc0[i][j] += Cc[i][j]
                        accr[i][j] += Cc[i][j] * e0
                        acci[i][j] += Cc[i][j] * e1
            for i in range(2):
                for j in range(2):
                    m0[ir + i, ix + j] = m0[ir + i, ix + j] + acc0[i][j]
                    v = m1[ir + i, ix + j]
                    m1[ir + i, ix + j] = complex(v.real + accr[i][j], v.imag + acci[i][j])
    return m0, m1


def test_c_deposit_equals_python_restatement(oracle_mod):
    s = _grid(oracle_mod, Nx=16, Nr=10)
    arr = _rand(s, 400, 11)
    p = _parts(oracle_mod, s, arr, charge=-1, immobile=True)
    p.sort_parts(s)
    _deposit_raw(oracle_mod, s, p)
    m0, m1 = _py_deposit_scalar(s, arr, p.data["sort_indx"], p.data["cell_offset"], -1.0)
    assert np.array_equal(s.f["rho_m0"], m0)
    assert np.array_equal(s.f["rho_m1"], m1)


def test_domain_fill_generator_known_answers(oracle_mod):
    """splitmix64's published first outputs for seed 0 (Steele, Lea, Flood
    2014; the reference implementation's test vector)."""
    O = oracle_mod
    words = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    u = O.rng_u01(0, np.arange(3))
    assert np.array_equal(u, np.array([(w >> 11) / 2.0 ** 53 for w in words]))


def test_det_sincos_matches_libm(oracle_mod):
    """The IEEE-only sin/cos of the domain fill agrees with numpy's to 2 ulp
    on the angles it is used for ([0, 4 pi)) and beyond."""
    O = oracle_mod
    a = np.concatenate([np.linspace(0, 4 * np.pi, 200001), np.random.default_rng(3).uniform(-50, 50, 100000),
                        np.array([0.0, np.pi / 4, np.pi / 2, np.pi, 1.5 * np.pi, 2 * np.pi])])
    s, c = O.det_sincos(a)
    es, ec = np.sin(a), np.cos(a)
    tol = lambda ref: 2 * np.spacing(np.maximum(np.abs(ref), 1e-16)) + 2e-16
    assert np.all(np.abs(s - es) <= tol(es))
    assert np.all(np.abs(c - ec) <= tol(ec))
    assert np.all(s * s + c * c - 1 < 4e-16)


def test_make_new_domain_geometry(oracle_mod):
    """fill_grid: regular sub-positions in every local cell, one theta offset
    per cell, w = w0 r; counters advance so two fills differ."""
    O = oracle_mod
    p = O.OracleParticles({"Nppc": (2, 2, 4), "dx": 0.02, "dr": 0.1, "dt": 0.02, "dens": 0.002,
                           "charge": -1})
    dom = {"Xmin": -0.4, "Xmax": -0.2, "Rmin": 0.0, "Rmax": 1.05}
    p.make_new_domain(dom)
    d = p.data
    Nx_loc = int(np.ceil((-0.2 + 0.4) / 0.02) + 1)
    Nr_loc = int(np.round(1.05 / 0.1) + 1)
    n = (Nx_loc - 1) * (Nr_loc - 1) * 16
    assert d["x_new"].size == n
    x, r = d["x_new"], np.hypot(d["y_new"], d["z_new"])
    cell = np.arange(n) // 16
    ix, ir = cell % (Nx_loc - 1), cell // (Nx_loc - 1)
    sub = np.arange(n) % 2
    assert np.allclose(x, -0.4 + 0.02 * ix + (0.5 + sub) * 0.01, atol=1e-13)
    sub_r = (np.arange(n) // 2) % 2
    assert np.allclose(r, 0.1 * ir + (0.5 + sub_r) * 0.05, atol=1e-13)
    assert np.allclose(d["w_new"], p.Args["w0"] * r, rtol=1e-13)
    th = np.arctan2(d["y_new"], d["z_new"]).reshape(-1, 4, 2, 2)[:, :, 0, 0]
    dth = np.mod(np.diff(th, axis=1), 2 * np.pi)
    assert np.allclose(dth, np.pi / 2, atol=1e-12)
    assert np.all(d["g_inv_new"] == 1) and np.all(d["px_new"] == 0)
    first = d["y_new"].copy()
    p.make_new_domain(dom)
    assert not np.array_equal(first, p.data["y_new"])


def test_det_log_matches_libm(oracle_mod):
    O = oracle_mod
    x = np.concatenate([np.random.default_rng(4).random(200000), np.logspace(-300, 300, 20001),
                        np.array([1.0, 0.5, 2.0, np.e, 1e-16, 1 - 2 ** -53])])
    x = x[x > 0]
    v, ref = O.det_log(x), np.log(x)
    assert np.all(np.abs(v - ref) <= 2 * np.spacing(np.abs(ref)) + 1e-300)
    assert O.det_log(np.array([1.0]))[0] == 0.0


def test_beam_statistics(oracle_mod):
    """make_new_beam: means and spreads as requested (Box-Muller normals)."""
    O = oracle_mod
    p = O.OracleParticles({"dx": 0.02, "dr": 0.1, "dt": 0.02, "charge": -1, "seed": 9})
    Np = 400000
    beam = {"Np": Np, "x_c": -1.0, "Lx": 0.2, "y_c": 0.0, "Ly": 0.5, "z_c": 0.1, "Lz": 0.5,
            "px_c": 100.0, "dpx": 5.0, "dpz": 1.0, "FullCharge": 3.0}
    p.make_new_beam(beam)
    d = p.data
    se = 5 / np.sqrt(Np)
    for k, mu, sig in (("x", -1.0, 0.2), ("y", 0.0, 0.5), ("z", 0.1, 0.5), ("px", 100.0, 5.0),
                       ("py", 0.0, 0.0), ("pz", 0.0, 1.0)):
        v = d[k + "_new"]
        assert abs(v.mean() - mu) <= se * max(sig, 1e-12) + 1e-14, k
        assert abs(v.std() - sig) <= se * sig + 1e-14, k
    assert np.allclose(d["w_new"], 3.0 / Np)
    u2 = d["px_new"] ** 2 + d["py_new"] ** 2 + d["pz_new"] ** 2
    assert np.allclose(d["g_inv_new"], 1 / np.sqrt(1 + u2))
    # two normals of a pair are uncorrelated
    assert abs(np.corrcoef(d["x_new"], d["y_new"])[0, 1]) < 5 / np.sqrt(Np)


def test_make_new_domain_thermal_statistics(oracle_mod):
    """make_new_domain with thermal spreads (particles_methods_cl.py:114-147):
    each momentum component is p_c + dp N(0, 1); the sample means and
    standard deviations match within 5 standard errors, the components are
    uncorrelated, g_inv = 1/sqrt(1 + p^2), positions and weights are those of
    the cold fill, and the counter advances by ncells + 4 Np."""
    O = oracle_mod
    cfg = {"Nppc": (2, 2, 4), "dx": 0.02, "dr": 0.1, "dt": 0.02, "dens": 0.002, "charge": -1,
           "seed": 3}
    dom = {"Xmin": -2.0, "Xmax": 0.0, "Rmin": 0.0, "Rmax": 4.0}
    cold = O.OracleParticles(dict(cfg))
    cold.make_new_domain(dict(dom))
    p = O.OracleParticles(dict(cfg))
    p.make_new_domain(dict(dom, px_c=0.5, dpx=0.1, dpz=0.03))
    d = p.data
    n = d["px_new"].size
    ncells = n // 16
    for k in ("x", "y", "z", "w"):
        assert np.array_equal(d[k + "_new"], cold.data[k + "_new"]), k
    se = 5 / np.sqrt(n)
    for k, mu, sig in (("px", 0.5, 0.1), ("py", 0.0, 0.0), ("pz", 0.0, 0.03)):
        v = d[k + "_new"]
        assert abs(v.mean() - mu) <= se * sig + 1e-15, k
        assert abs(v.std() - sig) <= se * sig / np.sqrt(2) * 2 + 1e-15, k
    assert abs(np.corrcoef(d["px_new"], d["pz_new"])[0, 1]) < 5 / np.sqrt(n)
    u2 = d["px_new"] ** 2 + d["py_new"] ** 2 + d["pz_new"] ** 2
    assert np.allclose(d["g_inv_new"], 1 / np.sqrt(1 + u2), rtol=1e-15)
    assert p.Args["rng_counter"] == ncells + 4 * n
    assert cold.Args["rng_counter"] == ncells
