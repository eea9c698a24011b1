"""Moving window (Frame, frame.py:4-64): the GPU domain fill reproduces the
oracle's bit for bit; a laser-plasma run with window shifts and plasma
injection matches the oracle's run."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _require_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


@pytest.mark.parametrize("precision", ["f64", "mixed"])
def test_fill_domain_bit_exact(precision):
    _require_gpu()
    import chimeracl_amd as CA
    from chimeracl_amd.methods.generic_methods_hip import Communicator
    from oracle import oracle as O
    cfg = {"Nppc": (2, 3, 4), "dx": 0.02, "dr": 0.1, "dt": 0.02, "dens": 0.002, "charge": -1,
           "seed": 5}
    comm = Communicator(precision=precision)
    g = CA.Particles(dict(cfg), comm)
    o = O.OracleParticles(dict(cfg))
    dom = {"Xmin": -1.37, "Xmax": -1.21, "Rmin": 0.0, "Rmax": 2.25, "px_c": 0.3}
    for _ in range(2):  # the counter advances between fills
        g.make_new_domain(dict(dom))
        o.make_new_domain(dict(dom))
        for k in ("x", "y", "z", "w", "px", "py", "pz", "g_inv"):
            a = g.DataDev[k + "_new"].cpu().numpy()
            b = o.data[k + "_new"]
            if precision == "mixed":
                b = b.astype(np.float32)
            assert a.shape == b.shape and np.array_equal(a, b), k
    assert g.Args["rng_counter"] == o.Args["rng_counter"]


@pytest.mark.parametrize("precision", ["f64", "mixed"])
def test_fill_domain_thermal_bit_exact(precision):
    """make_new_domain with thermal spreads (particles_methods_cl.py:114-147:
    p = p_c + dp randn per component): the GPU fill equals the oracle's bit
    for bit over two fills in a row (the counter advances by ncells + 4 Np),
    and the momenta have the requested means and spreads."""
    _require_gpu()
    import chimeracl_amd as CA
    from chimeracl_amd.methods.generic_methods_hip import Communicator
    from oracle import oracle as O
    cfg = {"Nppc": (2, 2, 4), "dx": 0.02, "dr": 0.1, "dt": 0.02, "dens": 0.002, "charge": -1,
           "seed": 9}
    comm = Communicator(precision=precision)
    g = CA.Particles(dict(cfg), comm)
    o = O.OracleParticles(dict(cfg))
    dom = {"Xmin": -2.0, "Xmax": -0.5, "Rmin": 0.0, "Rmax": 4.0, "px_c": 0.2, "dpx": 0.05,
           "dpy": 0.01, "dpz": 0.02}
    for _ in range(2):
        g.make_new_domain(dict(dom))
        o.make_new_domain(dict(dom))
        for k in ("x", "y", "z", "w", "px", "py", "pz", "g_inv"):
            a = g.DataDev[k + "_new"].cpu().numpy()
            b = o.data[k + "_new"]
            if precision == "mixed":
                b = b.astype(np.float32)
            assert a.shape == b.shape and np.array_equal(a, b), k
        assert g.Args["rng_counter"] == o.Args["rng_counter"]
    n = g.DataDev["px_new"].numel()
    for k, mu, sd in (("px", 0.2, 0.05), ("py", 0.0, 0.01), ("pz", 0.0, 0.02)):
        v = g.DataDev[k + "_new"].double().cpu().numpy()
        assert abs(v.mean() - mu) < 5 * sd / np.sqrt(n), k
        assert abs(v.std() - sd) < 5 * sd / np.sqrt(2 * n), k
    u = [g.DataDev[k + "_new"].double().cpu().numpy() for k in ("px", "py", "pz")]
    gi = g.DataDev["g_inv_new"].double().cpu().numpy()
    assert np.allclose(gi, 1 / np.sqrt(1 + u[0] ** 2 + u[1] ** 2 + u[2] ** 2),
                       rtol=1e-6 if precision == "mixed" else 1e-15)


def _frame_pair(precision="f64"):
    import chimeracl_amd as CA
    from chimeracl_amd import config as C
    from chimeracl_amd.methods.generic_methods_hip import Communicator
    from oracle import oracle as O
    from tests import simpair as S
    g = S.tiny_grid(Nx=64, Nr=24, M=1, damp=4)
    las = S.tiny_laser(g)
    comm = Communicator(precision=precision)
    solver = CA.Solver(dict(g), comm)
    CA.add_gausian_pulse(solver, dict(las))
    eo, io = C.species_config("C1", solver.Args, ppc=(2, 2, 2), dens=0.01)
    eo["Nppc"] = io["Nppc"] = (2, 2, 2)
    eo["dens"] = io["dens"] = 0.01
    arr = C.load_plasma(g, (2, 2, 2), 0.01, seed=0)
    eons, ions = CA.Particles(dict(eo), comm), CA.Particles(dict(io), comm)
    eons.add_particles(source=C.ParticleSource(arr))
    ions.add_particles(source=C.ParticleSource(arr))
    ions.Args["InjectorSource"] = eons
    prof = [{"coord": "x", "points": [-10.0, 0.5, 1.0, 10.0], "values": [1.0, 1.0, 1.5, 1.5]}]
    fr_cfg = {"Velocity": 1.0, "dt": g["dt"], "Steps": 3, "DensityProfiles": prof}
    frame = CA.Frame(dict(fr_cfg))
    loop = CA.PIC_loop(solvers=[solver], species=[eons, ions], frames=[frame])
    s = O.OracleSolver(dict(g))
    O.add_gausian_pulse(s, dict(las))
    e, i = O.OracleParticles(dict(eo)), O.OracleParticles(dict(io))
    e.add_particles(arr)
    i.add_particles(arr)
    i.Args["InjectorSource"] = e
    oloop = O.OracleLoop(s, [e, i], frames=[O.OracleFrame(dict(fr_cfg))])
    return (solver, eons, ions, loop), (s, e, i, oloop)


def test_moving_window_run_matches_oracle():
    _require_gpu()
    from tests import simpair as S
    (solver, eons, ions, loop), (s, e, i, oloop) = _frame_pair()
    np0 = eons.Args["Np"]
    for _ in range(9):
        loop.step()
        oloop.step()
    assert solver.Args["Xmin"] == s.Args["Xmin"]
    assert eons.Args["Np"] == e.Args["Np"] and ions.Args["Np"] == i.Args["Np"]
    assert eons.Args["Np"] != np0  # particles were dropped and injected
    assert eons.Args["right_lim"] == pytest.approx(e.Args["right_lim"], abs=1e-12)
    for k in ("x", "y", "z", "px", "py", "pz", "w"):
        assert S.rel_l2(eons.host(k), e.data[k]) < 1e-10, k
    errs = S.compare_fields(solver, s, S.real_names(1))
    assert max(errs.values()) < 1e-10, errs


@pytest.mark.parametrize("precision", ["f64", "mixed"])
def test_fill_beam_bit_exact(precision):
    _require_gpu()
    import chimeracl_amd as CA
    from chimeracl_amd.methods.generic_methods_hip import Communicator
    from oracle import oracle as O
    cfg = {"dx": 0.02, "dr": 0.1, "dt": 0.02, "charge": -1, "seed": 11}
    comm = Communicator(precision=precision)
    g = CA.Particles(dict(cfg), comm)
    o = O.OracleParticles(dict(cfg))
    beam = {"Np": 50000, "x_c": -1.0, "Lx": 0.2, "y_c": 0.0, "Ly": 0.5, "z_c": 0.1, "Lz": 0.5,
            "px_c": 100.0, "dpx": 5.0, "dpz": 1.0, "FullCharge": 3.0}
    for _ in range(2):
        g.make_new_beam(dict(beam))
        o.make_new_beam(dict(beam))
        for k in ("x", "y", "z", "px", "py", "pz", "w", "g_inv"):
            a = g.DataDev[k + "_new"].cpu().numpy()
            b = o.data[k + "_new"]
            if precision == "mixed":
                b = b.astype(np.float32)
            assert np.array_equal(a, b), k
    g.add_new_particles()
    assert g.Args["Np"] == 50000


def test_diagnostics_records(tmp_path, monkeypatch):
    """Diagnostics (diagnostics.py): one record per interval with the
    reference's paths; fields stacked [m0, Re m1, Im m1]; weights in pC."""
    _require_gpu()
    import chimeracl_amd as CA
    from chimeracl_amd import diagnostics as Dg
    monkeypatch.chdir(tmp_path)
    (solver, eons, ions, loop), _ = _frame_pair()
    diag = CA.Diagnostics({"Interval": 2, "ScalarFields": ["rho"], "VectorFields": ["E"],
                           "Species": {"Components": ["x", "px", "w"],
                                       "Selections": [("x", -0.5, None)]}},
                          solver, species=[eons], frame=loop.frames[0], path="diags")
    loop.diags = [diag]
    for _ in range(3):
        loop.step()
    ext = ".h5" if Dg.h5py is not None else ".npz"
    files = sorted(os.listdir(tmp_path / "diags"))
    assert files == ["000000000" + ext, "000000002" + ext]
    if ext == ".npz":
        rec = np.load(tmp_path / "diags" / files[-1])
        get = lambda k: rec[k.lstrip("/")]
    else:
        import h5py
        rec = h5py.File(tmp_path / "diags" / files[-1], "r")
        get = lambda k: rec[k][()]
    a = solver.Args
    assert int(get("/data/info/iteration")) == 2
    assert float(get("/data/info/FrameVelocity")) == 1.0
    assert np.array_equal(get("/data/info/Xgrid"), a["Xgrid"])
    ex = get("/data/fields/Ex")
    assert ex.shape == (3, a["Nr"], a["Nx"]) and ex.dtype == np.float32
    assert get("/data/fields/rho").shape == (3, a["Nr"], a["Nx"])
    x = get("/data/species/species_0/x")
    assert x.size > 0 and np.all(x > -0.5)
    w = get("/data/species/species_0/w")
    assert w.size == x.size and np.all(w > 0)
