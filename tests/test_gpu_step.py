"""Whole PIC steps on the GPU vs the CPU oracle from the same seeded inputs
(relative L2 <= 1e-10 for E, B, J, rho and the momenta), the reference's two
known-answer checks on the GPU path, and HIP-graph replay == eager."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

TOL = 1e-10


@pytest.fixture(scope="module")
def S():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from tests import simpair
    return simpair


def _check(S, gpu, orc, M, tol=TOL):
    errs = S.compare_fields(gpu["solver"], orc["solver"], S.real_names(M))
    for k in ("px", "py", "pz", "x", "y", "z"):
        errs[k] = S.rel_l2(gpu["eons"].host(k), orc["eons"].data[k])
    bad = {k: v for k, v in errs.items() if not v < tol}
    assert not bad, bad
    return errs


@pytest.mark.parametrize("M", [0, 1, 2])
def test_laser_init(S, M):
    g = S.tiny_grid(Nx=128, Nr=32, M=M)
    gpu, orc = S.make_pair(g, S.tiny_laser(g))
    errs = S.compare_fields(gpu["solver"], orc["solver"], S.real_names(M, flds=("E", "B"), scal=()))
    assert max(errs.values()) < 1e-12, errs


@pytest.mark.parametrize("M", [0, 1, 2])
def test_steps_tiny(S, M):
    g = S.tiny_grid(Nx=64, Nr=24, M=M)
    gpu, orc = S.make_pair(g, S.tiny_laser(g))
    for it in range(6):
        gpu["loop"].step()
        orc["loop"].step()
    _check(S, gpu, orc, M)
    srt_g = gpu["eons"].DataDev["sort_indx"].cpu().numpy().view(np.uint32)
    assert np.array_equal(srt_g, orc["eons"].data["sort_indx"])
    assert gpu["eons"].Args["Np_stay"] == orc["eons"].Args["Np_stay"]


@pytest.mark.parametrize("graph", [False, True])
def test_steps_aligned(S, graph):
    """align_every=2: the mobile species are permuted into their sorted order
    at the start of every second step; the oracle loop does the same and the
    steps still agree (fields, momenta and positions index by index, the
    sort)."""
    g = S.tiny_grid(Nx=64, Nr=24, M=1)
    gpu, orc = S.make_pair(g, S.tiny_laser(g))
    gpu["loop"].align_every = orc["loop"].align_every = 2
    gpu["loop"].graph = graph
    for it in range(7):
        gpu["loop"].step()
        orc["loop"].step()
    _check(S, gpu, orc, 1)
    srt_g = gpu["eons"].DataDev["sort_indx"].cpu().numpy().view(np.uint32)
    assert np.array_equal(srt_g, orc["eons"].data["sort_indx"])
    assert gpu["eons"].Args["Np_stay"] == orc["eons"].Args["Np_stay"]
    # a reorder leaves the storage in cell order and the sort the identity
    from chimeracl_amd import config as C
    gpu["eons"].reorder_parts()
    orc["eons"].reorder_parts()
    e = gpu["eons"]
    srt = e.DataDev["sort_indx"].cpu().numpy().view(np.uint32)
    assert np.array_equal(srt, np.arange(srt.size))
    keys = C.cell_keys(g, e.host("x"), e.host("y"), e.host("z"))
    assert np.all(np.diff(keys) >= 0)
    gpu["loop"].step()
    orc["loop"].step()
    _check(S, gpu, orc, 1)


def test_aligned_matches_unaligned_oracle(S):
    """Reordered storage (align_every=2, with pid) against the oracle run
    without any reordering: the same particles are skipped by the gather
    (50 electrons start outside the domain, so Np_stay < Np and the
    reference's skip of indices >= Np_stay applies), and fields and momenta
    agree to roundoff (only the summation order within a cell differs)."""
    g = S.tiny_grid(Nx=64, Nr=24, M=1)
    gpu, orc = S.make_pair(g, S.tiny_laser(g))
    n_out = 50
    gpu["eons"].DataDev["x"][:n_out] = g["Xmax"] + 1.0
    orc["eons"].data["x"][:n_out] = g["Xmax"] + 1.0
    gpu["loop"].align_every = 2
    for it in range(7):
        gpu["loop"].step()
        orc["loop"].step()
    e, o = gpu["eons"], orc["eons"]
    assert e.Args["Np_stay"] < e.Args["Np"] and "pid" in e.DataDev
    errs = S.compare_fields(gpu["solver"], orc["solver"], S.real_names(1))
    pid = e.DataDev["pid"].cpu().numpy().astype(np.int64)
    assert np.array_equal(np.sort(pid), np.arange(pid.size))
    for k in ("px", "py", "pz", "x", "y", "z"):
        errs[k] = S.rel_l2(e.host(k), o.data[k][pid])
    bad = {k: v for k, v in errs.items() if not v < TOL}
    assert not bad, bad


@pytest.mark.parametrize("graph", [False, True])
def test_run_equals_steps(S, graph):
    """PIC_loop.run(n) (each gather also applies the next first half push,
    graph variants) ends in the state of n calls of step(), bit for bit, with
    a reorder inside the run (align_every=3) and electrons outside the
    domain; the oracle agrees."""
    g = S.tiny_grid(Nx=64, Nr=24, M=1)
    a, orc = S.make_pair(g, S.tiny_laser(g))
    b, _ = S.make_pair(g, S.tiny_laser(g), with_oracle=False)
    for sim in (a, b):
        sim["eons"].DataDev["x"][:40] = g["Xmax"] + 1.0
        sim["loop"].align_every = 3
        sim["loop"].graph = graph
    orc["eons"].data["x"][:40] = g["Xmax"] + 1.0
    orc["loop"].align_every = 3
    for it in range(8):
        a["loop"].step()
        orc["loop"].step()
    b["loop"].step()
    b["loop"].run(7)
    torch.cuda.synchronize()
    assert b["loop"].it == 8 and not b["loop"]._pre
    for n in S.real_names(1):
        assert torch.equal(a["solver"].DataDev[n], b["solver"].DataDev[n]), n
    for k in ("x", "y", "z", "px", "py", "pz", "g_inv", "w", "pid", "sort_indx", "cell_offset"):
        assert torch.equal(a["eons"].DataDev[k], b["eons"].DataDev[k]), k
    _check(S, b, orc, 1)


def test_steps_c1_shape(S):
    """512 x 128, M=1, 2x2x4 ppc (the small laser-wakefield configuration),
    first 3 steps."""
    from chimeracl_amd import config as C
    g = C.grid_config("C1")
    gpu, orc = S.make_pair(g, C.laser_config("C1"))
    for it in range(3):
        gpu["loop"].step()
        orc["loop"].step()
    errs = _check(S, gpu, orc, 1)
    print({k: f"{v:.2e}" for k, v in errs.items()})


def test_graph_replay_equals_eager(S):
    g = S.tiny_grid(Nx=64, Nr=24, M=1)
    a, _ = S.make_pair(g, S.tiny_laser(g), with_oracle=False)
    b, _ = S.make_pair(g, S.tiny_laser(g), with_oracle=False)
    b["loop"].graph = True
    for it in range(5):
        a["loop"].step()
        b["loop"].step()
        torch.cuda.synchronize()
        for n in S.real_names(1):
            assert torch.equal(a["solver"].DataDev[n], b["solver"].DataDev[n]), (it, n)
        for k in ("x", "px", "g_inv"):
            assert torch.equal(a["eons"].DataDev[k], b["eons"].DataDev[k]), (it, k)
    assert b["loop"]._graphs


def test_round_trip_kat(S):
    """examples/test_transformer.py:35-43: forward then backward transform of
    a deposited beam density returns it (max error relative to max)."""
    import chimeracl_amd as CA
    from chimeracl_amd import config as C
    from chimeracl_amd.methods.generic_methods_hip import Communicator
    comm = Communicator(precision="f64")
    g = {"Xmin": -1.0, "Xmax": 1.0, "Nx": 1024, "Rmin": 0, "Rmax": 1.0, "Nr": 200, "M": 1}
    solver = CA.Solver(dict(g), comm)
    rng = np.random.default_rng(0)
    n = 2_000_000
    arr = {"x": rng.normal(0, 0.3, n), "y": rng.normal(0.2, 0.3, n), "z": rng.normal(0.2, 0.3, n),
           "w": np.full(n, 1.0 / n)}
    p = CA.Particles({"charge": -1, "Immobile": True, "dx": solver.Args["dx"], "dr": solver.Args["dr"]}, comm)
    p.add_particles(source=C.ParticleSource(arr))
    p.sort_parts(solver)
    solver.depose_charge([p])
    t0 = solver.DataDev["rho_m0"].cpu().numpy().copy()
    t1 = solver.DataDev["rho_m1"].cpu().numpy().copy()
    solver.fb_transform(scals=["rho"], dir=0)
    solver.DataDev["rho_m0"].zero_()
    solver.DataDev["rho_m1"].zero_()
    solver.fb_transform(scals=["rho"], dir=1)
    r0 = solver.DataDev["rho_m0"].cpu().numpy()
    r1 = solver.DataDev["rho_m1"].cpu().numpy()
    err = (np.abs(r0 - t0)[1:] / np.abs(t0[1:]).max() + np.abs(r1 - t1)[1:] / np.abs(t1[1:]).max()).max()
    assert err < 1e-11, err


def test_laser_group_velocity_kat(S):
    """examples/test_laser_veloc.py:9-59 on the GPU path: the laser centroid
    moves at 1 - (2 pi R)^-2 (r-weighted) and 1 - 2 (2 pi R)^-2 (on axis).
    Same grid as the reference's script (dx 0.05, dr 0.3, dt 0.05, R 12),
    Nx = 2048 (power of two) instead of 2000.  The reference prints the mean
    relative deviation without a threshold; 10% is the bar here."""
    import chimeracl_amd as CA
    from chimeracl_amd.methods.generic_methods_hip import Communicator
    comm = Communicator(precision="f64")
    dx, dr, dt = 0.05, 0.3, 0.05
    Nx = 2048
    g = {"Xmin": -Nx * dx / 2, "Xmax": -Nx * dx / 2 + (Nx - 1) * dx, "Rmax": 50.0, "dt": dt,
         "Nx": Nx, "Nr": int(50.0 / dr) // 2 * 2 + 1, "M": 0}
    g["Rmax"] = dr * (g["Nr"] - 1.5)
    laser = {"k0": 1.0, "a0": 1.0, "x0": 0, "Lx": 12.0, "R": 12.0, "x_foc": 20.0}
    solver = CA.Solver(dict(g), comm)
    CA.add_gausian_pulse(solver, dict(laser))
    Xg, Rg = solver.Args["Xgrid"], solver.Args["Rgrid"]
    c1, c2 = [], []
    for i in range(100):
        solver.push_fields()
        solver.fb_transform(scals=["Ez"], dir=1)
        v = solver.DataDev["Ez_m0"].cpu().numpy()
        Px = (Rg[1:, None] * v[1:, :] ** 2).sum(0)
        c1.append((Xg[None, :] * Px).sum() / Px.sum())
        Px = v[1, :] ** 2
        c2.append((Xg[None, :] * Px).sum() / Px.sum())
    th = (2 * np.pi * laser["R"]) ** -2
    e1 = np.abs(1 - np.diff(c1) / dt - th).mean() / th
    e2 = np.abs(1 - np.diff(c2) / dt - 2 * th).mean() / (2 * th)
    assert e1 < 0.1 and e2 < 0.1, (e1, e2)
