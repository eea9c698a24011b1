"""The medium laser-wakefield workload at its full size (2048 x 256, M=1,
2x2x8 particles per cell: 16,621,760 electrons and as many immobile ions,
FP64): two whole steps against the CPU oracle from the same seeded inputs,
then properties that hold at any size, checked at this one: the sort is the
stable argsort of the keys with consistent cell offsets, the deposits are
linear (doubling every weight doubles rho and J bit for bit), and the gather
+ push in a uniform electric field changes every pushed momentum by
FactorPush * E; and the mixed form against the FP64 form after 20 steps."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def c2():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import chimeracl_amd as CA
    from chimeracl_amd import config as C
    from chimeracl_amd.methods.generic_methods_hip import Communicator
    from oracle import oracle as O
    g, laser, c = C.grid_config("C2"), C.laser_config("C2"), C.CONFIGS["C2"]
    comm = Communicator(precision="f64")
    s = CA.Solver(dict(g), comm)
    CA.add_gausian_pulse(s, dict(laser))
    eo, io = C.species_config("C2", s.Args)
    arr = C.load_plasma(g, c["ppc"], c["dens"], seed=0, order="cell")
    e, i = CA.Particles(dict(eo), comm), CA.Particles(dict(io), comm)
    e.add_particles(source=C.ParticleSource(arr))
    i.add_particles(source=C.ParticleSource(arr))
    loop = CA.PIC_loop(solvers=[s], species=[e, i])
    os_ = O.OracleSolver(dict(g))
    O.add_gausian_pulse(os_, dict(laser))
    oe, oi = O.OracleParticles(dict(eo)), O.OracleParticles(dict(io))
    oe.add_particles(arr)
    oi.add_particles(arr)
    oloop = O.OracleLoop(os_, [oe, oi])
    del arr
    for _ in range(2):
        loop.step()
        oloop.step()
    torch.cuda.synchronize()
    return {"g": g, "s": s, "e": e, "os": os_, "oe": oe}


def test_c2_two_steps_vs_oracle(c2):
    from tests import simpair as S
    errs = S.compare_fields(c2["s"], c2["os"], S.real_names(1))
    for k in ("px", "py", "pz", "x", "y", "z", "g_inv"):
        errs[k] = S.rel_l2(c2["e"].host(k), c2["oe"].data[k])
    bad = {k: v for k, v in errs.items() if not v < 1e-10}
    assert not bad, bad
    d = c2["e"].DataDev
    assert np.array_equal(d["sort_indx"].cpu().numpy().view(np.uint32), c2["oe"].data["sort_indx"])
    assert np.array_equal(d["cell_offset"].cpu().numpy().view(np.uint32), c2["oe"].data["cell_offset"])
    assert c2["e"].Args["Np_stay"] == c2["oe"].Args["Np_stay"]


def test_c2_sort_invariants(c2):
    """sort_indx is a permutation that orders the keys stably; cell_offset[c]
    is the first sorted slot with key >= c; Np_stay = cell_offset[nbins-1]."""
    d, g = c2["e"].DataDev, c2["g"]
    key = d["indx_in_cell"].cpu().numpy().view(np.uint32)
    srt = d["sort_indx"].cpu().numpy().view(np.uint32).astype(np.int64)
    n = key.size
    assert np.array_equal(np.sort(srt), np.arange(n))
    ks = key[srt]
    assert np.all(ks[1:] >= ks[:-1])
    same = ks[1:] == ks[:-1]
    assert np.all(srt[1:][same] > srt[:-1][same])  # stable: ties in index order
    nbins = (g["Nx"] - 1) * (g["Nr"] - 1) + 1
    off = d["cell_offset"].cpu().numpy().view(np.uint32)
    assert np.array_equal(off, np.searchsorted(ks, np.arange(nbins + 1), side="left"))
    assert c2["e"].Args["Np_stay"] == off[nbins - 1]


def _planes(s, names):
    return {n: s.DataDev[n].clone() for n in names}


def test_c2_deposit_doubling(c2):
    """rho and J of the electrons with every weight doubled are exactly twice
    those with the weights as they are (the deposit is linear, each operation
    scales by 2 exactly)."""
    s, e = c2["s"], c2["e"]
    rho = ["rho_m0", "rho_m1"]
    J = [f"J{c}_m{m}" for c in "xyz" for m in (0, 1)]
    s.depose_charge([e])
    s.depose_currents([e])
    a = _planes(s, rho + J)
    w = e.DataDev["w"]
    w.mul_(2.0)
    try:
        s.depose_charge([e])
        s.depose_currents([e])
        b = _planes(s, rho + J)
    finally:
        w.mul_(0.5)
    for n in rho + J:
        assert torch.count_nonzero(a[n]) > 0, n
        assert torch.equal(b[n], 2.0 * a[n]), n


def test_c2_gather_uniform_field(c2):
    """E_m0 uniform, every other field zero: each pushed electron's momentum
    changes by FactorPush * E (Boris with B = 0, CIC weights summing to 1) to
    roundoff; the particles the reference's Np_stay test skips keep theirs;
    g_inv = 1/sqrt(1 + u^2) of the new momenta.  Runs last: it overwrites the
    fields and the momenta."""
    s, e = c2["s"], c2["e"]
    E = (0.37, -0.21, 0.55)
    for f in "EB":
        for c, v in zip("xyz", E):
            s.DataDev[f"{f}{c}_m0"].fill_(v if f == "E" else 0.0)
            s.DataDev[f"{f}{c}_m1"].zero_()
    u0 = {k: e.host(k) for k in ("px", "py", "pz")}
    s.gather_and_push([e])
    torch.cuda.synchronize()
    np_stay = e.Args["Np_stay"]
    srt = e.DataDev["sort_indx"].cpu().numpy().view(np.uint32)[:np_stay].astype(np.int64)
    pushed = np.zeros(e.Args["Np"], bool)
    pushed[srt[srt < np_stay]] = True
    F = e.Args["FactorPush"]
    u = {k: e.host(k) for k in ("px", "py", "pz")}
    for k, ek in zip(("px", "py", "pz"), E):
        du = u[k][pushed] - u0[k][pushed]
        assert np.max(np.abs(du - F * ek)) <= 1e-12 * abs(F * ek) + 1e-13 * np.max(np.abs(u0[k])), k
        assert np.array_equal(u[k][~pushed], u0[k][~pushed]), k
    g_inv = e.host("g_inv")[pushed]
    usq = u["px"][pushed] ** 2 + u["py"][pushed] ** 2 + u["pz"][pushed] ** 2
    assert np.max(np.abs(g_inv - 1 / np.sqrt(1 + usq))) < 1e-15
    assert pushed.sum() > 0.99 * e.Args["Np"]


def test_c2_mixed_vs_fp64_full_size():
    """The mixed form (FP32 particles and FFT, split-BF16 DHT) against the
    FP64 form at full size after 20 steps from the same seeded inputs, each
    field over all its components and modes: E and B within 1e-4 relative
    L2 (observed ~3e-6), J and the momenta within 1e-3 (~2e-5, ~1e-5), rho
    within 1e-2 (~2e-4: electrons and ions nearly cancel, so its relative
    error measures the cancellation)."""
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import chimeracl_amd as CA
    from chimeracl_amd import config as C
    from chimeracl_amd.methods.generic_methods_hip import Communicator
    from tests import simpair as S
    g, laser, c = C.grid_config("C2"), C.laser_config("C2"), C.CONFIGS["C2"]
    arr = C.load_plasma(g, c["ppc"], c["dens"], seed=0, order="cell")
    runs = {}
    for prec in ("f64", "mixed"):
        comm = Communicator(precision=prec)
        s = CA.Solver(dict(g), comm)
        CA.add_gausian_pulse(s, dict(laser))
        eo, io = C.species_config("C2", s.Args)
        e, i = CA.Particles(dict(eo), comm), CA.Particles(dict(io), comm)
        e.add_particles(source=C.ParticleSource(arr))
        i.add_particles(source=C.ParticleSource(arr))
        loop = CA.PIC_loop(solvers=[s], species=[e, i], graph=True, align_every=20)
        loop.run(20)
        torch.cuda.synchronize()
        runs[prec] = (s, e)
    (s64, e64), (s32, e32) = runs["f64"], runs["mixed"]
    errs = {}
    for fam in ("E", "B", "J", "rho"):
        names = [n for n in S.real_names(1) if n.startswith(fam)]
        a = np.concatenate([S.field_host(s64, n)[1:].ravel().astype(np.complex128) for n in names])
        b = np.concatenate([S.field_host(s32, n)[1:].ravel().astype(np.complex128) for n in names])
        errs[fam] = S.rel_l2(b, a)
    pid64, pid32 = e64.DataDev["pid"].cpu().numpy(), e32.DataDev["pid"].cpu().numpy()
    o64, o32 = np.argsort(pid64), np.argsort(pid32)  # both by the unreordered index
    for k in ("px", "py", "pz"):
        errs[k] = S.rel_l2(e32.host(k)[o32], e64.host(k)[o64])
    tol = {"E": 1e-4, "B": 1e-4, "J": 1e-3, "rho": 1e-2, "px": 1e-3, "py": 1e-3, "pz": 1e-3}
    print({k: f"{v:.2e}" for k, v in errs.items()})
    bad = {k: v for k, v in errs.items() if not v < tol[k]}
    assert not bad, bad
