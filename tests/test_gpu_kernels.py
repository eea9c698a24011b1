"""GPU kernels vs the CPU oracle on identical inputs (through the C ABI, via
the reference-API wrappers).  Integer work and the FP64 particle kernels are
bit-exact; the transforms and spectral operators agree to <= 1e-12 relative
L2 (summation order of the FFT / MFMA products differs)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _require_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


@pytest.fixture(scope="module")
def env():
    _require_gpu()
    from tests import simpair as S
    return S


def _rand_particles(g, n, seed, inside_frac=0.97):
    rng = np.random.default_rng(seed)
    dx = (g["Xmax"] - g["Xmin"]) / (g["Nx"] - 1)
    dr = g["Rmax"] / (g["Nr"] - 1.5)
    x = rng.uniform(g["Xmin"] - 2 * dx, g["Xmax"] + 2 * dx, n)
    r = rng.uniform(0.01 * dr, (g["Nr"] - 0.5) * dr, n)
    th = rng.uniform(0, 2 * np.pi, n)
    # some particles exactly on cell boundaries
    k = n // 20
    x[:k] = g["Xmin"] + dx * rng.integers(0, g["Nx"], k)
    return {"x": x, "y": r * np.sin(th), "z": r * np.cos(th),
            "px": rng.normal(0, 0.3, n), "py": rng.normal(0, 0.3, n), "pz": rng.normal(0, 0.3, n),
            "w": rng.uniform(0.5, 1.5, n) * 1e-4, "g_inv": rng.uniform(0.5, 1.0, n)}


def _pair_particles(env, g, n, seed, M):
    import chimeracl_amd as CA
    from chimeracl_amd import config as C
    from chimeracl_amd.methods.generic_methods_hip import Communicator
    from oracle import oracle as O
    comm = Communicator(precision="f64")
    solver = CA.Solver(dict(g), comm)
    cfg = {"Nppc": (2, 2, 4), "dx": solver.Args["dx"], "dr": solver.Args["dr"], "dt": solver.Args["dt"],
           "dens": 0.002, "charge": -1}
    arr = _rand_particles(g, n, seed)
    p = CA.Particles(dict(cfg), comm)
    p.add_particles(source=C.ParticleSource(arr))
    s = O.OracleSolver(dict(g))
    q = O.OracleParticles(dict(cfg))
    q.add_particles(arr)
    return solver, p, s, q


GRIDS = [dict(Nx=64, Nr=24, M=1), dict(Nx=128, Nr=40, M=0), dict(Nx=64, Nr=32, M=2)]


@pytest.mark.parametrize("gi", range(len(GRIDS)))
def test_push_sort_deposit_bit_exact(env, gi):
    g = env.tiny_grid(**GRIDS[gi])
    solver, p, s, q = _pair_particles(env, g, 20000, 10 + gi, g["M"])
    p.push_coords("half")
    q.push_coords("half")
    for k in ("x", "y", "z"):
        assert np.array_equal(p.host(k), q.data[k]), k
    p.sort_parts(solver)
    q.sort_parts(s)
    key = p.DataDev["indx_in_cell"].cpu().numpy().view(np.uint32)
    assert np.array_equal(key, q.data["indx_in_cell"])
    assert np.array_equal(p.DataDev["cell_offset"].cpu().numpy().view(np.uint32), q.data["cell_offset"])
    srt = p.DataDev["sort_indx"].cpu().numpy().view(np.uint32)
    assert np.array_equal(srt, np.argsort(key, kind="stable").astype(np.uint32))
    assert np.array_equal(srt, q.data["sort_indx"])
    assert p.Args["Np_stay"] == q.Args["Np_stay"]
    # deposits: bit-exact, including the node colour order
    solver.depose_charge([p])
    s.depose_charge([q])
    solver.depose_currents([p])
    s.depose_currents([q])
    for name in env.real_names(g["M"], flds=("J",)):
        a = env.field_host(solver, name)
        b = s.f[name]
        assert np.array_equal(a, b), name


@pytest.mark.parametrize("gi", range(len(GRIDS)))
def test_deposit_push_fused_matches_separate(env, gi):
    """depose_currents(push_half=True) equals depose_currents, then
    push_coords('half', grid), bit for bit: J, the pushed positions (also of
    the particles outside the domain), the keys and the next sort."""
    import chimeracl_amd as CA
    from chimeracl_amd import config as C
    g = env.tiny_grid(**GRIDS[gi])
    solver, p, s, q = _pair_particles(env, g, 30000, 40 + gi, g["M"])
    arr = _rand_particles(g, 30000, 40 + gi)
    p2 = CA.Particles({"Nppc": (2, 2, 4), "dx": solver.Args["dx"], "dr": solver.Args["dr"],
                       "dt": solver.Args["dt"], "dens": 0.002, "charge": -1}, solver.comm)
    p2.add_particles(source=C.ParticleSource(arr))
    out = {}
    for tag, parts, fused in (("sep", p, False), ("fused", p2, True)):
        parts.sort_parts(solver)
        solver.depose_currents(species=[parts], push_half=fused)
        if not fused:
            parts.push_coords(mode="half", grid=solver)
        parts.sort_parts(solver)
        d = parts.DataDev
        out[tag] = {n: env.field_host(solver, n).copy() for n in env.real_names(g["M"], flds=("J",))}
        for k in ("x", "y", "z"):
            out[tag][k] = parts.host(k)
        for k in ("indx_in_cell", "sort_indx", "cell_offset"):
            out[tag][k] = d[k].cpu().numpy().view(np.uint32).copy()
        out[tag]["Np_stay"] = parts.Args["Np_stay"]
    assert out["sep"]["Np_stay"] < 30000  # some particles outside the domain
    for k, v in out["sep"].items():
        assert np.array_equal(v, out["fused"][k]), k


@pytest.mark.parametrize("gi", range(len(GRIDS)))
def test_gather_push_bit_exact(env, gi):
    g = env.tiny_grid(**GRIDS[gi])
    solver, p, s, q = _pair_particles(env, g, 20000, 20 + gi, g["M"])
    rng = np.random.default_rng(5 + gi)
    for f in ("E", "B"):
        for c in "xyz":
            for m in range(g["M"] + 1):
                name = f"{f}{c}_m{m}"
                shp = s.f[name].shape
                v = rng.normal(0, 0.5, shp) if m == 0 else rng.normal(0, 0.5, shp) + 1j * rng.normal(0, 0.5, shp)
                s.f[name][:] = v
                solver.DataDev[name].copy_(torch.from_numpy(np.ascontiguousarray(v)))
    p.sort_parts(solver)
    q.sort_parts(s)
    solver.gather_and_push([p])
    s.gather_and_push([q])
    for k in ("px", "py", "pz", "g_inv"):
        assert np.array_equal(p.host(k), q.data[k]), k
    for name in env.real_names(g["M"], flds=("E", "B"), scal=()):
        assert np.array_equal(env.field_host(solver, name), s.f[name]), name


def test_align_multi_matches_numpy(env):
    """chm_align_multi permutes 8- and 4-byte arrays in one launch,
    dst[a][i] = src[a][sort_indx[i]], bit for bit as numpy's take and as
    chm_align_elems per array; a destination that is also a source, or
    another destination, is refused without a launch."""
    import ctypes
    from chimeracl_amd import _lib
    n = 100_003
    rng = np.random.default_rng(7)
    perm = rng.permutation(n).astype(np.uint32)
    host = [rng.normal(size=n), rng.integers(0, 2**31, n).astype(np.int32),
            rng.normal(size=n).astype(np.float32), rng.normal(size=n),
            rng.integers(0, 2**62, n).astype(np.int64)]
    dev = torch.device("cuda")
    srt = torch.from_numpy(perm.view(np.int32)).to(dev)
    src = [torch.from_numpy(h).to(dev) for h in host]
    dst = [torch.empty_like(t) for t in src]
    P = ctypes.c_void_p
    k = len(src)
    sp = (P * k)(*(t.data_ptr() for t in src))
    dp = (P * k)(*(t.data_ptr() for t in dst))
    eb = (ctypes.c_int32 * k)(*(t.element_size() for t in src))
    st = torch.cuda.current_stream().cuda_stream
    _lib.call("chm_align_multi", srt.data_ptr(), n, k, ctypes.addressof(sp), ctypes.addressof(dp),
              ctypes.addressof(eb), st)
    torch.cuda.synchronize()
    for h, t in zip(host, dst):
        assert np.array_equal(t.cpu().numpy(), h[perm])
    one = torch.empty_like(src[0])
    _lib.call("chm_align_elems", srt.data_ptr(), n, src[0].data_ptr(), one.data_ptr(), 8, st)
    torch.cuda.synchronize()
    assert torch.equal(one, dst[0])
    bad = (P * 2)(dst[0].data_ptr(), src[0].data_ptr())
    assert _lib.lib.chm_align_multi(srt.data_ptr(), n, 2, ctypes.addressof(sp), ctypes.addressof(bad),
                                    ctypes.addressof(eb), st) != 0
    bad2 = (P * 2)(dst[0].data_ptr(), dst[0].data_ptr())
    assert _lib.lib.chm_align_multi(srt.data_ptr(), n, 2, ctypes.addressof(sp), ctypes.addressof(bad2),
                                    ctypes.addressof(eb), st) != 0


def test_gather_stale_sorted_keys_match_per_particle(env):
    """The FP64 gather of the sorted slots given the sort's keys
    (chm_sort_sorted_keys) after the particles moved by up to 1.5 cells since
    that sort, so that the keys no longer match the positions: bit-identical
    to the per-particle gather (sort_indx = NULL), which reads each
    particle's fields at its own position (the staged fields are used only
    for particles inside the staged cells, the others read the planes).  All
    particles are kept inside the valid cells before and after the move, so
    that both visit every particle."""
    import ctypes
    from chimeracl_amd import _lib
    from chimeracl_amd.methods.generic_methods_hip import ptr
    g = env.tiny_grid(Nx=64, Nr=24, M=2)
    solver, p, s, q = _pair_particles(env, g, 20000, 91, 2)
    a, n = solver.Args, p.Args["Np"]
    rng = np.random.default_rng(91)
    for f in ("E", "B"):
        for c in "xyz":
            for m in range(3):
                t = solver.DataDev[f"{f}{c}_m{m}"]
                v = rng.normal(0, 0.5, tuple(t.shape))
                t.copy_(torch.from_numpy(v).to(t.dtype))
    d = p.DataDev
    dx, dr, xmin, rmin = a["dx"], a["dr"], a["Xmin"], a["Rmin"]
    xlo, xhi = xmin + 1.2 * dx, xmin + (a["Nx"] - 2.2) * dx
    rlo, rhi = 0.05 * dr, rmin + (a["Nr"] - 2.2) * dr

    def place(shift):
        x = np.clip(d["x"].cpu().numpy() + shift * rng.uniform(-1.5, 1.5, n) * dx, xlo, xhi)
        y, z = d["y"].cpu().numpy(), d["z"].cpu().numpy()
        r = np.hypot(y, z)
        rn = np.clip(r + shift * rng.uniform(-1.5, 1.5, n) * dr, rlo, rhi)
        fac = np.where(r > 0, rn / np.where(r > 0, r, 1.0), 1.0)
        for k, v in (("x", x), ("y", y * fac), ("z", z * fac)):
            d[k].copy_(torch.from_numpy(v))

    place(0.0)          # every particle in a valid cell: Np_stay = Np
    p.flag_sorted = False
    p.sort_parts(solver)
    assert p.Args["Np_stay"] == n
    skey = p.sorted_keys_ptr(solver)
    assert skey
    place(1.0)          # move after the sort
    mom = {k: d[k].clone() for k in ("px", "py", "pz", "g_inv")}
    comps = a["vec_comps"]
    E = solver._planes_arr(["E" + c for c in comps])
    B = solver._planes_arr(["B" + c for c in comps])
    out = []
    for sidx, key in ((ptr(d["sort_indx"]), skey), (None, None)):
        for k, v in mom.items():
            d[k].copy_(v)
        _lib.call("chm_gather_push_f64", ptr(d["x"]), ptr(d["y"]), ptr(d["z"]), ptr(d["px"]),
                  ptr(d["py"]), ptr(d["pz"]), ptr(d["g_inv"]), n, sidx, ptr(d["Np_stay"]), key,
                  p.Args["FactorPush"], ctypes.byref(solver._grid()), E, B,
                  torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        out.append({k: d[k].cpu().numpy().copy() for k in mom})
    assert not np.array_equal(out[0]["px"], mom["px"].cpu().numpy())  # pushed
    for k in mom:
        assert np.array_equal(out[0][k], out[1][k]), k


def test_sort_dense_cells_and_edges(env):
    """Half of the particles in one cell, a quarter outside the domain, the
    rest spread: keys, offsets and the stable order still match numpy; an
    empty species sorts to empty arrays."""
    g = env.tiny_grid(Nx=64, Nr=24, M=1)
    solver, p, s, q = _pair_particles(env, g, 60000, 77, 1)
    d = p.DataDev
    n = 60000
    x = d["x"].clone()
    x[: n // 2] = g["Xmin"] + 10.5 * (g["Xmax"] - g["Xmin"]) / (g["Nx"] - 1)
    x[n // 2: 3 * n // 4] = g["Xmax"] + 1.0
    d["x"].copy_(x)
    q.data["x"][:] = x.cpu().numpy()
    p.sort_parts(solver)
    q.sort_parts(s)
    key = d["indx_in_cell"].cpu().numpy().view(np.uint32)
    assert np.array_equal(key, q.data["indx_in_cell"])
    assert np.array_equal(d["sort_indx"].cpu().numpy().view(np.uint32), np.argsort(key, kind="stable"))
    assert np.array_equal(d["cell_offset"].cpu().numpy().view(np.uint32), q.data["cell_offset"])
    assert p.Args["Np_stay"] == q.Args["Np_stay"]
    import chimeracl_amd as CA
    e = CA.Particles({"charge": -1, "dx": 0.02, "dr": 0.1}, solver.comm)
    e.add_particles(source=type("S", (), {"DataDev": {k + "_new": np.zeros(0) for k in
                                                      ("x", "y", "z", "px", "py", "pz", "w", "g_inv")}})())
    e.sort_parts(solver)
    assert e.Args["Np"] == 0


@pytest.mark.parametrize("gi", range(len(GRIDS)))
def test_deposit_dense_cells_bit_exact(env, gi):
    """Half of the particles in three adjacent cells (thousands per cell, so
    a group of cells spans many LDS chunks of the deposit), the rest spread:
    rho, J and the fused half push with its keys match the oracle bit for
    bit."""
    g = env.tiny_grid(**GRIDS[gi])
    n = 40000
    solver, p, s, q = _pair_particles(env, g, n, 90 + gi, g["M"])
    rng = np.random.default_rng(190 + gi)
    dx = (g["Xmax"] - g["Xmin"]) / (g["Nx"] - 1)
    dr = g["Rmax"] / (g["Nr"] - 1.5)
    k = n // 2
    r = rng.uniform(1.6 * dr, 2.4 * dr, k)
    th = rng.uniform(0, 2 * np.pi, k)
    dense = {"x": g["Xmin"] + dx * rng.uniform(10.0, 13.0, k), "y": r * np.sin(th),
             "z": r * np.cos(th)}
    for c, v in dense.items():
        a = p.DataDev[c].clone()
        a[:k] = torch.as_tensor(v, dtype=a.dtype, device=a.device)
        p.DataDev[c].copy_(a)
        q.data[c][:k] = v
    p.sort_parts(solver)
    q.sort_parts(s)
    assert np.array_equal(p.DataDev["cell_offset"].cpu().numpy().view(np.uint32), q.data["cell_offset"])
    assert np.diff(q.data["cell_offset"].astype(np.int64)).max() > 4000
    solver.depose_charge([p])
    s.depose_charge([q])
    solver.depose_currents(species=[p], push_half=True)
    s.depose_currents([q])
    q.push_coords("half")
    for name in env.real_names(g["M"], flds=("J",)):
        assert np.array_equal(env.field_host(solver, name), s.f[name]), name
    for c in ("x", "y", "z"):
        assert np.array_equal(p.host(c), q.data[c]), c
    p.sort_parts(solver)
    q.sort_parts(s)
    assert np.array_equal(p.DataDev["sort_indx"].cpu().numpy().view(np.uint32), q.data["sort_indx"])


def test_all_particles_outside(env):
    """Every particle outside the valid cells: Np_stay = 0, the deposits add
    nothing, the gather + push leaves the momenta unchanged (the reference
    visits no slot), as in the oracle."""
    g = env.tiny_grid(Nx=64, Nr=24, M=1)
    solver, p, s, q = _pair_particles(env, g, 5000, 13, 1)
    x = np.full(5000, g["Xmax"] + 1.0)
    p.DataDev["x"].copy_(torch.from_numpy(x))
    q.data["x"][:] = x
    p.sort_parts(solver)
    q.sort_parts(s)
    assert p.Args["Np_stay"] == 0 == q.Args["Np_stay"]
    assert np.array_equal(p.DataDev["cell_offset"].cpu().numpy().view(np.uint32), q.data["cell_offset"])
    solver.depose_charge([p])
    solver.depose_currents([p])
    for n in ["rho_m0", "rho_m1"] + [f"J{c}_m{m}" for c in "xyz" for m in (0, 1)]:
        assert torch.count_nonzero(solver.DataDev[n]) == 0, n
    u0 = {k: p.host(k) for k in ("px", "py", "pz", "g_inv")}
    solver.gather_and_push([p])
    for k, v in u0.items():
        assert np.array_equal(p.host(k), v), k


@pytest.mark.parametrize("case", ["many_tiles", "sparse"])
def test_sort_large_grid(env, case):
    """A grid of 131,968 bins (18-bit keys, two radix passes): 300,000
    particles span many chunks of sort tiles; a sparse species (a few
    clusters, long stretches of empty cells) takes the gap search of the
    offsets.  Keys, stable order, cell offsets and Np_stay match the oracle."""
    g = env.tiny_grid(Nx=1024, Nr=130, M=1)
    n = 300000 if case == "many_tiles" else 3000
    solver, p, s, q = _pair_particles(env, g, n, 91, 1)
    if case == "sparse":
        d = p.DataDev
        dx = (g["Xmax"] - g["Xmin"]) / (g["Nx"] - 1)
        rng = np.random.default_rng(3)
        centres = g["Xmin"] + dx * np.array([5.5, 300.2, 301.7, 900.4])
        x = centres[rng.integers(0, 4, n)] + rng.uniform(-0.4, 0.4, n) * dx
        d["x"].copy_(torch.from_numpy(x))
        q.data["x"][:] = x
    p.sort_parts(solver)
    q.sort_parts(s)
    key = p.DataDev["indx_in_cell"].cpu().numpy().view(np.uint32)
    assert np.array_equal(key, q.data["indx_in_cell"])
    assert np.array_equal(p.DataDev["sort_indx"].cpu().numpy().view(np.uint32),
                          np.argsort(key, kind="stable"))
    off = p.DataDev["cell_offset"].cpu().numpy().view(np.uint32)
    assert np.array_equal(off, q.data["cell_offset"])
    if case == "sparse":
        assert np.max(np.diff(np.unique(key))) > 500  # long empty stretches
    assert p.Args["Np_stay"] == q.Args["Np_stay"]


# (Nx, Nr) -> key bits of nbins = (Nx-1)(Nr-1)+1: 11 (two 8-bit radix
# passes), 17, 19 and 21 (two passes of up to 11 bits) and 23 (three passes;
# 3000 clusters overflow the list of long empty stretches of the offsets)
SORT_GRIDS = [(64, 30), (1024, 66), (2048, 256), (4096, 500), (4096, 1100)]
SORT_CASES = ["random", "near_sorted", "few_cells", "clusters", "outside", "tiny", "empty"]


def _sort_keys_case(case, nbins, rng):
    if case == "random":
        return rng.integers(0, nbins, 400000)
    if case == "near_sorted":  # cell order with local disorder, as between realignments
        k = np.sort(rng.integers(0, nbins - 1, 600000))
        return np.clip(k + rng.integers(-3, 4, k.size), 0, nbins - 1)
    if case == "few_cells":  # buckets far larger than one chunk of the bucket pass
        h = nbins // 2
        return rng.choice(np.array([0, 7, h, h + 1, nbins - 2]), 300000)
    if case == "clusters":  # 3000 clusters: long stretches of empty cells
        c = rng.integers(0, nbins - 1, 3000)
        return c[rng.integers(0, 3000, 50000)]
    if case == "outside":  # most particles in the out-of-domain bin
        k = rng.integers(0, nbins, 100000)
        k[rng.random(k.size) < 0.9] = nbins - 1
        return k
    if case == "tiny":
        return rng.integers(0, nbins, 37)
    return np.zeros(0, np.int64)


@pytest.mark.parametrize("gi", range(len(SORT_GRIDS)))
def test_sort_from_keys_widths(env, gi):
    """chm_sort_from_keys on synthetic keys for every key width the sort
    distinguishes and every shape of key distribution: sort_indx is the
    stable argsort, cell_offset[c] the first sorted slot with key >= c,
    Np_stay = cell_offset[nbins - 1] (numpy as the checker)."""
    import ctypes
    from chimeracl_amd import _lib
    Nx, Nr = SORT_GRIDS[gi]
    nbins = (Nx - 1) * (Nr - 1) + 1
    g = _lib.Grid(Nx, Nr, 1, 0, 0.0, 1.0, 0.0, 1.0)
    rng = np.random.default_rng(11 + gi)
    dev = torch.device("cuda")
    for case in SORT_CASES:
        key = _sort_keys_case(case, nbins, rng).astype(np.uint32)
        n = key.size
        kd = torch.from_numpy(key.view(np.int32).copy()).to(dev)
        off = torch.full((nbins + 1,), -1, dtype=torch.int32, device=dev)
        srt = torch.full((max(n, 1),), -1, dtype=torch.int32, device=dev)
        nps = torch.full((1,), -1, dtype=torch.int32, device=dev)
        wsb = _lib.lib.chm_sort_workspace_bytes(n, nbins)
        ws = torch.empty(int(wsb), dtype=torch.uint8, device=dev)
        _lib.call("chm_sort_from_keys", n, ctypes.byref(g), kd.data_ptr(), off.data_ptr(),
                  srt.data_ptr(), nps.data_ptr(), ws.data_ptr(), ws.numel(),
                  torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        ref = np.argsort(key, kind="stable")
        got = srt.cpu().numpy().view(np.uint32)[:n]
        assert np.array_equal(got, ref), (case, Nx, Nr)
        ref_off = np.searchsorted(key[ref], np.arange(nbins + 1), side="left")
        assert np.array_equal(off.cpu().numpy().view(np.uint32), ref_off), (case, Nx, Nr)
        assert int(nps.item()) == ref_off[nbins - 1], (case, Nx, Nr)


@pytest.mark.parametrize("n", [16_777_316, 33_563_401])
def test_sort_from_keys_large(env, n):
    """chm_sort_from_keys at 16.8M and 33.6M keys (thousands of 4096-key
    tiles, a partial last tile): 21-bit keys of a 4096 x 512 grid, both
    uniform and nearly sorted (cell order with local shuffles and a few far
    moves, as after a step of a cell-ordered plasma).  sort_indx is the
    stable argsort, cell_offset and Np_stay as the reference defines them."""
    import ctypes
    from chimeracl_amd import _lib
    Nx, Nr = 4096, 512
    nbins = (Nx - 1) * (Nr - 1) + 1
    g = _lib.Grid(Nx, Nr, 1, 0, 0.0, 1.0, 0.0, 1.0)
    rng = np.random.default_rng(n)
    dev = torch.device("cuda")
    near = np.sort(rng.integers(0, nbins - 1, n)).astype(np.int64)
    near += rng.integers(-1, 2, n)
    far = rng.random(n) < 1e-3
    near[far] = rng.integers(0, nbins, int(far.sum()))
    for case, key in (("uniform", rng.integers(0, nbins, n)), ("near-sorted", np.clip(near, 0, nbins - 1))):
        key = key.astype(np.uint32)
        print(f"large sort n={n} {case}", flush=True)
        kd = torch.from_numpy(key.view(np.int32)).to(dev)
        off = torch.full((nbins + 1,), -1, dtype=torch.int32, device=dev)
        srt = torch.full((n,), -1, dtype=torch.int32, device=dev)
        nps = torch.full((1,), -1, dtype=torch.int32, device=dev)
        ws = torch.empty(int(_lib.lib.chm_sort_workspace_bytes(n, nbins)), dtype=torch.uint8, device=dev)
        _lib.call("chm_sort_from_keys", n, ctypes.byref(g), kd.data_ptr(), off.data_ptr(),
                  srt.data_ptr(), nps.data_ptr(), ws.data_ptr(), ws.numel(),
                  torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        ref = np.argsort(key, kind="stable")
        assert np.array_equal(srt.cpu().numpy().view(np.uint32), ref), case
        ref_off = np.searchsorted(key[ref], np.arange(nbins + 1), side="left")
        assert np.array_equal(off.cpu().numpy().view(np.uint32), ref_off), case
        assert int(nps.item()) == ref_off[nbins - 1], case
        del kd, off, srt, ws


def test_dens_profile_bit_exact(env):
    g = env.tiny_grid()
    solver, p, s, q = _pair_particles(env, g, 5000, 3, 1)
    xx = [g["Xmin"] - 1, g["Xmin"] + 0.3, g["Xmin"] + 0.5, g["Xmax"] + 1]
    ff = [0.0, 0.5, 1.0, 1.2]
    p.dens_profile(xx, ff)
    q.dens_profile(xx, ff)
    assert np.array_equal(p.host("w"), q.data["w"])


@pytest.mark.parametrize("gi", range(len(GRIDS)))
def test_tables(env, gi):
    import chimeracl_amd as CA
    from chimeracl_amd.methods.generic_methods_hip import Communicator
    from oracle import oracle as O
    g = env.tiny_grid(**GRIDS[gi])
    solver = CA.Solver(dict(g), Communicator(precision="f64"))
    s = O.OracleSolver(dict(g))
    D = solver.DataDev
    M = g["M"]
    assert np.array_equal(D["kx"].cpu().numpy(), s.Args["kx"])
    for m in range(M + 2):
        assert np.allclose(D[f"kr_m{m}"].cpu().numpy(), s.sp["kr"][m], rtol=1e-14, atol=0)
    for m in range(M + 1):
        for mine, ref in (("DHT_inv_m", "DHT_inv"), ("DHT_m", "DHT"), ("dDHT_plus_m", "dDp"),
                          ("dDHT_minus_m", "dDm")):
            a, b = D[mine + str(m)].cpu().numpy(), s.dht[ref][m]
            assert env.rel_l2(a, b) < 1e-12, (mine, m, env.rel_l2(a, b))
        c1, c2, c3 = s.coef[m]
        for key, ref in (("w_m", s.sp["w"][m]), ("MxSlv_cos(wdt)_m", c1), ("MxSlv_sin(wdt)*w_m", c2),
                         ("MxSlv_1/w**2_m", c3), ("SmoothingFilter_m", s.sp["Smooth"][m])):
            a = D[key + str(m)].cpu().numpy()
            assert env.rel_l2(a, ref) < 1e-14, (key, m, env.rel_l2(a, ref))
    assert np.allclose(D["DampProfile"].cpu().numpy(), s.damp, rtol=1e-14, atol=1e-15)
    assert np.allclose(D["phase_minus"].cpu().numpy(), s.phase(0), rtol=0, atol=1e-14)


# the spectral operators also at 4096-point rows (the register-resident
# row FFT) with three modes
SPEC_GRIDS = GRIDS + [dict(Nx=4096, Nr=10, M=2)]


@pytest.mark.parametrize("gi", range(len(SPEC_GRIDS)))
def test_spectral_ops(env, gi):
    import chimeracl_amd as CA
    from chimeracl_amd.methods.generic_methods_hip import Communicator
    from oracle import oracle as O
    g = env.tiny_grid(**SPEC_GRIDS[gi])
    M = g["M"]
    solver = CA.Solver(dict(g), Communicator(precision="f64"))
    s = O.OracleSolver(dict(g))
    rng = np.random.default_rng(40 + gi)

    def setf(name, cplx):
        shp = s.f[name].shape
        v = rng.normal(0, 1, shp) + (1j * rng.normal(0, 1, shp) if cplx else 0)
        s.f[name][...] = v
        solver.DataDev[name].copy_(torch.from_numpy(np.ascontiguousarray(v)))

    # forward transforms (full and half)
    for c in "xyz":
        for m in range(M + 1):
            setf(f"J{c}_m{m}", m > 0)
    for m in range(M + 1):
        setf(f"rho_m{m}", m > 0)
    solver.fb_transform(scals=["rho"], vects=["J"], dir=0)
    s.fb_transform(scals=["rho"], vects=["J"], dir=0)
    for m in range(M + 1):
        for n in ["rho"] + ["J" + c for c in "xyz"]:
            nm = f"{n}_fb_m{m}"
            assert env.rel_l2(env.field_host(solver, nm), s.f[nm]) < 1e-12, nm
    # smoothing, grad with the dN0 copy
    solver.fields_smooth(["rho", "Jx", "Jy", "Jz"])
    s.fields_smooth(["rho", "Jx", "Jy", "Jz"])
    for c in "xyz":
        for m in range(M + 1):
            setf(f"dN1{c}_fb_m{m}", True)
    old = {f"dN1{c}_fb_m{m}": s.f[f"dN1{c}_fb_m{m}"].copy() for c in "xyz" for m in range(M + 1)}
    solver.field_grad("rho", "dN1", copy_to="dN0")
    s.field_grad("rho", "dN1")
    for c in "xyz":
        for m in range(M + 1):
            nm = f"dN1{c}_fb_m{m}"
            assert env.rel_l2(env.field_host(solver, nm), s.f[nm]) < 1e-12, nm
            assert np.array_equal(env.field_host(solver, f"dN0{c}_fb_m{m}"), old[nm])
            s.f[f"dN0{c}_fb_m{m}"][:] = old[nm]
    # PSATD + damping fused vs the oracle's sequence
    for f in ("E", "G"):
        for c in "xyz":
            for m in range(M + 1):
                setf(f"{f}{c}_fb_m{m}", True)
    solver.damp_edges(("E", "G"), advance_with=("J", "dN0", "dN1"))
    s.push_fields()
    s.damp_fields()
    for f in ("E", "G"):
        for c in "xyz":
            for m in range(M + 1):
                nm = f"{f}{c}_fb_m{m}"
                assert env.rel_l2(env.field_host(solver, nm), s.f[nm]) < 1e-12, nm
    # rot + Poisson, div, backward transforms
    solver.restore_B_fb()
    s.restore_B_fb()
    for c in "xyz":
        for m in range(M + 1):
            nm = f"B{c}_fb_m{m}"
            assert env.rel_l2(env.field_host(solver, nm), s.f[nm]) < 1e-12, nm
    for m in range(M + 1):
        solver.DataDev[f"divJ_fb_m{m}"] = torch.zeros_like(solver.DataDev[f"rho_fb_m{m}"])
    solver.field_div("J", "divJ")
    dref = s.field_div("J")
    for m in range(M + 1):
        assert env.rel_l2(env.field_host(solver, f"divJ_fb_m{m}"), dref[m]) < 1e-12
    solver.fb_transform(vects=["E", "B"], dir=1)
    s.fb_transform(vects=["E", "B"], dir=1)
    for n in env.real_names(M, flds=("E", "B"), scal=()):
        assert env.rel_l2(env.field_host(solver, n)[1:], s.f[n][1:]) < 1e-12, n


@pytest.mark.parametrize("Nx", [64, 4096])
def test_forward_transform_with_smoothing(env, Nx):
    """fb_transform_smooth (the step's forward transform, the filter applied
    in the row FFT's store: at 4096 points the FFT kernels compiled for the
    real- and complex-input forms with the table) against the oracle's
    fb_transform followed by fields_smooth, within 1e-12 relative L2."""
    import chimeracl_amd as CA
    from chimeracl_amd.methods.generic_methods_hip import Communicator
    from oracle import oracle as O
    g = env.tiny_grid(Nx=Nx, Nr=10, M=2)
    solver = CA.Solver(dict(g), Communicator(precision="f64"))
    s = O.OracleSolver(dict(g))
    rng = np.random.default_rng(11)
    names = ["rho", "Jx", "Jy", "Jz"]
    for n in names:
        for m in range(3):
            shp = s.f[f"{n}_m{m}"].shape
            v = rng.normal(0, 1, shp) + (1j * rng.normal(0, 1, shp) if m else 0)
            s.f[f"{n}_m{m}"][...] = v
            solver.DataDev[f"{n}_m{m}"].copy_(torch.from_numpy(np.ascontiguousarray(v)))
    solver.fb_transform_smooth(scals=["rho"], vects=["J"])
    s.fb_transform(scals=["rho"], vects=["J"], dir=0)
    s.fields_smooth(names)
    for n in names:
        for m in range(3):
            nm = f"{n}_fb_m{m}"
            assert env.rel_l2(env.field_host(solver, nm), s.f[nm]) < 1e-12, nm


def test_dht_product_vs_numpy(env):
    """The FP64 DHT product on the matrix cores (chm_gemm_f64, k_gemm)
    against numpy's FP64 matmul, for a large launch (1000 x 130 @ 130 x 8192,
    128 blocks along the columns) and for its first 64 rows alone: both within
    1e-13 of numpy, and the shared rows bit-identical (every output is the same
    k-ordered MFMA accumulation whichever block and launch computes it)."""
    import ctypes  # noqa: F401
    from chimeracl_amd import _lib
    dev = torch.device("cuda")
    rng = np.random.default_rng(5)
    M, K, N = 1000, 130, 8192
    A = rng.normal(size=(M, K))
    B = rng.normal(size=(K, N))
    ref = A @ B
    At = torch.from_numpy(A).to(dev)
    Bt = torch.from_numpy(B).to(dev)
    out = {}
    for rows in (M, 64):
        n = int(_lib.lib.chm_dht_table_elems(rows, K))
        pad = torch.zeros(n, dtype=torch.float64, device=dev)
        Ar = At[:rows].contiguous()
        _lib.call("chm_pad_table", Ar.data_ptr(), rows, K, pad.data_ptr(), None, None, None, None)
        C = torch.full((rows, N), np.nan, dtype=torch.float64, device=dev)
        _lib.call("chm_gemm_f64", pad.data_ptr(), Bt.data_ptr(), C.data_ptr(), rows, K, N, None)
        torch.cuda.synchronize()
        out[rows] = C.cpu().numpy()
        err = np.abs(out[rows] - ref[:rows]).max() / np.abs(ref[:rows]).max()
        assert err < 1e-13, (rows, err)
    assert np.array_equal(out[M][:64], out[64])


@pytest.mark.parametrize("Nx", [2048, 4096])
def test_mixed_transforms_vs_oracle(env, Nx):
    """The FP32 row FFTs (LDS Stockham at 2048, register-resident at 4096)
    with the split-BF16 DHT: forward and backward transforms of random
    fields against the FP64 oracle within 2e-5 relative L2."""
    import chimeracl_amd as CA
    from chimeracl_amd.methods.generic_methods_hip import Communicator
    from oracle import oracle as O
    g = env.tiny_grid(Nx=Nx, Nr=12, M=1)
    solver = CA.Solver(dict(g), Communicator(precision="mixed"))
    s = O.OracleSolver(dict(g))
    rng = np.random.default_rng(7)
    for m in range(2):
        shp = s.f[f"rho_m{m}"].shape
        v = rng.normal(0, 1, shp) + (1j * rng.normal(0, 1, shp) if m else 0)
        s.f[f"rho_m{m}"][...] = v
        solver.DataDev[f"rho_m{m}"].copy_(torch.from_numpy(np.ascontiguousarray(v)))
    solver.fb_transform(scals=["rho"], dir=0)
    s.fb_transform(scals=["rho"], dir=0)
    for m in range(2):
        nm = f"rho_fb_m{m}"
        assert env.rel_l2(env.field_host(solver, nm).astype(np.complex128), s.f[nm]) < 2e-5, nm
    for m in range(2):
        solver.DataDev[f"rho_m{m}"].zero_()
    solver.fb_transform(scals=["rho"], dir=1)
    s.fb_transform(scals=["rho"], dir=1)
    for m in range(2):
        nm = f"rho_m{m}"
        assert env.rel_l2(env.field_host(solver, nm)[1:], s.f[nm][1:]) < 2e-5, nm
