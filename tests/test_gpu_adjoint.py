"""Adjoint kernels of the differentiable form, FP64.

Linear maps (transforms, grad, rot, PSATD + damping, axis post-processing,
ghost-row warp) are pinned by the dot-product identity <A x, y> = <x, A^T y>
with random inputs, which any wrong sign, index, table or conjugation
breaks.  The nonlinear particle maps (deposits, gather + Boris push, density
profile) are pinned by central finite differences of the forward kernels in
FP64 along random directions, with particles kept away from cell
boundaries so the CIC cells do not change under the perturbation.  The
whole differentiable run is checked against central finite differences of
the objective."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _require_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def ip(a, b):
    """Real inner product sum Re(conj(a) b) over lists of tensors."""
    tot = 0.0
    for x, y in zip(a, b):
        x, y = x.double() if not x.is_complex() else x.to(torch.complex128), \
               y.double() if not y.is_complex() else y.to(torch.complex128)
        if x.is_complex():
            tot += float((x.real * y.real + x.imag * y.imag).sum())
        else:
            tot += float((x * y).sum())
    return tot


def rand_like(t, gen):
    if t.is_complex():
        re = torch.randn(t.shape, generator=gen, dtype=torch.float64)
        im = torch.randn(t.shape, generator=gen, dtype=torch.float64)
        return torch.complex(re, im).to(t.device, t.dtype)
    return torch.randn(t.shape, generator=gen, dtype=torch.float64).to(t.device, t.dtype)


def rel(a, b):
    return abs(a - b) / max(abs(a), abs(b), 1e-300)


@pytest.fixture(scope="module", params=[0, 1, 2])
def solv(request):
    _require_gpu()
    import chimeracl_amd as CA
    from chimeracl_amd.adjoint import AdjointSpectral
    from chimeracl_amd.methods.generic_methods_hip import Communicator
    from tests import simpair as S
    g = S.tiny_grid(Nx=64, Nr=24, M=request.param, damp=6)
    s = CA.Solver(dict(g), Communicator(precision="f64"))
    return s, AdjointSpectral(s)


def _fb(s, name):
    return [s.DataDev[f"{name}_fb_m{m}"] for m in range(s.Args["M"] + 1)]


def _re(s, name):
    return [s.DataDev[f"{name}_m{m}"] for m in range(s.Args["M"] + 1)]


def _vec_fb(s, v):
    return [_fb(s, v + c) for c in "xyz"]


def test_grad_adjoint_dot(solv):
    from chimeracl_amd import _lib
    from chimeracl_amd.adjoint import _fb_arr
    s, adj = solv
    gen = torch.Generator().manual_seed(1)
    rho = _fb(s, "rho")
    for t in rho:
        t.copy_(rand_like(t, gen))
    s.field_grad("rho", "dN1")
    out = _vec_fb(s, "dN1")
    lam = [[rand_like(t, gen) for t in comp] for comp in out]
    lhs = ip([t for c in out for t in c], [t for c in lam for t in c])
    rb = [[torch.zeros_like(t) for t in rho]]
    _lib.call("chm_field_grad_adj_f64", adj.sp, _fb_arr(lam), _fb_arr(rb), 0, s._tws.data_ptr(),
              s._tws.numel(), 0)
    rhs = ip(rho, rb[0])
    assert rel(lhs, rhs) < 1e-12, (lhs, rhs)


@pytest.mark.parametrize("poisson", [0, 1])
def test_rot_adjoint_dot(solv, poisson):
    from chimeracl_amd import _lib
    from chimeracl_amd.adjoint import _fb_arr
    s, adj = solv
    gen = torch.Generator().manual_seed(2 + poisson)
    G = _vec_fb(s, "G")
    for comp in G:
        for t in comp:
            t.copy_(rand_like(t, gen))
    s.field_rot("G", "B", poisson=bool(poisson))
    B = _vec_fb(s, "B")
    lam = [[rand_like(t, gen) for t in comp] for comp in B]
    lhs = ip([t for c in B for t in c], [t for c in lam for t in c])
    Gb = [[torch.zeros_like(t) for t in comp] for comp in G]
    _lib.call("chm_field_rot_adj_f64", adj.sp, _fb_arr(lam), _fb_arr(Gb), poisson, 0,
              s._tws.data_ptr(), s._tws.numel(), 0)
    rhs = ip([t for c in G for t in c], [t for c in Gb for t in c])
    assert rel(lhs, rhs) < 1e-12, (lhs, rhs)


def test_rot_adjoint_accumulates(solv):
    from chimeracl_amd import _lib
    from chimeracl_amd.adjoint import _fb_arr
    s, adj = solv
    gen = torch.Generator().manual_seed(7)
    lam = [[rand_like(t, gen) for t in comp] for comp in _vec_fb(s, "B")]
    base = [[rand_like(t, gen) for t in comp] for comp in _vec_fb(s, "G")]
    out0 = [[torch.zeros_like(t) for t in comp] for comp in base]
    out1 = [[t.clone() for t in comp] for comp in base]
    lam2 = [[t.clone() for t in comp] for comp in lam]
    ws = (s._tws.data_ptr(), s._tws.numel())
    _lib.call("chm_field_rot_adj_f64", adj.sp, _fb_arr(lam), _fb_arr(out0), 1, 0, *ws, 0)
    _lib.call("chm_field_rot_adj_f64", adj.sp, _fb_arr(lam2), _fb_arr(out1), 1, 1, *ws, 0)
    for c0, c1, cb in zip(out0, out1, base):
        for a, b, c in zip(c0, c1, cb):
            assert torch.allclose(a + c, b, rtol=0, atol=1e-12 * float(b.abs().max()))


def test_backward_transform_adjoint_dot(solv):
    from chimeracl_amd import _lib
    from chimeracl_amd.adjoint import _fb_arr, _planes_arr
    s, adj = solv
    gen = torch.Generator().manual_seed(3)
    F = _fb(s, "Ex")
    for t in F:
        t.copy_(rand_like(t, gen))
    s.transform_field("Ex", dir=1)
    out = [t[1:] for t in _re(s, "Ex")]
    lam = [rand_like(t, gen) for t in _re(s, "Ex")]
    lhs = ip(out, [t[1:] for t in lam])
    Fb = [[torch.zeros_like(t) for t in F]]
    _lib.call("chm_transform_f64", adj.sp, _planes_arr([lam]), _fb_arr(Fb), 1, 0, 0, 0,
              s._tws.data_ptr(), s._tws.numel(), 0)
    rhs = ip(F, Fb[0])
    assert rel(lhs, rhs) < 1e-12, (lhs, rhs)


def test_forward_transform_smooth_adjoint_dot(solv):
    from chimeracl_amd import _lib
    from chimeracl_amd.adjoint import _fb_arr, _planes_arr
    s, adj = solv
    gen = torch.Generator().manual_seed(4)
    P = _re(s, "rho")
    for t in P:
        t.copy_(rand_like(t, gen))
    s.fb_transform_smooth(scals=["rho"])
    out = _fb(s, "rho")
    lam = [rand_like(t, gen) for t in out]
    lhs = ip(out, lam)
    lamf = _fb_arr([lam])
    _lib.call("chm_mult_table_f64", s._sp, lamf, 1, 0, 0)
    Pb = [torch.zeros_like(t) for t in P]
    _lib.call("chm_transform_f64", adj.sp, _planes_arr([Pb]), lamf, 1, 1, 0, 0,
              s._tws.data_ptr(), s._tws.numel(), 0)
    rhs = ip([t[1:] for t in P], [t[1:] for t in Pb])
    assert rel(lhs, rhs) < 1e-12, (lhs, rhs)


@pytest.mark.parametrize("with_add", [False, True])
def test_psatd_damp_adjoint_dot(solv, with_add):
    from chimeracl_amd import _lib
    from chimeracl_amd.adjoint import _fb_arr
    s, adj = solv
    gen = torch.Generator().manual_seed(5)
    ins = {v: _vec_fb(s, v) for v in ("E", "G", "J", "dN0", "dN1")}
    for v in ins.values():
        for comp in v:
            for t in comp:
                t.copy_(rand_like(t, gen))
    x0 = {k: [[t.clone() for t in c] for c in v] for k, v in ins.items()}
    s.damp_edges(("E", "G"), advance_with=("J", "dN0", "dN1"))
    lamE = [[rand_like(t, gen) for t in c] for c in ins["E"]]
    lamG = [[rand_like(t, gen) for t in c] for c in ins["G"]]
    flat = lambda L: [t for c in L for t in c]
    lhs = ip(flat(ins["E"]), flat(lamE)) + ip(flat(ins["G"]), flat(lamG))
    Eadd = None
    if with_add:
        half = [[0.5 * t for t in c] for c in lamE]
        lamE, Eadd = half, [[t.clone() for t in c] for c in half]
    z = lambda: [[torch.zeros_like(t) for t in c] for c in lamE]
    Jb, carry, N1b = z(), z(), z()
    _lib.call("chm_damp_fields_adj_f64", s._sp, _fb_arr(lamE), _fb_arr(lamG),
              _fb_arr(Eadd) if Eadd else None, _fb_arr(Jb), _fb_arr(carry), _fb_arr(N1b), 0)
    rhs = (ip(flat(x0["E"]), flat(lamE)) + ip(flat(x0["G"]), flat(lamG)) +
           ip(flat(x0["J"]), flat(Jb)) + ip(flat(x0["dN0"]), flat(carry)) +
           ip(flat(x0["dN1"]), flat(N1b)))
    assert rel(lhs, rhs) < 1e-12, (lhs, rhs)


def test_postproc_and_warp_adjoint_dot(solv):
    from chimeracl_amd import _lib
    from chimeracl_amd.adjoint import _planes_arr
    s, adj = solv
    a = s.Args
    gen = torch.Generator().manual_seed(6)
    for fwd, name in (("chm_postproc_deposit_f64", "chm_postproc_deposit_adj_f64"),
                      ("chm_warp_axis_f64", "chm_warp_axis_adj_f64")):
        P = [rand_like(t, gen) for t in _re(s, "rho")]
        P0 = [t.clone() for t in P]
        lam = [rand_like(t, gen) for t in P]
        dvi = s.DataDev["dV_inv"]
        if "postproc" in fwd:
            _lib.call(fwd, _planes_arr([P]), 1, dvi.data_ptr(), a["Nx"], a["Nr"], a["M"], 0)
            lhs = ip(P, lam)
            _lib.call(name, _planes_arr([lam]), 1, dvi.data_ptr(), a["Nx"], a["Nr"], a["M"], 0)
        else:
            _lib.call(fwd, _planes_arr([P]), 1, a["Nx"], a["Nr"], a["M"], 0)
            lhs = ip(P, lam)
            _lib.call(name, _planes_arr([lam]), 1, a["Nx"], a["Nr"], a["M"], 0)
        rhs = ip(P0, lam)
        assert rel(lhs, rhs) < 1e-13, (fwd, lhs, rhs)


# ---------------------------------------------------------------- particles
@pytest.fixture(scope="module", params=[1, 2])
def psim(request):
    """Solver + particles strictly inside valid cells, away from cell
    boundaries and the axis."""
    _require_gpu()
    import chimeracl_amd as CA
    from chimeracl_amd import config as C
    from chimeracl_amd.methods.generic_methods_hip import Communicator
    from tests import simpair as S
    M = request.param
    g = S.tiny_grid(Nx=48, Nr=20, M=M, damp=4)
    comm = Communicator(precision="f64")
    s = CA.Solver(dict(g), comm)
    a = s.Args
    rng = np.random.default_rng(10 + M)
    n = 3000
    ix = rng.integers(1, a["Nx"] - 2, n)
    ir = rng.integers(1, a["Nr"] - 2, n)
    x = a["Xmin"] + a["dx"] * (ix + rng.uniform(0.1, 0.9, n))
    r = a["Rmin"] + a["dr"] * (ir + rng.uniform(0.1, 0.9, n))
    th = rng.uniform(0, 2 * np.pi, n)
    arr = {"x": x, "y": r * np.sin(th), "z": r * np.cos(th), "px": rng.normal(0, 0.5, n),
           "py": rng.normal(0, 0.5, n), "pz": rng.normal(0, 0.5, n),
           "w": rng.uniform(0.5, 1.5, n), "g_inv": rng.uniform(0.4, 1.0, n)}
    eo, _ = C.species_config("C1", a)
    e = CA.Particles(dict(eo), comm)
    e.add_particles(source=C.ParticleSource(arr))
    return s, e, rng


def _fd(fun, arrays, dirs, h):
    """Central difference of fun() along dirs (dict name -> tensor)."""
    saved = {k: arrays[k].clone() for k in dirs}
    for k, d in dirs.items():
        arrays[k].add_(h * d)
    lp = fun()
    for k, d in dirs.items():
        arrays[k].copy_(saved[k] - h * d)
    lm = fun()
    for k in dirs:
        arrays[k].copy_(saved[k])
    return (lp - lm) / (2 * h)


def _dirs(e, rng, keys):
    d = e.DataDev
    return {k: torch.from_numpy(rng.normal(0, 1, d[k].numel())).to(d[k].device) for k in keys}


def test_deposit_adjoints_fd(psim):
    from chimeracl_amd import _lib
    from chimeracl_amd.adjoint import _planes_arr
    s, e, rng = psim
    a, d = s.Args, e.DataDev
    gen = torch.Generator().manual_seed(11)
    names_J = [f"J{c}" for c in "xyz"]
    lamJ = [[rand_like(t, gen) for t in _re(s, n)] for n in names_J]
    lamR = [rand_like(t, gen) for t in _re(s, "rho")]

    def L_J():
        e.index_sort(s)
        s.depose_currents(species=[e])
        return ip([t for n in names_J for t in _re(s, n)], [t for c in lamJ for t in c])

    def L_rho():
        e.index_sort(s)
        s.depose_charge(species=[e])
        return ip(_re(s, "rho"), lamR)

    grid = s._grid()
    Np = e.Args["Np"]
    dvi = s.DataDev["dV_inv"].data_ptr()
    zeros = lambda: torch.zeros_like(d["x"])
    for kind in ("J", "rho"):
        keys = ["x", "y", "z", "w"] + (["px", "py", "pz", "g_inv"] if kind == "J" else [])
        dirs = _dirs(e, rng, keys)
        h = 1e-6
        fd = _fd(L_J if kind == "J" else L_rho, d, dirs, h)
        bars = {k: zeros() for k in ["x", "y", "z", "w", "px", "py", "pz", "g_inv"]}
        lam = [[t.clone() for t in c] for c in lamJ] if kind == "J" else [[t.clone() for t in lamR]]
        P = _planes_arr(lam)
        _lib.call("chm_postproc_deposit_adj_f64", P, len(lam), dvi, a["Nx"], a["Nr"], a["M"], 0)
        p = lambda k: d[k].data_ptr()
        b = lambda k: bars[k].data_ptr()
        if kind == "J":
            _lib.call("chm_deposit_vector_adj_f64", p("x"), p("y"), p("z"), p("px"), p("py"), p("pz"),
                      p("g_inv"), p("w"), -1, Np, grid, P, b("x"), b("y"), b("z"), b("px"), b("py"),
                      b("pz"), b("g_inv"), b("w"), 0)
        else:
            _lib.call("chm_deposit_scalar_adj_f64", p("x"), p("y"), p("z"), p("w"), -1, Np, grid, P,
                      b("x"), b("y"), b("z"), b("w"), 0)
        adj = sum(float((bars[k] * dirs[k]).sum()) for k in keys)
        assert rel(fd, adj) < 1e-6, (kind, fd, adj)


def test_gather_push_adjoint_fd(psim):
    from chimeracl_amd import _lib
    from chimeracl_amd.adjoint import _planes_arr
    s, e, rng = psim
    a, d = s.Args, e.DataDev
    gen = torch.Generator().manual_seed(12)
    E = [_re(s, f"E{c}") for c in "xyz"]
    B = [_re(s, f"B{c}") for c in "xyz"]
    for comp in E + B:
        for t in comp:
            t.copy_(0.3 * rand_like(t, gen))
    e.index_sort(s)
    Np = e.Args["Np"]
    lam = {k: torch.from_numpy(rng.normal(0, 1, Np)).to(d["x"].device) for k in ("px", "py", "pz", "g_inv")}
    keep = {k: d[k].clone() for k in ("px", "py", "pz", "g_inv")}
    flds = {f"F{i}": t for i, t in enumerate([t for c in E + B for t in c])}

    def L():
        for k in keep:
            d[k].copy_(keep_cur[k])
        s._gather_and_push(e, ["E", "B"])
        out = sum(float((d[k] * lam[k]).sum()) for k in lam)
        return out

    keep_cur = {k: v.clone() for k, v in keep.items()}
    dirs = _dirs(e, rng, ["x", "y", "z"])
    du = {k: torch.from_numpy(rng.normal(0, 1, Np)).to(d["x"].device) for k in ("px", "py", "pz")}
    dF = {k: 0.3 * rand_like(t, gen) for k, t in flds.items()}
    h = 1e-7
    # perturb x, y, z, the momenta before the push and the field planes together
    arrays = dict(d)
    arrays.update(flds)
    for k in du:
        arrays["u_" + k] = keep_cur[k]
    alld = dict(dirs)
    alld.update(dF)
    alld.update({"u_" + k: v for k, v in du.items()})
    fd = _fd(L, arrays, alld, h)
    # adjoint at the unperturbed point
    for k in keep:
        d[k].copy_(keep_cur[k])
    bars = {k: lam[k].clone() for k in ("px", "py", "pz", "g_inv")}
    for k in ("x", "y", "z"):
        bars[k] = torch.zeros_like(d["x"])
    fb = [torch.zeros_like(d["x"]) for _ in range(6)]
    fb_arr = (ctypes.c_void_p * 6)(*[t.data_ptr() for t in fb])
    p = lambda k: d[k].data_ptr()
    b = lambda k: bars[k].data_ptr()
    _lib.call("chm_gather_push_adj_f64", p("x"), p("y"), p("z"), p("px"), p("py"), p("pz"), Np,
              d["Np_stay"].data_ptr(), e.Args["FactorPush"], s._grid(), _planes_arr(E),
              _planes_arr(B), b("px"), b("py"), b("pz"), b("g_inv"), b("x"), b("y"), b("z"),
              ctypes.cast(fb_arr, ctypes.c_void_p), 0)
    Eb = [[torch.zeros_like(t) for t in c] for c in E]
    Bb = [[torch.zeros_like(t) for t in c] for c in B]
    ws = s._dep_ws()
    for k0, out in ((0, Eb), (3, Bb)):
        _lib.call("chm_gather_adj_scatter_f64", d["sort_indx"].data_ptr(), d["cell_offset"].data_ptr(),
                  p("x"), p("y"), p("z"), fb[k0].data_ptr(), fb[k0 + 1].data_ptr(),
                  fb[k0 + 2].data_ptr(), s._grid(), _planes_arr(out), ws.data_ptr(), ws.numel(), 0)
    adj = sum(float((bars[k] * dirs[k]).sum()) for k in ("x", "y", "z"))
    adj += sum(float((bars[k] * du[k]).sum()) for k in du)
    adj += ip([t for c in Eb + Bb for t in c], [dF[f"F{i}"] for i in range(len(dF))])
    assert rel(fd, adj) < 1e-6, (fd, adj)


def test_push_x_and_profile_adjoints_fd(psim):
    from chimeracl_amd import _lib
    s, e, rng = psim
    d = e.DataDev
    Np = e.Args["Np"]
    dev = d["x"].device
    lam = {k: torch.from_numpy(rng.normal(0, 1, Np)).to(dev) for k in ("x", "y", "z")}
    keep = {k: d[k].clone() for k in ("x", "y", "z")}

    def L():
        for k in keep:
            d[k].copy_(keep[k])
        e.push_coords(mode="half")
        out = sum(float((d[k] * lam[k]).sum()) for k in lam)
        for k in keep:
            d[k].copy_(keep[k])
        return out

    dirs = _dirs(e, rng, ["px", "py", "pz", "g_inv"])
    # x' is bilinear in (u, g_inv): the central difference is exact up to
    # rounding, so a large step keeps the rounding small
    fd = _fd(L, d, dirs, 1e-2)
    bars = {k: torch.zeros_like(d["x"]) for k in ("px", "py", "pz", "g_inv")}
    p = lambda k: d[k].data_ptr()
    _lib.call("chm_push_x_adj_f64", lam["x"].data_ptr(), lam["y"].data_ptr(), lam["z"].data_ptr(),
              p("px"), p("py"), p("pz"), p("g_inv"), *(bars[k].data_ptr() for k in ("px", "py", "pz", "g_inv")),
              e.Args["dt_2"], Np, 0)
    adj = sum(float((bars[k] * dirs[k]).sum()) for k in dirs)
    assert rel(fd, adj) < 1e-9, (fd, adj)
    # density profile: L = sum c_p w_p(knots), linear in the knot values
    a = s.Args
    xx = np.linspace(a["Xmin"] - 0.05, a["Xmax"] + 0.01, 7)
    ff = 1 + 0.2 * rng.uniform(-1, 1, 7)
    c = torch.from_numpy(rng.normal(0, 1, Np)).to(dev)
    w0 = d["w"].clone()
    x0 = d["x"].clone()

    def Lk(vals):
        d["w"].copy_(w0)
        e.dens_profile(xx, vals)
        return float((d["w"] * c).sum())

    g_fd = np.array([(Lk(ff + 1e-3 * np.eye(7)[j]) - Lk(ff - 1e-3 * np.eye(7)[j])) / 2e-3 for j in range(7)])
    d["w"].copy_(w0)
    kg = torch.zeros(7, dtype=torch.float64, device=dev)
    ws = torch.empty(_lib.lib.chm_profile_adj_workspace_bytes(7), dtype=torch.uint8, device=dev)
    xt, dxm1 = torch.from_numpy(xx).to(dev), torch.from_numpy(1 / np.diff(xx)).to(dev)
    _lib.call("chm_dens_profile_adj_f64", x0.data_ptr(), w0.data_ptr(), c.data_ptr(), Np,
              xt.data_ptr(), dxm1.data_ptr(), 7, kg.data_ptr(), ws.data_ptr(), ws.numel(), 0)
    np.testing.assert_allclose(kg.cpu().numpy(), g_fd, rtol=1e-9, atol=1e-9 * np.abs(g_fd).max())


# ---------------------------------------------------------------- whole run
def _tiny_diff(precision="f64", nsteps=6, ckpt=4, gather_skip="index"):
    import chimeracl_amd.adjoint as A
    from chimeracl_amd import config as C
    from chimeracl_amd.methods.generic_methods_hip import Communicator
    from tests import simpair as S
    g = S.tiny_grid(Nx=64, Nr=24, M=1, damp=4)
    las = S.tiny_laser(g, a0=2.0)
    comm = Communicator(precision=precision)
    import chimeracl_amd as CA
    s0 = CA.Solver(dict(g), comm)
    eo, io = C.species_config("C1", s0.Args, ppc=(2, 2, 2), dens=0.01)
    eo["Nppc"] = io["Nppc"] = (2, 2, 2)
    eo["dens"] = io["dens"] = 0.01
    del s0
    plasma = C.load_plasma(g, (2, 2, 2), 0.01, seed=0)
    knots_x = np.linspace(g["Xmin"], g["Xmax"], 5)
    sim = A.DiffPIC(g, las, (eo, io), plasma, knots_x, comm, ckpt=ckpt, gather_skip=gather_skip)
    knots = 1.0 + 0.2 * np.random.default_rng(1).uniform(-1, 1, 5)
    return sim, las, knots


@pytest.mark.parametrize("gather_skip", ["index", "cell"])
def test_run_gradient_vs_finite_differences(gather_skip):
    """The whole run's adjoint against central finite differences, with the
    reference's gather skip (particles with index >= Np_stay are not pushed)
    and with the differentiable form's alternative that pushes every particle
    of a valid cell."""
    _require_gpu()
    import chimeracl_amd.adjoint as A
    sim, las, knots = _tiny_diff(gather_skip=gather_skip)
    nsteps = 6
    rg = sim.solver.Args["Rgrid"]

    def objective(outs):
        px, py, pz = outs[:3]
        return A.objective_gamma2(px, py, pz, threshold=1.5, tau=0.25) + \
            1e3 * A.objective_wake_energy(list(outs[3:]), rg, ncols=16)

    a0 = torch.tensor(las["a0"], dtype=torch.float64, requires_grad=True)
    R = torch.tensor(las["R"], dtype=torch.float64, requires_grad=True)
    k = torch.tensor(knots, dtype=torch.float64, requires_grad=True)
    L = objective(A.run(sim, a0, R, k, nsteps))
    L.backward()
    grads = np.concatenate([[a0.grad.item(), R.grad.item()], k.grad.numpy()])

    def f(p):
        with torch.no_grad():
            return float(objective(sim.forward(p[0], p[1], p[2:], nsteps)))

    p0 = np.concatenate([[las["a0"], las["R"]], knots])
    fd = np.zeros_like(p0)
    for j in range(len(p0)):
        h = 1e-5 * abs(p0[j])
        e = np.zeros_like(p0)
        e[j] = h
        fd[j] = (f(p0 + e) - f(p0 - e)) / (2 * h)
    scale = np.abs(fd).max()
    err = np.abs(grads - fd) / scale
    print("adjoint", grads, "\nfd     ", fd, "\nerr", err)
    assert float(L.detach()) > 0
    assert err.max() < 1e-5, err


def test_mixed_adjoint_matches_fp64_adjoint():
    """The FP32 adjoint (split-BF16 DHT products) of the tiny run agrees with
    the FP64 adjoint to the precision of the mixed forward."""
    _require_gpu()
    import chimeracl_amd.adjoint as A
    out = {}
    for prec in ("f64", "mixed"):
        sim, las, knots = _tiny_diff(prec)
        rg = sim.solver.Args["Rgrid"]
        a0 = torch.tensor(las["a0"], dtype=torch.float64, requires_grad=True)
        R = torch.tensor(las["R"], dtype=torch.float64, requires_grad=True)
        k = torch.tensor(knots, dtype=torch.float64, requires_grad=True)
        outs = A.run(sim, a0, R, k, 6)
        L = A.objective_gamma2(*outs[:3], threshold=1.5, tau=0.25) + \
            1e3 * A.objective_wake_energy(list(outs[3:]), rg, ncols=16)
        L.backward()
        out[prec] = np.concatenate([[a0.grad.item(), R.grad.item()], k.grad.numpy()])
    g64, g32 = out["f64"], out["mixed"]
    err = np.abs(g32 - g64) / np.abs(g64).max()
    assert err.max() < 1e-3, (g64, g32)
