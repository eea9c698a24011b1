"""CPU FP64 oracle of the chimeraCL PIC step.

TEST INFRASTRUCTURE ONLY.  This module restates the reference algorithm
(hightower8083/chimeraCL v0.0.1) on the CPU so that the HIP path can be
checked against it.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it, and only as the checker or
as the timed CPU baseline; the product path (``chimeracl_amd``) never imports
or calls anything under ``oracle/``.

What it restates (reference file:line in each docstring):
  * geometry and derived arguments         grid.py:68-103, particles.py:53-100
  * spectral axes, DHT matrices           transformer.py:28-97
  * PSATD coefficient tables              solver.py:41-52
  * damping profile                       solver_methods_cl.py:32-41
  * transforms (full and half)            transformer_methods_cl.py:38-64, 290-455
  * get_m1, grad, div, rot, Poisson, smoothing
                                          transformer_methods_cl.py:66-288,
                                          transformer_generic.cl:3-24
  * PSATD advance, edge damping            solver_ms_pic.cl:5-143
  * particle push, cell keys, sort, deposit, axis post-processing, gather
    and Boris push                        particles_oracle.c (C, IEEE, no FMA),
                                          grid_generic.cl:4-86
  * the step order                        pic_loop.py:57-142
  * the laser initialiser                 laser.py:3-37

Numerics: plain IEEE FP64.  The particle loops are C built without FMA
contraction; the transforms use numpy.fft (the numpy inverse-FFT
normalisation 1/Nx, which the reference's round-trip test implies,
examples/test_transformer.py:35-43) and numpy matmul.  The reference builds
its kernels with -cl-fast-relaxed-math (generic_methods_cl.py:15); this
oracle is the exact-arithmetic refinement of that.

The in-cell particle order is the canonical stable one (ascending particle
index).  The reference's atomic scatter leaves it nondeterministic
(particles_generic.cl:186-200), so any fixed order is a valid refinement.

Parity pins (tests/test_oracle_*.py): every function here is checked against
something the reference or mathematics fixes -- closed forms, the two
known-answer scripts of the reference (round-trip identity and laser group
velocity), committed scipy/mpmath fixtures for the Bessel tables,
numpy.fft / brute-force loops on tiny inputs, sum rules and single-particle
closed forms for the deposit, and textbook Boris identities.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import time

import numpy as np
from scipy.special import jn_zeros, jv

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBNAME = os.path.join(_HERE, "liboracle_particles.so")
_lib = None


def build():
    """Compile the C part of the oracle (make -C oracle)."""
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIBNAME):
            build()
        L = ctypes.CDLL(_LIBNAME)
        P = ctypes.c_void_p
        i64, i32, f64 = ctypes.c_int64, ctypes.c_int32, ctypes.c_double
        L.orc_push_xyz.argtypes = [P, P, P, P, P, P, P, f64, i64]
        L.orc_cell_keys.argtypes = [P, P, P, i64, i32, f64, f64, i32, f64, f64, P]
        L.orc_stable_sort.argtypes = [P, i64, i64, P, P]
        L.orc_deposit_scalar.argtypes = [P, P, P, P, P, P, i32, i32, f64, f64,
                                         i32, f64, f64, i32, P, P]
        L.orc_deposit_vector.argtypes = [P, P, P, P, P, P, P, P, P, P, i32, i32,
                                         f64, f64, i32, f64, f64, i32, P, P]
        L.orc_gather_push.argtypes = [P, P, P, P, P, P, P, P, P, i64, i64, f64, i32,
                                      f64, f64, i32, f64, f64, i32, P, P, P, P]
        L.orc_profile_by_interpolant.argtypes = [P, P, i64, P, P, P, i32]
        L.orc_num_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


# --------------------------------------------------------------------------
# geometry
# --------------------------------------------------------------------------
def grid_args(cfg):
    """Grid._process_configs, grid.py:68-103."""
    a = dict(cfg)
    a.setdefault("M", 0)
    a["dx"] = (a["Xmax"] - a["Xmin"]) / (a["Nx"] - 1)
    a["dx_inv"] = 1.0 / a["dx"]
    a["dr"] = a["Rmax"] / (a["Nr"] - 1.5)
    a["dr_inv"] = 1.0 / a["dr"]
    if "dt" not in a:
        a["dt"] = a["dx"]
    a["dt_inv"] = 1.0 / a["dt"]
    a["Xgrid"] = a["Xmin"] + a["dx"] * np.arange(a["Nx"])
    a["Rmin"] = -0.5 * a["dr"]
    a["Rgrid"] = a["Rmin"] + a["dr"] * np.arange(a["Nr"])
    a["Rmax"] = a["Rgrid"].max()
    with np.errstate(divide="ignore"):
        a["dV_inv"] = (a["Rgrid"] > 0) / (2 * np.pi * a["dx"] * a["dr"] * a["Rgrid"])
    a["NxNr"] = a["Nr"] * a["Nx"]
    a["Nxm1Nrm1"] = (a["Nr"] - 1) * (a["Nx"] - 1)
    a["NxNrm1"] = (a["Nr"] - 1) * a["Nx"]
    return a


def particle_args(cfg):
    """Particles._process_configs, particles.py:53-100."""
    a = dict(cfg)
    a.setdefault("dt", 1.0)
    a.setdefault("dx", 1.0)
    a.setdefault("dr", 1.0)
    a.setdefault("charge", -1.0)
    a.setdefault("mass", 1.0)
    a.setdefault("dens", 1.0)
    a["dt_2"] = 0.5 * a["dt"]
    if "Nppc" in a:
        nppc = np.array(a["Nppc"], dtype=np.uint32)
        a["Nppc"] = nppc
        a["w0"] = 2 * np.pi * a["dx"] * a["dr"] * a["dens"] / np.prod(nppc)
        a["ddx"] = a["dx"] / nppc[0]
    else:
        a["ddx"] = 1.0
    a["FactorPush"] = 2 * np.pi * a["dt"] * a["charge"] / a["mass"]
    a["right_lim"] = 0.0
    return a


def spectral_axes(ga):
    """Transformer._make_spectral_axes, transformer.py:28-63."""
    Nx, Nr, M, dx, dr = ga["Nx"], ga["Nr"], ga["M"], ga["dx"], ga["dr"]
    kx = 2 * np.pi * np.fft.fftfreq(Nx, dx)
    R_period = ga["Rgrid"][-1] + 0.5 * dr
    kr = [jn_zeros(m, Nr - 1) / R_period for m in range(M + 2)]
    w = [np.sqrt(kx[None, :] ** 2 + kr[m][:, None] ** 2) for m in range(M + 1)]
    poiss = [1.0 / w[m] ** 2 for m in range(M + 1)]
    smooth = [(1 - np.sin(0.5 * np.pi * kx[None, :] / kx.max()) ** 2)
              * (1 - np.sin(0.5 * np.pi * kr[m][:, None] / kr[m].max()) ** 2)
              for m in range(M + 1)]
    return dict(kx=kx, kr=kr, w=w, Poiss=poiss, Smooth=smooth)


def dht_matrices(ga):
    """Transformer._make_DHT, transformer.py:65-97.

    jn_zeros(-1, n) is the zeros of J_1 (scipy takes |n|)."""
    Nr, M, dr = ga["Nr"], ga["M"], ga["dr"]
    Rg = ga["Rgrid"][1:, None]
    R_period = Rg[-1] + 0.5 * dr
    out = dict(DHT=[], DHT_inv=[], dDp=[], dDm=[])
    for m in range(M + 1):
        kr0 = jn_zeros(m, Nr - 1) / R_period
        krp = jn_zeros(m + 1, Nr - 1) / R_period
        krm = jn_zeros(abs(m - 1), Nr - 1) / R_period
        inv = jv(m, Rg * kr0)
        fwd = np.linalg.pinv(inv)
        out["DHT_inv"].append(inv)
        out["DHT"].append(fwd)
        out["dDp"].append(fwd.dot(0.5 * krp * jv(m, Rg * krp)))
        out["dDm"].append(fwd.dot(0.5 * krm * jv(m, Rg * krm)))
    return out


def ms_coefficients(w, dt):
    """Solver._make_ms_coefficients, solver.py:41-52."""
    return np.cos(w * dt), np.sin(w * dt) * w, 1.0 / w ** 2


def psatd_update(e0, g0, j, n0, n1, c1, c2, c3, dti):
    """One PSATD step of one spectral component, advance_e_g_m
    (solver_ms_pic.cl:114-134): the current j and the previous / current
    density gradients n0 / n1 are scaled by 2 pi (:114-122), then
      E' = c1 E + c2 c3 (G - j) + c3 (c1 n0 - n1 - (n0 - n1) dt_inv c2 c3)
      G' = -c2 E + c1 (G - j) + j + c3 (dt_inv (1 - c1) (n0 - n1) - c2 n0)
    with c1 = cos(w dt), c2 = w sin(w dt), c3 = 1/w^2.  Returns (E', G')."""
    pi2 = 2 * np.pi
    j0, n0, n1 = j * pi2, n0 * pi2, n1 * pi2
    e1 = (c1 * e0 + c2 * c3 * (g0 - j0)
          + c3 * (c1 * n0 - n1 - (n0 - n1) * dti * c2 * c3))
    g1 = (-c2 * e0 + c1 * (g0 - j0) + j0
          + c3 * (dti * (1.0 - c1) * (n0 - n1) - c2 * n0))
    return e1, g1


def damp_profile(N):
    """SolverMethodsCL._init_field_damping, solver_methods_cl.py:32-41."""
    z = np.arange(2 * N)
    z_shft = 3.0 * (z - N + 1) / (N + 1)
    return ((z < 4.0 * N / 3) * (z >= N) * np.sin(0.5 * np.pi * z_shft) ** 2
            + (z >= 4.0 * N / 3))


def get_m1(f):
    """get_m1, transformer_generic.cl:3-24: F_{-1}[ir,ix] = -conj(F_1[ir,-ix])."""
    Nx = f.shape[1]
    idx = (Nx - np.arange(Nx)) % Nx
    g = f[:, idx]
    return -g.real + 1j * g.imag


# --------------------------------------------------------------------------
# deterministic generator and sin/cos of the domain fill
# --------------------------------------------------------------------------
def rng_u01(seed, counters):
    """splitmix64 (Steele, Lea, Flood 2014) at seed + (counter+1)*golden, top
    53 bits as a double in [0, 1); the GPU's chm_fill_domain draws the same."""
    c = np.asarray(counters, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (c + np.uint64(1)) * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def det_sincos(a):
    """sin, cos from IEEE +, -, *, / only: Cody-Waite reduction by pi/2 with
    fdlibm's three-part constant, Taylor polynomials on |r| <= pi/4; the same
    expressions in the same order as the GPU kernel, hence the same bits."""
    a = np.asarray(a, dtype=np.float64)
    P1, P2, P2T = 1.57079632673412561417e+00, 6.07710050630396597660e-11, 2.02226624879595063154e-21
    k = np.floor(a * (2.0 / 3.14159265358979323846) + 0.5)
    r = ((a - k * P1) - k * P2) - k * P2T
    r2 = r * r
    s = np.full_like(r, 1.0 / 355687428096000.0)
    for cf in (-1.0 / 1307674368000.0, 1.0 / 6227020800.0, -1.0 / 39916800.0, 1.0 / 362880.0,
               -1.0 / 5040.0, 1.0 / 120.0, -1.0 / 6.0):
        s = cf + r2 * s
    s = r + r * (r2 * s)
    c = np.full_like(r, -1.0 / 6402373705728000.0)
    for cf in (1.0 / 20922789888000.0, -1.0 / 87178291200.0, 1.0 / 479001600.0, -1.0 / 3628800.0,
               1.0 / 40320.0, -1.0 / 720.0, 1.0 / 24.0, -0.5):
        c = cf + r2 * c
    c = 1.0 + r2 * c
    q = k.astype(np.int64) & 3
    sv = np.select([q == 0, q == 1, q == 2], [s, c, -s], -c)
    cv = np.select([q == 0, q == 1, q == 2], [c, -s, -c], s)
    return sv, cv


def det_log(x):
    """log from IEEE operations and the exact frexp split (the GPU's det_log):
    x = m 2^e, m in [sqrt(1/2), sqrt(2)), log m = 2 atanh((m-1)/(m+1)) to f^19,
    e ln2 with fdlibm's two-part ln2."""
    m, e = np.frexp(np.asarray(x, dtype=np.float64))
    low = m < 0.70710678118654752440
    m = np.where(low, m * 2.0, m)
    e = np.where(low, e - 1, e)
    f = (m - 1.0) / (m + 1.0)
    f2 = f * f
    s = np.full_like(f, 1.0 / 19.0)
    for cf in (1.0 / 17.0, 1.0 / 15.0, 1.0 / 13.0, 1.0 / 11.0, 1.0 / 9.0, 1.0 / 7.0, 1.0 / 5.0,
               1.0 / 3.0):
        s = cf + f2 * s
    lm = 2.0 * (f + f * (f2 * s))
    de = e.astype(np.float64)
    return de * 6.93147180369123816490e-01 + (lm + de * 1.90821492927058770002e-10)


def det_normal2(seed, c):
    """Box-Muller pair from counters (c, c+1), as the GPU draws it."""
    u1, u2 = rng_u01(seed, c), rng_u01(seed, c + 1)
    rad = np.sqrt(-2.0 * det_log(1.0 - u1))
    s, co = det_sincos((2.0 * 3.14159265358979323846) * u2)
    return rad * co, rad * s


# --------------------------------------------------------------------------
# particles
# --------------------------------------------------------------------------
class OracleParticles:
    """Particles (particles.py) with ParticleMethodsCL semantics."""

    def __init__(self, cfg):
        self.Args = particle_args(cfg)
        self.immobile = "Immobile" in self.Args
        self.data = {}
        self.Args["Np"] = 0
        self.Args["Np_stay"] = 0
        self.flag_sorted = False

    def add_particles(self, arrays):
        """add_new_particles(source=...), particles_methods_cl.py:40-61."""
        keys = ["x", "y", "z", "w"] if self.immobile else \
            ["x", "y", "z", "px", "py", "pz", "w", "g_inv"]
        for k in keys:
            new = np.ascontiguousarray(arrays[k], dtype=np.float64)
            old = self.data.get(k)
            self.data[k] = new.copy() if old is None else np.concatenate([old, new])
        if "pid" in self.data:  # appended particles take the next indices
            n0 = self.data["pid"].size
            self.data["pid"] = np.concatenate(
                [self.data["pid"], np.arange(n0, self.data["x"].size, dtype=np.uint32)])
        self.Args["Np"] = self.data["x"].size
        self.Args["Np_stay"] = self.Args["Np"]
        self.flag_sorted = False

    def dens_profile(self, x_prf, f_prf, xmin=None, xmax=None, coord="x", weight="w"):
        """dens_profile + profile_by_interpolant, particles_methods_cl.py:179-204,
        particles_generic.cl:6-30 (constant extension outside the knots)."""
        xx = np.ascontiguousarray(x_prf, dtype=np.float64)
        ff = np.ascontiguousarray(f_prf, dtype=np.float64)
        dxm1 = np.ascontiguousarray(1.0 / (xx[1:] - xx[:-1]))
        d = self.data
        lib().orc_profile_by_interpolant(_p(d[coord]), _p(d[weight]), d[coord].size,
                                         _p(xx), _p(ff), _p(dxm1), xx.size)

    def _keys(self):
        return ["x", "y", "z", "w"] if self.immobile else \
            ["x", "y", "z", "px", "py", "pz", "w", "g_inv"]

    def make_new_domain(self, parts_in, density_profiles=None):
        """make_new_domain, particles_methods_cl.py:66-147, with fill_grid,
        particles_generic.cl:33-84 (theta offsets from rng_u01, sin/cos from
        det_sincos, as the GPU fills it)."""
        a = self.Args
        xmin, xmax, rmin, rmax = (parts_in[k] for k in ("Xmin", "Xmax", "Rmin", "Rmax"))
        Nx_loc = int(np.ceil((xmax - xmin) / a["dx"]) + 1)
        Nr_loc = int(np.round((rmax - rmin) / a["dr"]) + 1)
        xg = xmin + a["dx"] * np.arange(Nx_loc)
        rg = rmin + a["dr"] * np.arange(Nr_loc)
        a["right_lim"] = float(xg[-1])
        ncells = (Nx_loc - 1) * (Nr_loc - 1)
        npx, npr, npt = (int(v) for v in a["Nppc"])
        counter = int(a.get("rng_counter", 0))
        cell = np.arange(ncells, dtype=np.int64)
        ir, ix = cell // (Nx_loc - 1), cell % (Nx_loc - 1)
        thmin = (2.0 * 3.14159265358979323846) * rng_u01(int(a.get("seed", 0)), counter + cell)
        a["rng_counter"] = counter + ncells
        Lx = xg[ix + 1] - xg[ix]
        Lr = rg[ir + 1] - rg[ir]
        dx, dr, dth = 1. / float(npx), 1. / float(npr), 2 * 3.14159265358979323846 / float(npt)
        it = np.arange(npt)[None, :, None, None]
        irr = np.arange(npr)[None, None, :, None]
        ixx = np.arange(npx)[None, None, None, :]
        th = thmin[:, None, None, None] + it * dth
        sin_th, cos_th = det_sincos(th)
        rp = rg[ir][:, None, None, None] + (0.5 + irr) * dr * Lr[:, None, None, None]
        shp = (ncells, npt, npr, npx)
        x = np.broadcast_to(xg[ix][:, None, None, None] + (0.5 + ixx) * dx * Lx[:, None, None, None], shp)
        new = {"x": x.ravel().copy(), "y": np.broadcast_to(rp * sin_th, shp).ravel().copy(),
               "z": np.broadcast_to(rp * cos_th, shp).ravel().copy(),
               "w": np.broadcast_to(rp * a["w0"], shp).ravel().copy()}
        if not self.immobile:
            # particles_methods_cl.py:114-147: p_c, or p_c + dp * randn when a
            # spread dp is given; normals of particle p from counters
            # counter + ncells + 4p .. + 2 (det_normal2 pairs), as the GPU draws them
            pc = [float(parts_in.get(k + "_c", 0.0)) for k in ("px", "py", "pz")]
            dp = [float(parts_in.get("d" + k, 0.0)) for k in ("px", "py", "pz")]
            n = new["x"].size
            if any(v != 0.0 for v in dp):
                c = np.uint64(counter + ncells) + np.uint64(4) * np.arange(n, dtype=np.uint64)
                n0, n1 = det_normal2(int(a.get("seed", 0)), c)
                n2, _ = det_normal2(int(a.get("seed", 0)), c + np.uint64(2))
                ux, uy, uz = pc[0] + dp[0] * n0, pc[1] + dp[1] * n1, pc[2] + dp[2] * n2
                new.update(px=ux, py=uy, pz=uz, g_inv=1. / np.sqrt(1 + ux * ux + uy * uy + uz * uz))
                a["rng_counter"] += 4 * n
            else:
                for k, v in zip(("px", "py", "pz"), pc):
                    new[k] = np.full(n, v)
                g = 1.0
                if any(v != 0.0 for v in pc):
                    g = 1. / np.sqrt(1 + pc[0] * pc[0] + pc[1] * pc[1] + pc[2] * pc[2])
                new["g_inv"] = np.full(n, g)
        for k, v in new.items():
            self.data[k + "_new"] = v
        if density_profiles is not None:
            for profile in density_profiles:
                if profile["coord"] == "x":
                    self.dens_profile(profile["points"], profile["values"], coord="x_new",
                                      weight="w_new")

    def make_new_beam(self, parts_in):
        """make_new_beam, particles_methods_cl.py:149-177 (normals of particle p
        from counters counter0 + 6p .. 6p+5, as the GPU fills them)."""
        a = self.Args
        Np = int(parts_in["Np"])
        counter = int(a.get("rng_counter", 0))
        c = counter + 6 * np.arange(Np, dtype=np.uint64)
        n0, n1 = det_normal2(int(a.get("seed", 0)), c)
        n2, n3 = det_normal2(int(a.get("seed", 0)), c + np.uint64(2))
        n4, n5 = det_normal2(int(a.get("seed", 0)), c + np.uint64(4))
        a["rng_counter"] = counter + 6 * Np
        g = lambda k, d=0.0: float(parts_in.get(k, d))
        new = {"x": g("x_c") + g("Lx") * n0, "y": g("y_c") + g("Ly") * n1, "z": g("z_c") + g("Lz") * n2,
               "px": g("px_c") + g("dpx") * n3, "py": g("py_c") + g("dpy") * n4,
               "pz": g("pz_c") + g("dpz") * n5}
        new["w"] = np.full(Np, float(parts_in["FullCharge"]) / Np)
        ux, uy, uz = new["px"], new["py"], new["pz"]
        new["g_inv"] = 1. / np.sqrt(1 + ux * ux + uy * uy + uz * uz)
        for k, v in new.items():
            self.data[k + "_new"] = v

    def add_new_particles(self, source=None):
        """add_new_particles, particles_methods_cl.py:35-61."""
        src = (self if source is None else source).data
        self.add_particles({k: src[k + "_new"] for k in self._keys()})

    def free_added(self):
        for k in list(self.data):
            if k.endswith("_new"):
                del self.data[k]

    def align_parts(self):
        """align_parts -> align_and_damp, particles.py:42-51,
        particles_methods_cl.py:263-293: keep the Np_stay particles of the
        valid cells, in sorted order."""
        if self.Args["Np"] == 0:
            return
        idx = self.data["sort_indx"][: self.Args["Np_stay"]].astype(np.int64)
        for k in self._keys():
            self.data[k] = np.ascontiguousarray(self.data[k][idx])
        self.data.pop("pid", None)  # the reference renumbers the particles here
        self.Args["Np"] = self.Args["Np_stay"] = idx.size
        self.flag_sorted = False

    def reorder_parts(self):
        """The storage reorder of the product (not in the reference): every
        component of all Np particles permuted by sort_indx with the
        data_align gather (particles_generic.cl:156-186), none dropped;
        sort_indx becomes the identity, cell offsets and Np_stay stay.  pid
        keeps each particle's index in the unreordered numbering (for the
        gather's K5 skip)."""
        if self.Args["Np"] == 0 or not self.flag_sorted:
            return
        idx = self.data["sort_indx"].astype(np.int64)
        if "pid" not in self.data:
            self.data["pid"] = np.arange(self.Args["Np"], dtype=np.uint32)
        for k in self._keys() + ["pid"]:
            self.data[k] = np.ascontiguousarray(self.data[k][idx])
        self.data["sort_indx"] = np.arange(self.Args["Np"], dtype=np.uint32)

    def push_coords(self, mode="half"):
        """push_coords, particles_methods_cl.py:206-223."""
        if self.Args["Np"] == 0 or self.immobile:
            return
        dt = self.Args["dt_2"] if mode == "half" else self.Args["dt"]
        d = self.data
        lib().orc_push_xyz(_p(d["x"]), _p(d["y"]), _p(d["z"]), _p(d["px"]),
                           _p(d["py"]), _p(d["pz"]), _p(d["g_inv"]), dt,
                           self.Args["Np"])
        self.flag_sorted = False

    def sort_parts(self, grid):
        """sort_parts, particles.py:23-29 -> index_sort, particles_methods_cl.py:225-261."""
        if self.Args["Np"] == 0:
            self.flag_sorted = True
        if not self.flag_sorted:
            self.index_sort(grid)
            self.flag_sorted = True

    def index_sort(self, grid):
        g, d, Np = grid.Args, self.data, self.Args["Np"]
        key = np.empty(Np, np.uint32)
        lib().orc_cell_keys(_p(d["x"]), _p(d["y"]), _p(d["z"]), Np, g["Nx"],
                            g["Xmin"], g["dx_inv"], g["Nr"], g["Rmin"],
                            g["dr_inv"], _p(key))
        nbins = g["Nxm1Nrm1"] + 1
        off = np.empty(nbins + 1, np.uint32)
        srt = np.empty(Np, np.uint32)
        lib().orc_stable_sort(_p(key), Np, nbins, _p(off), _p(srt))
        d["indx_in_cell"] = key
        d["cell_offset"] = off
        d["sort_indx"] = srt
        self.Args["Np_stay"] = int(off[nbins - 1])


# --------------------------------------------------------------------------
# grid + transformer + solver
# --------------------------------------------------------------------------
class OracleSolver:
    """Solver(Grid, Transformer, SolverMethodsCL), solver.py:9-68."""

    comps = ("x", "y", "z")

    def __init__(self, cfg):
        self.Args = a = grid_args(cfg)
        Nx, Nr, M = a["Nx"], a["Nr"], a["M"]
        self.M = M
        sp = spectral_axes(a)
        a.update(kx=sp["kx"])
        self.sp = sp
        self.dht = dht_matrices(a)
        self.coef = [ms_coefficients(sp["w"][m], a["dt"]) for m in range(M + 1)]
        self.damp = damp_profile(int(a["DampCells"])) if "DampCells" in a else None
        self.f = {}
        for fld in ["E", "B", "J", "G"]:
            for c in self.comps:
                self._alloc_real(fld + c)
        self._alloc_real("rho")
        for fld in ["E", "B", "G", "J", "dN0", "dN1"]:
            for c in self.comps:
                for m in range(M + 1):
                    self.f[f"{fld}{c}_fb_m{m}"] = np.zeros((Nr - 1, Nx), complex)
        for m in range(M + 1):
            self.f[f"rho_fb_m{m}"] = np.zeros((Nr - 1, Nx), complex)

    def _alloc_real(self, name):
        Nx, Nr = self.Args["Nx"], self.Args["Nr"]
        self.f[name + "_m0"] = np.zeros((Nr, Nx))
        for m in range(1, self.M + 1):
            self.f[name + f"_m{m}"] = np.zeros((Nr, Nx), complex)

    # ---- phases, transformer_generic.cl:28-55
    def phase(self, dir):
        arg = self.Args["Xmin"] * self.Args["kx"]
        return np.cos(arg) + (1j if dir == 1 else -1j) * np.sin(arg)

    # ---- transforms, transformer_methods_cl.py:38-64, 290-455
    def transform_field(self, name, dir, mode="full"):
        phs = self.phase(dir)[None, :]
        f, M = self.f, self.M
        for m in range(M + 1):
            rs, fb = f"{name}_m{m}", f"{name}_fb_m{m}"
            if dir == 0:
                src = f[rs][1:]
                if mode == "full":
                    src = self.dht["DHT"][m] @ src
                f[fb][:] = np.fft.fft(src.astype(complex), axis=1) * phs
            else:
                buf = np.fft.ifft(f[fb] * phs, axis=1)
                if m == 0:
                    # contiguous, so that the product below runs in BLAS (a
                    # strided .real view takes numpy's slow generic loop)
                    buf = np.ascontiguousarray(buf.real)
                if mode == "full":
                    buf = self.dht["DHT_inv"][m] @ buf
                f[rs][1:] = buf

    def fb_transform(self, scals=(), vects=(), dir=0, mode="full"):
        """Transformer.fb_transform, transformer.py:17-26."""
        for s in scals:
            self.transform_field(s, dir, mode)
        for v in vects:
            for c in self.comps:
                self.transform_field(v + c, dir, mode)

    # ---- spectral operators, transformer_methods_cl.py:66-263
    def fields_smooth(self, flds):
        for m in range(self.M + 1):
            for fl in flds:
                self.f[f"{fl}_fb_m{m}"] *= self.sp["Smooth"][m]

    def field_poiss_vec(self, fld):
        for m in range(self.M + 1):
            for c in self.comps:
                self.f[f"{fld}{c}_fb_m{m}"] *= self.sp["Poiss"][m]

    def field_grad(self, scl, vec):
        f, M, kx = self.f, self.M, self.Args["kx"][None, :]
        mm1 = get_m1(f[f"{scl}_fb_m1"]) if M > 0 else None
        for m in range(M + 1):
            x = kx * (1j * f[f"{scl}_fb_m{m}"])
            y = np.zeros_like(x)
            z = np.zeros_like(x)
            if m > 0 or M > 0:
                src = f[f"{scl}_fb_m{m-1}"] if m > 0 else mm1
                b = self.dht["dDm"][m] @ src
                y += -1.0 * b
                z += -1j * b
                if m < M:
                    b = self.dht["dDp"][m] @ f[f"{scl}_fb_m{m+1}"]
                    y += b
                    z += -1j * b
            f[f"{vec}x_fb_m{m}"][:] = x
            f[f"{vec}y_fb_m{m}"][:] = y
            f[f"{vec}z_fb_m{m}"][:] = z

    def field_div(self, vec):
        """field_div, transformer_methods_cl.py:135-183 (returns the list of modes)."""
        f, M, kx = self.f, self.M, self.Args["kx"][None, :]
        if M > 0:
            my = get_m1(f[f"{vec}y_fb_m1"])
            mz = get_m1(f[f"{vec}z_fb_m1"])
        out = []
        for m in range(M + 1):
            o = kx * (1j * f[f"{vec}x_fb_m{m}"])
            if m > 0 or M > 0:
                fy = f[f"{vec}y_fb_m{m-1}"] if m > 0 else my
                fz = f[f"{vec}z_fb_m{m-1}"] if m > 0 else mz
                o = o + self.dht["dDm"][m] @ (-1j * fz - 1.0 * fy)
                if m < M:
                    fy, fz = f[f"{vec}y_fb_m{m+1}"], f[f"{vec}z_fb_m{m+1}"]
                    o = o + self.dht["dDp"][m] @ (-1j * fz + 1.0 * fy)
            out.append(o)
        return out

    def field_rot(self, fin, fout):
        f, M, kx = self.f, self.M, self.Args["kx"][None, :]
        if M > 0:
            mm1 = {c: get_m1(f[f"{fin}{c}_fb_m1"]) for c in self.comps}
        for m in range(M + 1):
            bx = np.zeros_like(f[f"{fin}x_fb_m{m}"])
            by = kx * (-1j * f[f"{fin}z_fb_m{m}"])
            bz = kx * (1j * f[f"{fin}y_fb_m{m}"])
            if m > 0 or M > 0:
                if m > 0:
                    gx, gy, gz = (f[f"{fin}{c}_fb_m{m-1}"] for c in self.comps)
                else:
                    gx, gy, gz = mm1["x"], mm1["y"], mm1["z"]
                bx = bx + self.dht["dDm"][m] @ (-1 * gz + 1j * gy)
                b = self.dht["dDm"][m] @ gx
                by = by + (-1j) * b
                bz = bz + 1 * b
                if m < M:
                    gx, gy, gz = (f[f"{fin}{c}_fb_m{m+1}"] for c in self.comps)
                    bx = bx + self.dht["dDp"][m] @ (1 * gz + 1j * gy)
                    b = self.dht["dDp"][m] @ gx
                    by = by + (-1j) * b
                    bz = bz + (-1) * b
            f[f"{fout}x_fb_m{m}"][:] = bx
            f[f"{fout}y_fb_m{m}"][:] = by
            f[f"{fout}z_fb_m{m}"][:] = bz

    # ---- solver, solver.py:29-39 and solver_ms_pic.cl
    def push_fields(self):
        """advance_e_g_m, solver_ms_pic.cl:57-143, per mode."""
        f, dti = self.f, self.Args["dt_inv"]
        for m in range(self.M + 1):
            c1, c2, c3 = self.coef[m]
            for c in self.comps:
                e0 = f[f"E{c}_fb_m{m}"]
                g0 = f[f"G{c}_fb_m{m}"]
                e1, g1 = psatd_update(e0, g0, f[f"J{c}_fb_m{m}"], f[f"dN0{c}_fb_m{m}"],
                                      f[f"dN1{c}_fb_m{m}"], c1, c2, c3, dti)
                e0[:] = e1
                g0[:] = g1

    def profile_edges(self, flds):
        """profile_edges_d/c, solver_ms_pic.cl:5-55 (whole [Nr,Nx] plane)."""
        Nx, prof = self.Args["Nx"], self.damp
        Nf = prof.size
        fac = np.ones(Nx)
        ix = np.arange(Nx)
        lo = ix < Nf
        fac[lo] *= prof[ix[lo]]
        hi = ix > Nx - Nf
        fac[hi] *= prof[Nx - ix[hi]]
        for fl in flds:
            for c in self.comps:
                for m in range(self.M + 1):
                    self.f[f"{fl}{c}_m{m}"] *= fac[None, :]

    def damp_fields(self):
        """Solver.damp_fields, solver.py:32-35."""
        self.fb_transform(vects=["E", "G"], dir=1, mode="half")
        self.profile_edges(["E", "G"])
        self.fb_transform(vects=["E", "G"], dir=0, mode="half")

    def restore_B_fb(self):
        self.field_rot("G", "B")
        self.field_poiss_vec("B")

    # ---- grid, grid.py:23-66 and grid_methods_cl.py
    def _postproc(self, name):
        """treat_axis then divide_by_dv, grid_generic.cl:4-60."""
        dvi = self.Args["dV_inv"][:, None]
        for m in range(self.M + 1):
            a = self.f[f"{name}_m{m}"]
            a[1] -= a[0]
            a *= dvi

    def _modes(self, name):
        M = self.M
        m0 = self.f[name + "_m0"]
        mc = (np.stack([self.f[name + f"_m{m}"] for m in range(1, M + 1)])
              if M > 0 else np.zeros((1, 1, 1), complex))
        return m0, np.ascontiguousarray(mc)

    def depose_charge(self, species):
        a, M = self.Args, self.M
        m0 = np.zeros((a["Nr"], a["Nx"]))
        mc = np.zeros((max(M, 1), a["Nr"], a["Nx"]), complex)
        for parts in species:
            if parts.Args["Np"] <= 0:
                continue
            d = parts.data
            lib().orc_deposit_scalar(_p(d["sort_indx"]), _p(d["cell_offset"]),
                                     _p(d["x"]), _p(d["y"]), _p(d["z"]), _p(d["w"]),
                                     int(parts.Args["charge"]), a["Nx"], a["Xmin"],
                                     a["dx_inv"], a["Nr"], a["Rmin"], a["dr_inv"], M,
                                     _p(m0), _p(mc))
        self.f["rho_m0"][:] = m0
        for m in range(1, M + 1):
            self.f[f"rho_m{m}"][:] = mc[m - 1]
        self._postproc("rho")

    def depose_currents(self, species):
        a, M = self.Args, self.M
        m0 = np.zeros((3, a["Nr"], a["Nx"]))
        mc = np.zeros((3, max(M, 1), a["Nr"], a["Nx"]), complex)
        for parts in species:
            if parts.immobile or parts.Args["Np"] <= 0:
                continue
            d = parts.data
            lib().orc_deposit_vector(_p(d["sort_indx"]), _p(d["cell_offset"]),
                                     _p(d["x"]), _p(d["y"]), _p(d["z"]),
                                     _p(d["px"]), _p(d["py"]), _p(d["pz"]),
                                     _p(d["g_inv"]), _p(d["w"]),
                                     int(parts.Args["charge"]), a["Nx"], a["Xmin"],
                                     a["dx_inv"], a["Nr"], a["Rmin"], a["dr_inv"], M,
                                     _p(m0), _p(mc))
        for k, c in enumerate(self.comps):
            self.f[f"J{c}_m0"][:] = m0[k]
            for m in range(1, M + 1):
                self.f[f"J{c}_m{m}"][:] = mc[k, m - 1]
            self._postproc("J" + c)

    def preproc_project_vec(self, fld):
        """warp_axis_m0_d / warp_axis_m1plus_c, grid_generic.cl:63-86."""
        for c in self.comps:
            a = self.f[f"{fld}{c}_m0"]
            a[0] = a[1]
            for m in range(1, self.M + 1):
                a = self.f[f"{fld}{c}_m{m}"]
                a[0] = -a[1]

    def gather_and_push(self, species):
        for fld in ("E", "B"):
            self.preproc_project_vec(fld)
        a, M = self.Args, self.M
        E0 = np.ascontiguousarray(np.stack([self.f[f"E{c}_m0"] for c in self.comps]))
        B0 = np.ascontiguousarray(np.stack([self.f[f"B{c}_m0"] for c in self.comps]))
        if M > 0:
            Ec = np.ascontiguousarray(np.stack([np.stack([self.f[f"E{c}_m{m}"] for m in range(1, M + 1)]) for c in self.comps]))
            Bc = np.ascontiguousarray(np.stack([np.stack([self.f[f"B{c}_m{m}"] for m in range(1, M + 1)]) for c in self.comps]))
        else:
            Ec = Bc = np.zeros(2)
        for parts in species:
            if parts.immobile or parts.Args["Np"] == 0:
                continue
            d = parts.data
            lib().orc_gather_push(_p(d["x"]), _p(d["y"]), _p(d["z"]), _p(d["px"]),
                                  _p(d["py"]), _p(d["pz"]), _p(d["g_inv"]),
                                  _p(d["sort_indx"]), _p(d["pid"]) if "pid" in d else None,
                                  parts.Args["Np"],
                                  parts.Args["Np_stay"], parts.Args["FactorPush"],
                                  a["Nx"], a["Xmin"], a["dx_inv"], a["Nr"], a["Rmin"],
                                  a["dr_inv"], M, _p(E0), _p(B0), _p(Ec), _p(Bc))


def add_gausian_pulse(solver, laser):
    """laser.py:3-37."""
    k0 = 2 * np.pi * laser["k0"]
    a0, Lx, R, x0, x_foc = (laser[k] for k in ("a0", "Lx", "R", "x0", "x_foc"))
    X_focus = x0 - x_foc
    a = solver.Args
    Xg, Rg = a["Xgrid"], a["Rgrid"]
    kx, w = a["kx"][None, :], solver.sp["w"][0]
    solver.f["Ez_m0"][1:] = (a0 * np.sin(k0 * (Xg[None, :] - x0))
                             * np.exp(-(Xg[None, :] - x0) ** 2 / Lx ** 2 - Rg[1:, None] ** 2 / R ** 2)
                             * (abs(Rg[1:, None]) < 3.5 * R) * (abs(Xg[None, :] - x0) < 3.5 * Lx))
    solver.fb_transform(scals=["Ez"], dir=0)
    EE = solver.f["Ez_fb_m0"].copy()
    DT = -1j * w * np.sign(kx + (kx == 0))
    GG = DT * EE
    EE_tmp = np.cos(w * X_focus) * EE + np.sin(w * X_focus) / w * GG
    GG = -w * np.sin(w * X_focus) * EE + np.cos(w * X_focus) * GG
    EE = EE_tmp
    EE = EE * np.exp(1j * kx * X_focus)
    GG = GG * np.exp(1j * kx * X_focus)
    solver.f["Ez_fb_m0"][:] = EE
    solver.f["Gz_fb_m0"][:] = GG
    solver.restore_B_fb()
    solver.fb_transform(vects=["B", "E"], dir=1)


class OracleFrame:
    """Frame, frame.py:4-64 (moving window and plasma injection)."""

    def __init__(self, cfg):
        self.Args = dict(cfg)
        self.Args.setdefault("Steps", 1.)
        self.Args.setdefault("Velocity", 0.)
        self.Args.setdefault("dt", 1)
        self.Args.setdefault("DensityProfiles", None)

    def shift_grids(self, grids=(), steps=None):
        steps = self.Args["Steps"] if steps is None else steps
        x_shift = steps * self.Args["dt"] * self.Args["Velocity"]
        for g in grids:
            for k in ("Xmax", "Xmin", "Xgrid"):
                g.Args[k] = g.Args[k] + x_shift

    def inject_plasma(self, species, grid, steps=None):
        steps = self.Args["Steps"] if steps is None else steps
        x_shift = steps * self.Args["dt"] * self.Args["Velocity"]
        for sp in species:
            if sp.Args["Np"] == 0:
                sp.Args["right_lim"] = grid.Args["Xmax"] - x_shift
            dom = {"Xmin": sp.Args["right_lim"]}
            dom["Xmax"] = dom["Xmin"] + x_shift
            dom["Rmin"] = grid.Args["Rmin"] * (grid.Args["Rmin"] > 0)
            dom["Rmax"] = grid.Args["Rmax"]
            sp.make_new_domain(dom, density_profiles=self.Args["DensityProfiles"])
            sp.add_new_particles(sp.Args.get("InjectorSource"))
        for sp in species:
            sp.free_added()
            sp.sort_parts(grid)
            sp.align_parts()
            num_ppc = int(np.prod(sp.Args["Nppc"]) + 1)
            sp.Args["right_lim"] = float(sp.data["x"][-num_ppc:].max()) + 0.5 * sp.Args["ddx"]


class OracleLoop:
    """PIC_loop.step, pic_loop.py:57-142 (frames, no diagnostics)."""

    def __init__(self, solver, species, frames=(), align_every=0):
        self.solver, self.species, self.frames, self.it = solver, species, list(frames), 0
        self.align_every = align_every
        # wall time per phase, keyed by the reference's loop_steps names
        # (pic_loop.py:5-8)
        self.Timer = dict.fromkeys(["frame", "push-x", "sort", "depose", "transform", "smooth",
                                    "data_copy", "grad", "push-eb", "damp-eb", "restore_B",
                                    "gather + push-p"], 0.0)

    def _t(self, name, t0):
        t1 = time.perf_counter()
        self.Timer[name] += t1 - t0
        return t1

    def step(self):
        s = self.solver
        t = time.perf_counter()
        if self.align_every and self.it > 0 and self.it % self.align_every == 0:
            for p in self.species:
                if not p.immobile:
                    p.reorder_parts()
        for fr in self.frames:
            if np.mod(self.it, fr.Args["Steps"]) == 0:
                fr.shift_grids(grids=[s])
                fr.inject_plasma(self.species, s)
        t = self._t("frame", t)
        for p in self.species:
            p.push_coords("half")
            t = self._t("push-x", t)
            p.sort_parts(s)
            t = self._t("sort", t)
        s.depose_currents(self.species)
        t = self._t("depose", t)
        for p in self.species:
            p.push_coords("half")
            t = self._t("push-x", t)
            p.sort_parts(s)
            t = self._t("sort", t)
        s.depose_charge(self.species)
        t = self._t("depose", t)
        s.fb_transform(scals=["rho"], vects=["J"], dir=0)
        t = self._t("transform", t)
        s.fields_smooth(["rho", "Jx", "Jy", "Jz"])
        t = self._t("smooth", t)
        for m in range(s.M + 1):
            for c in s.comps:
                s.f[f"dN0{c}_fb_m{m}"][:] = s.f[f"dN1{c}_fb_m{m}"]
        t = self._t("data_copy", t)
        s.field_grad("rho", "dN1")
        t = self._t("grad", t)
        s.push_fields()
        t = self._t("push-eb", t)
        s.damp_fields()
        t = self._t("damp-eb", t)
        s.restore_B_fb()
        t = self._t("restore_B", t)
        s.fb_transform(vects=["E", "B"], dir=1)
        t = self._t("transform", t)
        s.gather_and_push(self.species)
        self._t("gather + push-p", t)
        self.it += 1
        return self.it
