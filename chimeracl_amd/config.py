"""Workload definitions: the five laser-wakefield configurations of the step,
mapped onto the reference's configuration dictionaries, plus the seeded
plasma loader that feeds ``Particles.add_particles(source=...)``.

Mapping choices (the reference leaves these open):
  * Xmax = 0, Xmin = -(Nx-1)*dz, so dx = dz (grid.py:74);
    Rmax = dr*(Nr-1.5), so the radial step is dr (grid.py:78).
  * dt = dz; DampCells = 50 as in examples/lpa_script_small.py:55
    (damping is mandatory in PIC_loop.step, pic_loop.py:122-124).
  * Static window (no Frame).
  * Laser (laser.py:3-37): k0 = 1, Lx = FWHM/sqrt(2 ln 2) (field
    exp(-x^2/Lx^2) has intensity FWHM = FWHM), centre x0 = Xmax - 3,
    focus x_foc = x0.
  * Plasma: electrons (charge -1, mass 1) and immobile ions (charge +1) at
    identical positions.  Loading follows the reference's domain geometry
    (make_new_domain, particles_methods_cl.py:66-147): radial loading cells
    [ir*dr, (ir+1)*dr) for ir = 0..Nr-3 and longitudinal cells ix = 1..Nx-3,
    Nppc = (x, r, theta) sub-positions per cell, weights w = w0*r*f(x).
    Instead of the reference's regular sub-positions plus a Threefry theta
    offset, each particle is jittered uniformly inside its sub-cell with
    numpy.random.default_rng(seed) (PCG64): one uniform triple (x, r, theta)
    per particle, particles ordered cell (ir-major, ix-minor) -> theta-sub ->
    r-sub -> x-sub (the nesting of particles_generic.cl:68-82).
"""
from __future__ import annotations

import math

import numpy as np

DZ = 0.02
DR = 0.1


def _lx(fwhm):
    return fwhm / math.sqrt(2.0 * math.log(2.0))


CONFIGS = {
    "C1": dict(Nx=512, Nr=128, M=1, ppc=(2, 2, 4), dens=0.002, steps=200,
               a0=2.0, w0=4.0, fwhm=6.0, profile=None,
               desc="plain FP64 small laser-wakefield step"),
    "C2": dict(Nx=2048, Nr=256, M=1, ppc=(2, 2, 8), dens=0.002, steps=500,
               a0=2.0, w0=4.0, fwhm=6.0, profile=None,
               desc="medium laser-wakefield step, FP64 and mixed"),
    "C3": dict(Nx=4096, Nr=512, M=2, ppc=(2, 2, 8), dens=0.003, steps=1000,
               a0=3.0, w0=5.0, fwhm=8.0, profile="ramp512",
               desc="large mixed-precision step, M=2, density ramp"),
    "C4": dict(Nx=1024, Nr=128, M=1, ppc=(2, 2, 4), dens=0.002, steps=100,
               a0=2.0, w0=4.0, fwhm=6.0, profile="knots32", ckpt=10,
               desc="differentiable small step"),
    "C5": dict(Nx=4096, Nr=256, M=1, ppc=(2, 2, 8), dens=0.002, steps=400,
               a0=2.0, w0=4.0, fwhm=6.0, profile="knots64", ckpt=20,
               desc="differentiable large step, mixed precision"),
}


def grid_config(name, **over):
    """Reference Solver/Grid configuration dictionary."""
    c = dict(CONFIGS[name], **over)
    Nx, Nr = c["Nx"], c["Nr"]
    return {"Xmin": -(Nx - 1) * DZ, "Xmax": 0.0, "Nx": Nx,
            "Rmin": 0.0, "Rmax": DR * (Nr - 1.5), "Nr": Nr,
            "M": c["M"], "dt": DZ, "DampCells": c.get("DampCells", 50)}


def laser_config(name, **over):
    c = dict(CONFIGS[name], **over)
    x0 = 0.0 - 3.0
    return {"k0": 1.0, "a0": c["a0"], "x0": x0, "Lx": _lx(c["fwhm"]),
            "R": c["w0"], "x_foc": x0}


def species_config(name, solver_args, **over):
    """(electrons, ions) reference Particles configuration dictionaries."""
    c = dict(CONFIGS[name], **over)
    eons = {"Nppc": tuple(c["ppc"]), "dx": solver_args["dx"],
            "dr": solver_args["dr"], "dt": solver_args["dt"],
            "dens": c["dens"], "charge": -1}
    ions = dict(eons)
    ions["charge"] = 1
    ions["Immobile"] = True
    return eons, ions


def density_profile(name, grid_cfg, seed=1, **over):
    """Knots (points, values) of the longitudinal density multiplier or None."""
    c = dict(CONFIGS[name], **over)
    prof = c.get("profile")
    if prof is None:
        return None
    xmin, xmax = grid_cfg["Xmin"], grid_cfg["Xmax"]
    if prof == "ramp512":
        return ([xmin - 1.0, xmin, xmin + 512 * DZ, xmax + 1.0], [0.0, 0.0, 1.0, 1.0])
    if prof.startswith("knots"):
        n = int(prof[5:])
        u = np.random.default_rng(seed).uniform(-1.0, 1.0, n)
        return (list(np.linspace(xmin, xmax, n)), list(1.0 + 0.2 * u))
    raise ValueError(prof)


def n_particles(Nx, Nr, ppc):
    return (Nx - 3) * (Nr - 2) * int(np.prod(ppc))


def load_plasma(grid_cfg, ppc, dens, seed=0, max_particles=None, order="load"):
    """Seeded uniform plasma on the valid cells, weights w0*r (no profile).

    Returns a dict of float64 arrays x, y, z, px, py, pz, w, g_inv in the
    particle order described in the module docstring (order="load"), or
    permuted stably into the order of the grid's cell keys (order="cell",
    what Particles.align_parts would leave after a sort): the radial loading
    cells [ir*dr, (ir+1)*dr) straddle two grid rows (Rmin = -dr/2), so in
    load order each cell's particles form two runs far apart in memory."""
    Nx, Nr = grid_cfg["Nx"], grid_cfg["Nr"]
    dx = (grid_cfg["Xmax"] - grid_cfg["Xmin"]) / (Nx - 1)
    dr = grid_cfg["Rmax"] / (Nr - 1.5)
    npx, npr, npt = (int(v) for v in ppc)
    nppc = npx * npr * npt
    ncx, ncr = Nx - 3, Nr - 2
    Np = ncx * ncr * nppc
    if max_particles is not None:
        Np = min(Np, max_particles)
    rng = np.random.default_rng(seed)
    U = rng.random((Np, 3))
    p = np.arange(Np, dtype=np.int64)
    sx = p % npx
    sr = (p // npx) % npr
    st = (p // (npx * npr)) % npt
    cell = p // nppc
    ix = 1 + cell % ncx
    ir = cell // ncx
    xg = grid_cfg["Xmin"] + dx * ix
    x = xg + (sx + U[:, 0]) * (dx / npx)
    r = (ir + (sr + U[:, 1]) / npr) * dr
    th = (st + U[:, 2]) * (2 * np.pi / npt)
    w0 = 2 * np.pi * dx * dr * dens / nppc
    out = {"x": x, "y": r * np.sin(th), "z": r * np.cos(th),
           "w": w0 * r, "px": np.zeros(Np), "py": np.zeros(Np),
           "pz": np.zeros(Np), "g_inv": np.ones(Np)}
    if order == "cell":
        perm = np.argsort(cell_keys(grid_cfg, out["x"], out["y"], out["z"]), kind="stable")
        out = {k: np.ascontiguousarray(v[perm]) for k, v in out.items()}
    elif order != "load":
        raise ValueError(order)
    return out


def cell_keys(grid_cfg, x, y, z):
    """The sort key of index_and_sum_in_cell (particles_generic.cl:105-124)
    in float64 with the Grid's derived geometry (grid.py:68-103)."""
    Nx, Nr = grid_cfg["Nx"], grid_cfg["Nr"]
    dx = (grid_cfg["Xmax"] - grid_cfg["Xmin"]) / (Nx - 1)
    dr = grid_cfg["Rmax"] / (Nr - 1.5)
    r = np.sqrt(y * y + z * z)
    ix = np.floor((x - grid_cfg["Xmin"]) * (1.0 / dx)).astype(np.int64)
    ir = np.floor((r - (-0.5 * dr)) * (1.0 / dr)).astype(np.int64)
    ok = (ix > 0) & (ix < Nx - 2) & (ir >= 0) & (ir < Nr - 2)
    return np.where(ok, ix + ir * (Nx - 1), (Nr - 1) * (Nx - 1))


class ParticleSource:
    """Object with a ``DataDev`` of ``<name>_new`` arrays, the form
    ``add_new_particles(source=...)`` reads (particles_methods_cl.py:35-61)."""

    def __init__(self, arrays):
        self.DataDev = {k + "_new": v for k, v in arrays.items()}
