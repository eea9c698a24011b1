"""Builds a GPU simulation and its CPU-oracle twin from the same seeded
inputs, and compares their states.  Test helper only."""
from __future__ import annotations

import numpy as np

from chimeracl_amd import config as C


def tiny_grid(Nx=64, Nr=24, M=1, damp=4, name="C1"):
    g = C.grid_config(name, Nx=Nx, Nr=Nr, M=M)
    g["DampCells"] = damp
    return g


def tiny_laser(g, a0=2.0):
    x0 = g["Xmax"] - 0.3
    return {"k0": 1.0, "a0": a0, "x0": x0, "Lx": 0.3, "R": 0.8, "x_foc": x0}


def make_pair(g, laser, ppc=(2, 2, 4), dens=0.002, precision="f64", profile=None, seed=0,
              with_oracle=True):
    import chimeracl_amd as CA
    from chimeracl_amd.methods.generic_methods_hip import Communicator
    from oracle import oracle as O
    comm = Communicator(precision=precision)
    solver = CA.Solver(dict(g), comm)
    if laser is not None:
        CA.add_gausian_pulse(solver, dict(laser))
    eo, io = C.species_config("C1", solver.Args, ppc=ppc, dens=dens)
    eo["Nppc"], io["Nppc"] = ppc, ppc
    eo["dens"] = io["dens"] = dens
    arr = C.load_plasma(g, ppc, dens, seed=seed)
    eons, ions = CA.Particles(dict(eo), comm), CA.Particles(dict(io), comm)
    eons.add_particles(source=C.ParticleSource(arr))
    ions.add_particles(source=C.ParticleSource(arr))
    if profile is not None:
        eons.dens_profile(*profile)
        ions.dens_profile(*profile)
    loop = CA.PIC_loop(solvers=[solver], species=[eons, ions])
    gpu = dict(solver=solver, eons=eons, ions=ions, loop=loop)
    if not with_oracle:
        return gpu, None
    s = O.OracleSolver(dict(g))
    if laser is not None:
        O.add_gausian_pulse(s, dict(laser))
    e, i = O.OracleParticles(dict(eo)), O.OracleParticles(dict(io))
    e.add_particles(arr)
    i.add_particles(arr)
    if profile is not None:
        e.dens_profile(*profile)
        i.dens_profile(*profile)
    orc = dict(solver=s, eons=e, ions=i, loop=O.OracleLoop(s, [e, i]))
    return gpu, orc


def rel_l2(a, b):
    a, b = np.asarray(a), np.asarray(b)
    d = np.linalg.norm((a - b).ravel())
    n = np.linalg.norm(b.ravel())
    return d / n if n > 0 else d


def field_host(solver, name):
    return solver.DataDev[name].detach().cpu().numpy()


def compare_fields(gpu_solver, orc_solver, names, rows=slice(1, None)):
    out = {}
    for n in names:
        out[n] = rel_l2(field_host(gpu_solver, n)[rows], orc_solver.f[n][rows])
    return out


def real_names(M, flds=("E", "B", "J"), scal=("rho",)):
    names = []
    for f in flds:
        for c in "xyz":
            for m in range(M + 1):
                names.append(f"{f}{c}_m{m}")
    for s in scal:
        for m in range(M + 1):
            names.append(f"{s}_m{m}")
    return names
