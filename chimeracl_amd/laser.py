"""Gaussian laser pulse initialiser (chimeraCL/laser.py:3-37), on the GPU."""
from __future__ import annotations

import numpy as np

from . import _lib
from .methods.generic_methods_hip import ptr


def add_gausian_pulse(solver, laser):
    """Ez_m0 pulse, forward transform, G = -i w sign(kx) E, vacuum
    propagation over x0 - x_foc, then B = rot(G)/w^2 and the backward
    transforms of B and E (laser.py:3-37).  The pulse and the spectral
    set-up run in FP64; in the mixed form the FP64 result is rounded once
    into the FP32 planes."""
    k0 = 2 * np.pi * laser["k0"]
    a0, Lx, R, x0, x_foc = (laser[k] for k in ("a0", "Lx", "R", "x0", "x_foc"))
    a, D = solver.Args, solver.DataDev
    stream = solver.comm.stream()
    if not solver.comm.mixed:
        Ez, Ez_fb, Gz_fb, sp = D["Ez_m0"], D["Ez_fb_m0"], D["Gz_fb_m0"], solver._sp
        _lib.call("chm_laser_pulse_f64", ptr(Ez), a["Nx"], a["Nr"], a["Xmin"], a["dx"], a["Rmin"],
                  a["dr"], k0, a0, x0, Lx, R, stream)
        solver.fb_transform(scals=["Ez"], dir=0)
        _lib.call("chm_laser_spectral_f64", sp, ptr(D["w_m0"]), ptr(Ez_fb), ptr(Gz_fb), x0 - x_foc, stream)
    else:
        _mixed_pulse(solver, k0, a0, x0, Lx, R, x_foc)
    solver.restore_B_fb()
    solver.fb_transform(vects=["B", "E"], dir=1)


def _mixed_pulse(solver, k0, a0, x0, Lx, R, x_foc):
    """Builds the FP64 spectral pulse with an FP64 twin of the solver's
    tables, then rounds E_fb, G_fb (mode 0) into the FP32 solver."""
    import torch
    from .methods.generic_methods_hip import Communicator
    from .solver import Solver
    cfg = {k: solver.Args[k] for k in ("Xmin", "Xmax", "Nx", "Nr", "M", "dt")}
    cfg["Rmax"] = solver.Args["dr"] * (solver.Args["Nr"] - 1.5)
    twin = Solver(cfg, Communicator(device=solver.comm.device.index, precision="f64"))
    laser = {"k0": k0 / (2 * np.pi), "a0": a0, "x0": x0, "Lx": Lx, "R": R, "x_foc": x_foc}
    a, D = twin.Args, twin.DataDev
    stream = twin.comm.stream()
    _lib.call("chm_laser_pulse_f64", ptr(D["Ez_m0"]), a["Nx"], a["Nr"], a["Xmin"], a["dx"], a["Rmin"],
              a["dr"], k0, a0, x0, Lx, R, stream)
    twin.fb_transform(scals=["Ez"], dir=0)
    _lib.call("chm_laser_spectral_f64", twin._sp, ptr(D["w_m0"]), ptr(D["Ez_fb_m0"]), ptr(D["Gz_fb_m0"]),
              x0 - x_foc, stream)
    solver.DataDev["Ez_fb_m0"].copy_(D["Ez_fb_m0"].to(torch.complex64))
    solver.DataDev["Gz_fb_m0"].copy_(D["Gz_fb_m0"].to(torch.complex64))
    del twin
