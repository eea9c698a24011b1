"""chimeracl_amd: the chimeraCL quasi-cylindrical PSATD PIC step on MI355X.

Python host code with the reference's wrapper API (Particles, Grid, Solver,
Transformer, PIC_loop, add_gausian_pulse) over the hand-written gfx950
kernels of libchimera_hip.so.  Importing the package does not touch the GPU;
the wrapper classes need the built library and a ROCm device.
"""
__version__ = "0.1.0"


def __getattr__(name):
    if name in ("Particles",):
        from .particles import Particles
        return Particles
    if name in ("Grid",):
        from .grid import Grid
        return Grid
    if name in ("Solver",):
        from .solver import Solver
        return Solver
    if name in ("PIC_loop", "loop_steps"):
        from . import pic_loop
        return getattr(pic_loop, name)
    if name == "add_gausian_pulse":
        from .laser import add_gausian_pulse
        return add_gausian_pulse
    if name == "Diagnostics":
        from .diagnostics import Diagnostics
        return Diagnostics
    if name == "Frame":
        from .frame import Frame
        return Frame
    if name == "Communicator":
        from .methods.generic_methods_hip import Communicator
        return Communicator
    raise AttributeError(name)
