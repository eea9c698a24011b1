"""Solver: the PSATD field solver (chimeraCL/solver.py:9-68)."""
from __future__ import annotations

from .grid import Grid
from .methods.generic_methods_hip import default_comm
from .methods.solver_methods_hip import SolverMethodsHIP
from .transformer import Transformer


class Solver(Grid, Transformer, SolverMethodsHIP):
    def __init__(self, configs_in, comm=None):
        self.import_comm(comm if comm is not None else default_comm())
        self.set_global_working_group_size()
        self._process_configs(configs_in)
        self.Args["vec_comps"] = ["x", "y", "z"]
        self.init_solver_methods()
        self.init_grid_methods()
        self.DataDev = {}
        self._init_grid_data_on_dev()
        self.send_args_to_dev()
        self.init_transformer()

    def push_fields(self):
        """solver.py:29-30."""
        self.advance_fields(vecs=["E", "G", "J", "dN0", "dN1"])

    def damp_fields(self):
        """solver.py:32-35.  Needs 'DampCells' in the configuration, as in the
        reference (solver.py:32-35 uses the profile _init_field_damping makes)."""
        if "DampCells" not in self.Args:
            raise KeyError("DampProfile")
        self.damp_edges(("E", "G"))

    def restore_B_fb(self):
        """solver.py:37-39: B = rot(G) / w^2."""
        self.field_rot("G", "B", poisson=True)
