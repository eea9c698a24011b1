"""PSATD solver methods (replaces chimeraCL/methods/solver_methods_cl.py)."""
from __future__ import annotations

from .generic_methods_hip import GenericMethodsHIP


class SolverMethodsHIP(GenericMethodsHIP):
    def init_solver_methods(self):
        self.set_global_working_group_size()

    def advance_fields(self, vecs):
        """solver_methods_cl.py:43-62 (advance_e_g_m for every mode).
        vecs = [E, G, J, dN0, dN1]."""
        E, G, J, N0, N1 = (self._fb_arr(self._vec(v)) for v in vecs)
        self.call("chm_advance_fields", self._sp, E, G, J, N0, N1)

    def damp_edges(self, flds=("E", "G"), advance_with=None):
        """Solver.damp_fields (solver.py:32-35): half backward transform,
        profile_edges (solver_methods_cl.py:64-83), half forward transform,
        fused per spectral row.  advance_with = [J, dN0, dN1] applies the
        PSATD advance first in the same kernel."""
        E, G = (self._fb_arr(self._vec(v)) for v in flds)
        if advance_with is None:
            self.call("chm_damp_fields", self._sp, E, G, 0, None, None, None)
        else:
            J, N0, N1 = (self._fb_arr(self._vec(v)) for v in advance_with)
            self.call("chm_damp_fields", self._sp, E, G, 1, J, N0, N1)
