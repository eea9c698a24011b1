"""Grid: geometry and real-space field planes (chimeraCL/grid.py:5-127)."""
from __future__ import annotations

import numpy as np
import torch

from .methods.generic_methods_hip import LazyArgs, default_comm
from .methods.grid_methods_hip import GridMethodsHIP


class Grid(GridMethodsHIP):
    def __init__(self, configs_in, comm=None):
        self.import_comm(comm if comm is not None else default_comm())
        self.set_global_working_group_size()
        self._process_configs(configs_in)
        if "vec_comps" not in self.Args:
            self.Args["vec_comps"] = ["x", "y", "z"]
        self.init_grid_methods()
        self.DataDev = {}
        self._init_grid_data_on_dev()
        self.send_args_to_dev()

    def depose_charge(self, species=[]):
        """grid.py:23-31: rho of every species (immobile ones included)."""
        self.zero_planes(["rho"])
        for parts in species:
            self.depose_scalar(parts, "w", "rho", charge=parts.Args["charge"])
        self.postproc_depose_scalar("rho")

    def depose_currents(self, species=[], push_half=False, half_in=False):
        """grid.py:33-54: J of every mobile species.  With push_half, each
        mobile species is also pushed by dt/2 and keyed on this grid in the
        same pass, as push_coords('half', grid=self) would do next.  With
        half_in, the deposit reads the positions from DataDev 'xh', 'yh',
        'zh' (left there by gather_and_push(prepush=True)) and the push writes
        x, y, z."""
        self.zero_planes(["J" + c for c in self.Args["vec_comps"]])
        momenta = ["p" + c for c in self.Args["vec_comps"]]
        for parts in species:
            if "Immobile" in parts.Args:
                continue
            if push_half:
                self.depose_vector_push(parts, momenta, ["g_inv", "w"], "J",
                                        charge=parts.Args["charge"], dt=parts.Args["dt_2"],
                                        half_in=half_in)
            else:
                self.depose_vector(parts, momenta, ["g_inv", "w"], "J", charge=parts.Args["charge"])
        self.postproc_depose_vector("J")

    def gather_and_push(self, species=[], prepush=False):
        """grid.py:56-66.  With prepush (consecutive steps with nothing in
        between), each mobile species' next first half push and its sort keys
        on this grid are computed in the same pass, into DataDev 'xh', 'yh',
        'zh' and 'indx_in_cell'; x, y, z keep the step's positions."""
        for fld in ("E", "B"):
            self.preproc_project_vec(fld)
        for parts in species:
            if "Immobile" in parts.Args:
                continue
            self._gather_and_push(parts, ["E", "B"], prepush=prepush)

    def _process_configs(self, configs_in):
        """grid.py:68-103."""
        a = self.Args = LazyArgs(configs_in)
        if "M" not in a:
            a["M"] = 0
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
        a["NxNr_4"] = a["Nr"] // 2 * a["Nx"] // 2
        a["dont_send"] = []
        a["dont_keep"] = []
        if a["M"] > 4:
            raise ValueError("modes above m=4 are not supported")

    def _init_grid_data_on_dev(self):
        """grid.py:106-127: E, B, J, G, rho planes; m0 real, m>=1 complex."""
        names = [f + c for f in ("E", "B", "J", "G") for c in self.Args["vec_comps"]] + ["rho"]
        shp = (self.Args["Nr"], self.Args["Nx"])
        for n in names:
            self.DataDev[n + "_m0"] = torch.zeros(shp, dtype=self.comm.real, device=self.comm.device)
            for m in range(1, self.Args["M"] + 1):
                self.DataDev[f"{n}_m{m}"] = torch.zeros(shp, dtype=self.comm.cplx, device=self.comm.device)
