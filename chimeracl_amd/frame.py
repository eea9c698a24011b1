"""Frame: the moving window of chimeraCL (chimeraCL/frame.py:4-64).

shift_grids moves the window origin of the grids by Steps*dt*Velocity;
inject_plasma fills the strip uncovered at the right edge with new plasma
(make_new_domain on the GPU), then sorts and aligns every species so that the
particles that left the window are dropped.  PIC_loop calls both every
``Steps`` steps (pic_loop.py:61-66).
"""
from __future__ import annotations

import numpy as np


class Frame:
    def __init__(self, configs_in, comm=None):
        self._process_configs(configs_in)

    def _process_configs(self, configs_in):
        """frame.py:8-20."""
        self.Args = configs_in
        if "Steps" not in self.Args:
            self.Args["Steps"] = 1.
        if "Velocity" not in self.Args:
            self.Args["Velocity"] = 0.
        if "dt" not in self.Args:
            self.Args["dt"] = 1
        if "DensityProfiles" not in self.Args:
            self.Args["DensityProfiles"] = None

    def shift_grids(self, grids=[], steps=None):
        """frame.py:22-30: Xmin, Xmax, Xgrid += steps*dt*Velocity (host and
        device copies); the grids then refresh their cached geometry and
        the transform phases."""
        if steps is None:
            steps = self.Args["Steps"]
        x_shift = steps * self.Args["dt"] * self.Args["Velocity"]
        for grid in grids:
            for arg in ("Xmax", "Xmin", "Xgrid"):
                grid.Args[arg] = grid.Args[arg] + x_shift
                if arg in grid.DataDev:
                    grid.DataDev[arg] += x_shift
            if hasattr(grid, "_frame_moved"):
                grid._frame_moved()

    def inject_plasma(self, species, grid, steps=None):
        """frame.py:32-64."""
        if steps is None:
            steps = self.Args["Steps"]
        x_shift = steps * self.Args["dt"] * self.Args["Velocity"]
        for specie in species:
            if specie.Args["Np"] == 0:
                specie.Args["right_lim"] = grid.Args["Xmax"] - x_shift
            inject_domain = {"Xmin": specie.Args["right_lim"]}
            inject_domain["Xmax"] = inject_domain["Xmin"] + x_shift
            inject_domain["Rmin"] = grid.Args["Rmin"] * (grid.Args["Rmin"] > 0)
            inject_domain["Rmax"] = grid.Args["Rmax"]
            specie.make_new_domain(inject_domain, density_profiles=self.Args["DensityProfiles"])
            if "InjectorSource" in specie.Args:
                specie.add_new_particles(specie.Args["InjectorSource"])
            else:
                specie.add_new_particles()
        for specie in species:
            specie.free_added()
            specie.sort_parts(grid=grid)
            specie.align_parts()
            num_ppc = int(np.prod(specie.Args["Nppc"]) + 1)
            x_max = float(specie.DataDev["x"][-num_ppc:].max().item())
            specie.Args["right_lim"] = x_max + 0.5 * specie.Args["ddx"]
