"""Moving window of the PIC step (the reference's Frame, chimeraCL/frame.py:4-64).

A window that moves at ``Velocity`` advances by Velocity * dt * Steps once
every ``Steps`` steps (PIC_loop calls both methods below on those steps,
pic_loop.py:61-66).  An advance has two halves:

1. ``shift_grids``: each grid translates its longitudinal axis by the
   advance (``Grid.translate_window``), then rebuilds its cached geometry and,
   for a solver, the transform phases of the new origin.
2. ``inject_plasma``: for every species, fresh plasma fills the strip that
   the window has uncovered at its leading edge -- from the species' current
   right limit over the advance, radially over the whole grid -- and is
   appended; then every species is sorted on the grid and compacted, which
   drops the particles that fell behind the trailing edge, and its right
   limit is taken again from the particles of the last cell.

The configuration keys and their defaults are the reference's (Steps 1,
Velocity 0, dt 1, DensityProfiles None).  A species with
``Args['InjectorSource']`` is refilled from that source instead of its own
newly made domain (the reference's add_new_particles(source)).
"""
from __future__ import annotations

import numpy as np

_DEFAULTS = {"Steps": 1.0, "Velocity": 0.0, "dt": 1, "DensityProfiles": None}


class Frame:
    def __init__(self, configs_in, comm=None):
        self.Args = configs_in
        for key, value in _DEFAULTS.items():
            self.Args.setdefault(key, value)

    def advance(self, steps=None):
        """Distance the window moves over `steps` steps (default: Steps)."""
        n = self.Args["Steps"] if steps is None else steps
        return n * self.Args["dt"] * self.Args["Velocity"]

    def shift_grids(self, grids=(), steps=None):
        dist = self.advance(steps)
        for grid in grids:
            grid.translate_window(dist)

    def inject_plasma(self, species, grid, steps=None):
        dist = self.advance(steps)
        for specie in species:
            strip = _leading_strip(specie, grid, dist)
            specie.make_new_domain(strip, density_profiles=self.Args["DensityProfiles"])
            specie.add_new_particles(specie.Args.get("InjectorSource"))
        for specie in species:
            _compact(specie, grid)


def _leading_strip(specie, grid, dist):
    """Domain of the plasma to add: [right_lim, right_lim + dist) along x
    over the grid's whole radius.  A species without particles starts the
    strip one advance behind the grid's right edge."""
    if specie.Args["Np"] == 0:
        specie.Args["right_lim"] = grid.Args["Xmax"] - dist
    x0 = specie.Args["right_lim"]
    return {"Xmin": x0, "Xmax": x0 + dist, "Rmin": max(grid.Args["Rmin"], 0.0),
            "Rmax": grid.Args["Rmax"]}


def _compact(specie, grid):
    """Sort the species on the grid, keep only the particles in its valid
    cells (align), and set the right limit half a sub-cell beyond the
    largest x among the last prod(Nppc) + 1 particles (those of the last
    cells in sort order)."""
    specie.free_added()
    specie.sort_parts(grid=grid)
    specie.align_parts()
    if specie.Args["Np"] == 0:
        return
    tail = int(np.prod(specie.Args["Nppc"])) + 1
    last_x = float(specie.DataDev["x"][-tail:].max().item())
    specie.Args["right_lim"] = last_x + 0.5 * specie.Args["ddx"]
