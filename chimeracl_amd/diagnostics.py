"""Field and species records of a run (the reference's Diagnostics,
chimeraCL/diagnostics.py:21-141).

One record per recorded iteration, in a file named after the iteration
(nine digits).  The record layout is the reference's on-disk format, so that
its readers (e.g. chimera2VTK.py) work on these files:

  /data/info/iteration             the iteration
  /data/info/<key>                 Xgrid Rgrid dx dr dt Nx Nr M of the solver,
                                   FrameVelocity (0 without a frame)
  /data/fields/<name>              float32 stack [m0, Re m1, Im m1, Re m2, ...]
                                   of the real-space planes (rows 0..Nr-1)
  /data/species/species_<i>/<c>    component c of species i (float64); the
                                   weight w in pC per micron of wavelength

The reference writes HDF5 through h5py.  When h5py is importable, so does
this module (``<it>.h5``); otherwise the same paths become the keys (without
the leading '/') of a NumPy archive (``<it>.npz``).

A record is built in two stages: ``collect(it)`` evaluates every entry (the
field planes through the backward transform, the selected particle
components) into a {path: array} table, and a writer stores that table.
"""
from __future__ import annotations

import os

import numpy as np

try:  # the reference's on-disk container, if this installation has it
    import h5py
except ImportError:  # pragma: no cover - depends on the installation
    h5py = None

INFO_KEYS = ("Xgrid", "Rgrid", "dx", "dr", "dt", "Nx", "Nr", "M")


def info_path(key):
    return "/data/info/" + key


def field_path(name):
    return "/data/fields/" + name


def species_path(index, comp):
    return f"/data/species/species_{index}/{comp}"


def record_name(it):
    return f"{int(it):09d}"


# ---------------------------------------------------------------- writers
def _write_h5(fname, table):
    with h5py.File(fname + ".h5", "w") as f:
        for path, value in table.items():
            f[path] = value


def _write_npz(fname, table):
    np.savez(fname + ".npz", **{path.lstrip("/"): np.asarray(v) for path, v in table.items()})


def write_record(fname, table):
    """Stores a {path: array} table as one record file (fname + suffix)."""
    (_write_h5 if h5py is not None else _write_npz)(fname, table)


# ---------------------------------------------------------------- diagnostics
class Diagnostics:
    """Diagnostics(configs_in, solver, species=[], frame=None, path='diags',
    dtype_flds=float32, dtype_parts=float64).

    configs_in keys (defaults in brackets): 'Interval' [1], 'ScalarFields'
    [[]], 'VectorFields' [[]] (each expands to its x, y, z components),
    'Species' [{'Components': []}] with optional 'Selections' [(component,
    vmin, vmax), ...] (a particle is kept when vmin < value < vmax for every
    bound that is not None), 'w2pC' [the species' own w2pC].  The output
    folder is created, or emptied if it exists."""

    def __init__(self, configs_in, solver, species=[], frame=None, path="diags",
                 dtype_flds=np.float32, dtype_parts=np.float64):
        self.Args = configs_in
        for key, value in (("ScalarFields", []), ("VectorFields", []),
                           ("Species", {"Components": []}), ("Interval", 1)):
            self.Args.setdefault(key, value)
        self.solver, self.species, self.frame = solver, species, frame
        self.dtype_flds, self.dtype_parts = dtype_flds, dtype_parts
        self.path = os.path.join(os.getcwd(), path)
        os.makedirs(self.path, exist_ok=True)
        for name in os.listdir(self.path):
            os.remove(os.path.join(self.path, name))

    # names of the recorded fields, vectors expanded into components
    def field_names(self):
        names = list(self.Args["ScalarFields"])
        for vec in self.Args["VectorFields"]:
            names += [vec + c for c in "xyz"]
        return names

    def make_record(self, it):
        if it % self.Args["Interval"] != 0:
            return
        write_record(os.path.join(self.path, record_name(it)), self.collect(it))

    def collect(self, it):
        table = {info_path("iteration"): it}
        for key in INFO_KEYS:
            table[info_path(key)] = self.solver.Args[key]
        table[info_path("FrameVelocity")] = 0.0 if self.frame is None else self.frame.Args["Velocity"]
        for name in self.field_names():
            table[field_path(name)] = self.field_stack(name)
        for i, part in enumerate(self.species):
            for comp, vals in self.species_columns(part).items():
                table[species_path(i, comp)] = vals
        return table

    def field_stack(self, name):
        """Real-space planes of a field as float32 [m0, Re m1, Im m1, ...].
        As in the reference (diagnostics.py:125-135), the charge density and
        the current components first take a fresh charge deposit and a
        forward transform with smoothing of that field; every field is then
        taken through the backward transform."""
        s = self.solver
        if name == "rho" or name.startswith("J"):
            s.depose_charge(self.species)   # the reference re-deposits rho here, also for J
            s.fb_transform(scals=[name], dir=0)
            s.fields_smooth(flds=[name])
        s.fb_transform(scals=[name], dir=1)
        modes = [s.DataDev[f"{name}_m{m}"].detach().cpu().numpy() for m in range(s.Args["M"] + 1)]
        parts = [modes[0]]
        for v in modes[1:]:
            parts += [v.real, v.imag]
        return np.stack(parts).astype(self.dtype_flds)

    def species_columns(self, part):
        """{component: values} of one species, after the optional
        selections; weights converted with w2pC."""
        spec = self.Args["Species"]
        comps = spec["Components"]
        if part.Args["Np"] == 0:
            return {c: np.zeros(0, self.dtype_parts) for c in comps}
        keep = None
        for comp, lo, hi in spec.get("Selections", ()):
            v = part.host(comp)
            ok = np.ones(v.shape, bool)
            if lo is not None:
                ok &= v > lo
            if hi is not None:
                ok &= v < hi
            keep = ok if keep is None else keep & ok
        out = {}
        w2pc = self.dtype_parts(self.Args.get("w2pC", part.Args["w2pC"]))
        for comp in comps:
            v = part.host(comp).astype(self.dtype_parts)
            if keep is not None:
                v = v[keep]
            out[comp] = v * w2pc if comp == "w" else v
        return out
