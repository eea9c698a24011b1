"""Diagnostics: field and species records (chimeraCL/diagnostics.py:21-141).

Same record layout as the reference:
  /data/fields/<name>             stack [m0, Re m1, Im m1, ...] (float32)
  /data/species/species_<i>/<c>   particle components (w scaled by w2pC)
  /data/info/{Xgrid, Rgrid, dx, dr, dt, Nx, Nr, M, iteration, FrameVelocity}
one file per recorded iteration, named <iteration:09d>.  The reference
writes HDF5 with h5py; when h5py is importable this module does too
(<it>.h5), otherwise it writes the same paths as the keys of a NumPy archive
(<it>.npz), so the data and its naming are the same either way.
"""
from __future__ import annotations

import os

import numpy as np

try:  # optional: the reference's on-disk format
    import h5py
except ImportError:  # pragma: no cover - depends on the installation
    h5py = None


class _NpzRecord(dict):
    """Collects path -> array pairs and writes them as one .npz file."""

    def __init__(self, fname):
        super().__init__()
        self.fname = fname

    def close(self):
        np.savez(self.fname, **{k.lstrip("/"): np.asarray(v) for k, v in self.items()})


class Diagnostics:
    def __init__(self, configs_in, solver, species=[], frame=None, path="diags",
                 dtype_flds=np.float32, dtype_parts=np.float64):
        """diagnostics.py:22-55 (the output folder is created and emptied)."""
        self.Args = configs_in
        self.solver = solver
        self.species = species
        self.frame = frame
        self.dtype_flds = dtype_flds
        self.dtype_parts = dtype_parts
        self.base_str = "/data/"
        self.flds_str = "fields/"
        self.parts_str = "species/"
        self.info_str = "info/"
        self.generic_keys = ["Xgrid", "Rgrid", "dx", "dr", "dt", "Nx", "Nr", "M"]
        self.path = os.path.join(os.getcwd(), path) + "/"
        if not os.path.exists(self.path):
            os.makedirs(self.path)
        else:
            for fl in os.listdir(self.path):
                os.remove(self.path + fl)
        self.Args.setdefault("ScalarFields", [])
        self.Args.setdefault("VectorFields", [])
        self.Args.setdefault("Species", {"Components": []})
        self.Args.setdefault("Interval", 1)

    def make_record(self, it):
        """diagnostics.py:57-79."""
        if np.mod(it, self.Args["Interval"]) != 0:
            return
        it_str = str(it).zfill(9)
        if h5py is not None:
            self.record = h5py.File(self.path + it_str + ".h5", "w")
        else:
            self.record = _NpzRecord(self.path + it_str + ".npz")
        self.record[self.base_str + self.info_str + "iteration"] = it
        self.add_generic_info()
        for fld in self.Args["ScalarFields"]:
            self.add_field(fld)
        for fld in self.Args["VectorFields"]:
            for comp in ["x", "y", "z"]:
                self.add_field(fld + comp)
        self.add_species()
        self.record.close()

    def add_species(self):
        """diagnostics.py:81-113 (optional selections (comp, vmin, vmax))."""
        for i, part in enumerate(self.species):
            h5_path = self.base_str + self.parts_str + "species_" + str(i) + "/"
            indx = None
            if "Selections" in self.Args["Species"]:
                mask = np.ones(part.Args["Np"], dtype=bool)
                for comp_select, vmin, vmax in self.Args["Species"]["Selections"]:
                    val = part.host(comp_select)
                    if vmin is not None:
                        mask &= val > vmin
                    if vmax is not None:
                        mask &= val < vmax
                indx = np.nonzero(mask)
            for comp in self.Args["Species"]["Components"]:
                if part.Args["Np"] == 0:
                    vals = np.array([], dtype=self.dtype_parts)
                else:
                    vals = part.host(comp).astype(self.dtype_parts)
                    if indx is not None:
                        vals = vals[indx]
                if comp == "w":
                    vals = vals * self.dtype_parts(part.Args["w2pC"])
                self.record[h5_path + comp] = vals

    def add_generic_info(self):
        """diagnostics.py:115-123."""
        h5_path = self.base_str + self.info_str
        for key in self.generic_keys:
            self.record[h5_path + key] = self.solver.Args[key]
        velocity = 0. if self.frame is None else self.frame.Args["Velocity"]
        self.record[h5_path + "FrameVelocity"] = velocity

    def add_field(self, fld):
        """diagnostics.py:125-141: rho and J are deposited and transformed
        (with smoothing) again; every field is taken through the backward
        transform before it is written."""
        s = self.solver
        h5_path = self.base_str + self.flds_str + fld
        if fld == "rho" or fld[0] == "J":
            s.depose_charge(self.species)
            s.fb_transform(scals=[fld], dir=0)
            s.fields_smooth(flds=[fld])
        s.fb_transform(scals=[fld], dir=1)
        stack = [s.DataDev[fld + "_m0"].detach().cpu().numpy()[None].astype(self.dtype_flds)]
        for m in range(1, s.Args["M"] + 1):
            v = s.DataDev[f"{fld}_m{m}"].detach().cpu().numpy()[None]
            stack.append(v.real.astype(self.dtype_flds))
            stack.append(v.imag.astype(self.dtype_flds))
        self.record[h5_path] = np.concatenate(stack, axis=0)
