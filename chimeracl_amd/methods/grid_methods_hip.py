"""Grid methods (replaces chimeraCL/methods/grid_methods_cl.py): deposits,
axis post-processing, ghost-row warp and the gather + push."""
from __future__ import annotations

import torch

from .. import _lib
from .generic_methods_hip import GenericMethodsHIP, ptr
from .particles_methods_hip import _grid_struct


class GridMethodsHIP(GenericMethodsHIP):
    def init_grid_methods(self):
        self.set_global_working_group_size()
        if "vec_comps" not in self.Args:
            self.Args["vec_comps"] = ["x", "y", "z"]
        self._pcache = {}

    # ---------------------------------------------------------------- structs
    def _planes(self, name):
        """chm_planes_t of the real-space planes '<name>_m<m>'."""
        st = self._pcache.get(name)
        if st is None:
            M = self.Args["M"]
            st = _lib.Planes()
            st.m0 = ptr(self.DataDev[name + "_m0"])
            for m in range(1, M + 1):
                st.mc[m - 1] = ptr(self.DataDev[f"{name}_m{m}"])
            self._pcache[name] = st
        return st

    def _planes_arr(self, names):
        key = ("arr",) + tuple(names)
        arr = self._pcache.get(key)
        if arr is None:
            arr = (_lib.Planes * len(names))(*[self._planes(n) for n in names])
            self._pcache[key] = arr
        return arr

    def _grid(self):
        g = self._pcache.get("grid")
        if g is None:
            g = self._pcache["grid"] = _grid_struct(self)
        return g

    def translate_window(self, dist):
        """Moves the longitudinal axis by dist (Frame.shift_grids,
        frame.py:22-30: Xmin, Xmax and Xgrid, host and device copies), then
        rebuilds the cached geometry and, for a solver, the transform phases
        e^(-/+ i kx Xmin) of the new origin."""
        for key in ("Xmin", "Xmax", "Xgrid"):
            self.Args[key] = self.Args[key] + dist
            if key in self.DataDev:
                self.DataDev[key] += dist
        self._pcache.pop("grid", None)
        if hasattr(self, "_refresh_phases"):
            self._refresh_phases()

    def _dep_ws(self):
        ws = self._pcache.get("dep_ws")
        if ws is None:
            a = self.Args
            n = _lib.lib.chm_deposit_workspace_bytes(a["Nx"], a["Nr"], a["M"], 1)
            ws = self._pcache["dep_ws"] = torch.empty(n, dtype=torch.uint8, device=self.comm.device)
        return ws

    # ---------------------------------------------------------------- deposits
    def depose_scalar(self, parts, src_scalar, dest_fld, charge):
        """grid_methods_cl.py:44-69 (adds into '<dest_fld>_m*').

        An immobile species keeps its per-cell sums between calls: while its
        particles are unchanged (Particles.version) only the node pass runs,
        which adds the same sums in the same order."""
        if parts.Args["Np"] <= 0:
            return
        if not parts.flag_sorted:  # deposits read the sorted order of this grid
            parts.sort_parts(self)
        d = parts.DataDev
        planes = self._planes_arr([dest_fld])
        if "Immobile" in parts.Args:
            key = (id(parts), src_scalar, int(charge))
            cache = self._pcache.setdefault("staged", {})
            ent = cache.get(key)
            if ent is not None and ent[0] == parts.version:
                ws = ent[1]
                self.call("chm_deposit_add_staged", ptr(d["cell_offset"]), self._grid(), planes, 0,
                          ptr(ws), ws.numel())
                return
            a = self.Args
            n = _lib.lib.chm_deposit_workspace_bytes(a["Nx"], a["Nr"], a["M"], 0)
            ws = ent[1] if ent is not None else torch.empty(n, dtype=torch.uint8, device=self.comm.device)
            cache[key] = (parts.version, ws)
        else:
            ws = self._dep_ws()
        self.call("chm_deposit_scalar", ptr(d["sort_indx"]), ptr(d["cell_offset"]), ptr(d["x"]),
                  ptr(d["y"]), ptr(d["z"]), ptr(d[src_scalar]), int(charge), self._grid(),
                  planes, ptr(ws), ws.numel())

    def depose_vector(self, parts, vec, factors, vec_fld, charge):
        """grid_methods_cl.py:71-95 (adds w*g_inv*charge*u into '<vec_fld><c>_m*')."""
        if parts.Args["Np"] <= 0:
            return
        if not parts.flag_sorted:
            parts.sort_parts(self)
        d, ws = parts.DataDev, self._dep_ws()
        comps = self.Args["vec_comps"]
        self.call("chm_deposit_vector", ptr(d["sort_indx"]), ptr(d["cell_offset"]), ptr(d["x"]),
                  ptr(d["y"]), ptr(d["z"]), *(ptr(d[v]) for v in vec), ptr(d[factors[0]]),
                  ptr(d[factors[1]]), int(charge), self._grid(),
                  self._planes_arr([vec_fld + c for c in comps]), ptr(ws), ws.numel())

    def depose_vector_push(self, parts, vec, factors, vec_fld, charge, dt, half_in=False):
        """depose_vector, then the species' next coordinate push by dt and
        the cell keys of its next sort on this grid, in one pass (the
        push_coords + key pass that follow depose_currents in the step,
        pic_loop.py:82-91); positions and keys equal those of push_coords
        with grid=self."""
        if parts.Args["Np"] <= 0:
            return
        if not parts.flag_sorted:
            parts.sort_parts(self)
        parts._ensure_sort_buffers(self)
        d, ws = parts.DataDev, self._dep_ws()
        comps = self.Args["vec_comps"]
        src = ("xh", "yh", "zh") if half_in else ("x", "y", "z")
        self.call("chm_deposit_vector_push", ptr(d["sort_indx"]), ptr(d["cell_offset"]),
                  *(ptr(d[k]) for k in src), ptr(d["x"]), ptr(d["y"]), ptr(d["z"]),
                  *(ptr(d[v]) for v in vec), ptr(d[factors[0]]),
                  ptr(d[factors[1]]), int(charge), dt, parts.Args["Np"], ptr(d["indx_in_cell"]),
                  self._grid(), self._planes_arr([vec_fld + c for c in comps]), ptr(ws), ws.numel())
        parts._keys_for = self
        parts.flag_sorted = False
        parts.version += 1

    def zero_planes(self, names):
        """set_to(.., 0) of every mode plane of the named fields, one launch."""
        a = self.Args
        self.call("chm_zero_planes", self._planes_arr(names), len(names), a["Nx"], a["Nr"], a["M"])

    def _postproc(self, names):
        a = self.Args
        self.call("chm_postproc_deposit", self._planes_arr(names), len(names),
                  ptr(self.DataDev["dV_inv"]), a["Nx"], a["Nr"], a["M"])

    def postproc_depose_scalar(self, fld):
        """grid_methods_cl.py:98-124: treat_axis then divide by dV."""
        self._postproc([fld])

    def postproc_depose_vector(self, vec_fld):
        """grid_methods_cl.py:126-151."""
        self._postproc([vec_fld + c for c in self.Args["vec_comps"]])

    def preproc_project_vec(self, vec_fld):
        """grid_methods_cl.py:153-166: ghost row of E or B."""
        a = self.Args
        names = [vec_fld + c for c in self.Args["vec_comps"]]
        self.call("chm_warp_axis", self._planes_arr(names), len(names), a["Nx"], a["Nr"], a["M"])

    def _gather_and_push(self, parts, flds, prepush=False):
        """grid_methods_cl.py:168-192 (visits the sorted slots, as the
        reference kernel does; the species must be sorted for this grid).
        prepush: see Grid.gather_and_push."""
        if parts.Args["Np"] == 0:
            return
        d = parts.DataDev
        comps = self.Args["vec_comps"]
        E = self._planes_arr([flds[0] + c for c in comps])
        B = self._planes_arr([flds[1] + c for c in comps])
        skey = parts.sorted_keys_ptr(self)
        if prepush:
            parts.ensure_prepush_buffers(self)
            self.call("chm_gather_push_pre", ptr(d["x"]), ptr(d["y"]), ptr(d["z"]), ptr(d["px"]),
                      ptr(d["py"]), ptr(d["pz"]), ptr(d["g_inv"]), parts.Args["Np"],
                      ptr(d["sort_indx"]), ptr(d["cell_offset"]), ptr(d["Np_stay"]),
                      ptr(parts.gather_skip_ids()), skey, parts.Args["FactorPush"],
                      parts.Args["dt_2"], ptr(d["xh"]), ptr(d["yh"]), ptr(d["zh"]),
                      ptr(d["indx_in_cell"]), self._grid(), E, B)
            return
        args = (ptr(d["x"]), ptr(d["y"]), ptr(d["z"]), ptr(d["px"]), ptr(d["py"]), ptr(d["pz"]),
                ptr(d["g_inv"]), parts.Args["Np"], ptr(d["sort_indx"]), ptr(d["Np_stay"]))
        ids = parts.gather_skip_ids()
        if ids is not None:  # reordered storage (or the "cell" skip rule)
            self.call("chm_gather_push_ids", *args, ptr(ids), skey, parts.Args["FactorPush"],
                      self._grid(), E, B)
        else:
            self.call("chm_gather_push", *args, skey, parts.Args["FactorPush"], self._grid(), E, B)
