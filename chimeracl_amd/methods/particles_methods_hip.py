"""Particle methods (replaces chimeraCL/methods/particles_methods_cl.py).

push_coords, index_sort, dens_profile, align and the particle bookkeeping
call the HIP kernels of libchimera_hip.so; make_new_domain / make_new_beam
fill new particles on the device from a counter-based generator.  Particles
can also come in through add_particles(source=...), e.g. from
chimeracl_amd.config.load_plasma.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from .. import _lib
from .generic_methods_hip import GenericMethodsHIP, ptr

ALIGN_MAXARR = 16  # arrays per chm_align_multi launch (CHM_ALIGN_MAXARR)


def _grid_struct(grid):
    a = grid.Args
    return _lib.Grid(Nx=a["Nx"], Nr=a["Nr"], M=a["M"], xmin=a["Xmin"], dx_inv=a["dx_inv"],
                     rmin=a["Rmin"], dr_inv=a["dr_inv"])


class ParticleMethodsHIP(GenericMethodsHIP):
    def init_particle_methods(self):
        self.set_global_working_group_size()
        self._sort_ws = None
        self._keys_for = None
        self._iota = None
        self._spare = {}
        # bumped whenever positions, weights, momenta or the order change
        self.version = 0

    def _arg_names(self):
        if "Immobile" in self.Args:
            return ["x", "y", "z", "w"]
        return ["x", "y", "z", "px", "py", "pz", "w", "g_inv"]

    # ---------------------------------------------------------------- data
    def add_new_particles(self, source=None):
        """particles_methods_cl.py:40-61 (appends source.DataDev['<c>_new'])."""
        src = self.DataDev if source is None else source.DataDev
        real = self.comm.real
        for arg in self._arg_names():
            new = src[arg + "_new"]
            new = (torch.from_numpy(np.ascontiguousarray(new)) if isinstance(new, np.ndarray) else new)
            new = new.to(self.comm.device, real)
            old = self.DataDev.get(arg)
            self.DataDev[arg] = new.contiguous().clone() if (old is None or old.numel() == 0) \
                else torch.cat([old, new])
        pid = self.DataDev.get("pid")
        if pid is not None:  # appended particles take the next indices
            n = self.DataDev["x"].numel()
            self.DataDev["pid"] = torch.cat([pid, torch.arange(pid.numel(), n, dtype=torch.int32,
                                                               device=pid.device)])
        self.reset_num_parts()
        self.flag_sorted = False
        self.version += 1

    def make_new_domain(self, parts_in, density_profiles=None):
        """particles_methods_cl.py:66-147: a plasma block on the local grids of
        parts_in (Xmin, Xmax, Rmin, Rmax) into DataDev['<c>_new'], filled on the
        GPU (fill_grid, particles_generic.cl:33-84).  The per-cell theta offsets
        come from a counter-based generator (Args 'seed', advancing
        Args['rng_counter']) instead of PyOpenCL's.  Momenta p_c (px_c, ...)
        plus, for a component with a spread (dpx, ...), dp times a standard
        normal per particle (_fill_arr_randn(mu=p_c, sigma=dp),
        particles_methods_cl.py:114-147), drawn by Box-Muller from the same
        generator; as in the reference, missing p_c / dp keys are set to 0 in
        parts_in."""
        a = self.Args
        xmin, xmax, rmin, rmax = (parts_in[k] for k in ("Xmin", "Xmax", "Rmin", "Rmax"))
        Nx_loc = int(np.ceil((xmax - xmin) / a["dx"]) + 1)
        Nr_loc = int(np.round((rmax - rmin) / a["dr"]) + 1)
        xg = xmin + a["dx"] * np.arange(Nx_loc)
        rg = rmin + a["dr"] * np.arange(Nr_loc)
        a["right_lim"] = float(xg[-1])
        ncells = (Nx_loc - 1) * (Nr_loc - 1)
        npx, npr, npt = (int(v) for v in a["Nppc"])
        Np = ncells * npx * npr * npt
        dev, real = self.comm.device, self.comm.real
        new = {k: torch.empty(Np, dtype=real, device=dev) for k in ("x", "y", "z", "w")}
        moving = "Immobile" not in a
        pc, dp = [0.0, 0.0, 0.0], [0.0, 0.0, 0.0]
        if moving:
            for i, k in enumerate(("px", "py", "pz")):
                pc[i] = float(parts_in.setdefault(k + "_c", 0))
                dp[i] = float(parts_in.setdefault("d" + k, 0))
            for k in ("px", "py", "pz", "g_inv"):
                new[k] = torch.empty(Np, dtype=real, device=dev)
        grids = [torch.from_numpy(v).to(dev) for v in (xg, rg)]
        counter = int(a.get("rng_counter", 0))
        p = lambda k: ptr(new[k]) if k in new else None
        self.call("chm_fill_domain", p("x"), p("y"), p("z"), p("w"), p("px"), p("py"), p("pz"),
                  p("g_inv"), ptr(grids[0]), ptr(grids[1]), Nx_loc, ncells, npx, npr, npt,
                  float(a["w0"]), pc[0], pc[1], pc[2], dp[0], dp[1], dp[2], int(a.get("seed", 0)),
                  counter)
        a["rng_counter"] = counter + ncells + (4 * Np if any(dp) else 0)
        for k, t in new.items():
            self.DataDev[k + "_new"] = t
        if density_profiles is not None:
            for profile in density_profiles:
                if profile["coord"] != "x":
                    continue  # only longitudinal profiles, as in the reference
                self.dens_profile(profile["points"], profile["values"], parts_in["Xmin"],
                                  parts_in["Xmax"], coord=profile["coord"] + "_new",
                                  weight="w_new")

    def make_new_beam(self, parts_in):
        """particles_methods_cl.py:149-177: Np Gaussian particles (x_c, Lx, ...,
        px_c, dpx, ...; weights FullCharge/Np) into DataDev['<c>_new'], filled
        on the GPU from the counter-based generator (Args 'seed',
        advancing Args['rng_counter'] by 6 Np)."""
        a = self.Args
        Np = int(parts_in["Np"])
        b = _lib.Beam()
        for i, c in enumerate("xyz"):
            b.c[i] = float(parts_in[c + "_c"])
            b.L[i] = float(parts_in["L" + c])
            b.pc[i] = float(parts_in.get("p" + c + "_c", 0.0))
            b.dp[i] = float(parts_in.get("dp" + c, 0.0))
        b.w = float(parts_in["FullCharge"]) / Np
        dev, real = self.comm.device, self.comm.real
        new = {k: torch.empty(Np, dtype=real, device=dev)
               for k in ("x", "y", "z", "px", "py", "pz", "w", "g_inv")}
        counter = int(a.get("rng_counter", 0))
        self.call("chm_fill_beam", *(ptr(new[k]) for k in ("x", "y", "z", "px", "py", "pz", "w", "g_inv")),
                  Np, b, int(a.get("seed", 0)), counter)
        a["rng_counter"] = counter + 6 * Np
        for k, t in new.items():
            self.DataDev[k + "_new"] = t
    def reset_num_parts(self, Np=None):
        if Np is None:
            Np = int(self.DataDev["x"].numel())
        self.Args["Np"] = Np
        self.Args["Np_stay"] = Np
        self._sort_ws = None

    # ---------------------------------------------------------------- profile
    def dens_profile(self, x_prf, f_prf, xmin=None, xmax=None, coord="x", weight="w"):
        """particles_methods_cl.py:179-204 + profile_by_interpolant
        (particles_generic.cl:6-30); constant extension outside the knots."""
        xx = np.ascontiguousarray(x_prf, dtype=np.float64)
        ff = np.ascontiguousarray(f_prf, dtype=np.float64)
        dxm1 = 1.0 / (xx[1:] - xx[:-1])
        dev = self.comm.device
        self._prof = [torch.from_numpy(a).to(dev) for a in (xx, ff, dxm1)]
        x, w = self.DataDev[coord], self.DataDev[weight]
        self.call("chm_dens_profile", ptr(x), ptr(w), x.numel(), *(ptr(t) for t in self._prof),
                  xx.size)
        self.version += 1

    # ---------------------------------------------------------------- push
    def push_coords(self, mode="half", grid=None):
        """particles_methods_cl.py:206-223.  With grid, the cell keys of the
        next sort_parts(grid) are computed in the same pass."""
        if self.Args["Np"] == 0 or "Immobile" in self.Args:
            return
        dt = self.Args["dt_2"] if mode == "half" else self.Args["dt"]
        d = self.DataDev
        args = (ptr(d["x"]), ptr(d["y"]), ptr(d["z"]), ptr(d["px"]), ptr(d["py"]), ptr(d["pz"]),
                ptr(d["g_inv"]), dt, self.Args["Np"])
        if grid is None:
            self.call("chm_push_x", *args)
            self._keys_for = None
        else:
            self._ensure_sort_buffers(grid)
            _, ws, _ = self._sort_ws
            self.call("chm_push_x_keys", *args, grid._grid(), ptr(d["indx_in_cell"]), ptr(ws),
                      ws.numel())
            self._keys_for = grid
        self.flag_sorted = False
        self.version += 1

    # ---------------------------------------------------------------- sort
    def _ensure_sort_buffers(self, grid):
        Np, nbins = self.Args["Np"], grid.Args["Nxm1Nrm1"] + 1
        key = (Np, nbins)
        if self._sort_ws is not None and self._sort_ws[0] == key:
            return
        dev = self.comm.device
        nbytes = _lib.lib.chm_sort_workspace_bytes(Np, nbins)
        ws = torch.zeros(nbytes, dtype=torch.uint8, device=dev)
        self.DataDev["indx_in_cell"] = torch.empty(Np, dtype=torch.int32, device=dev)
        self.DataDev["sort_indx"] = torch.empty(Np, dtype=torch.int32, device=dev)
        self.DataDev["cell_offset"] = torch.empty(nbins + 1, dtype=torch.int32, device=dev)
        self.DataDev["Np_stay"] = torch.zeros(1, dtype=torch.int32, device=dev)
        self.DataDev["sum_in_cell"] = ws[: nbins * 4].view(torch.int32)
        self._sort_ws = (key, ws, None)

    def index_sort(self, grid):
        """particles_methods_cl.py:225-261: cell keys, histogram, offsets and
        the stable in-cell order; Np_stay stays on the device and is read
        lazily from Args."""
        self._ensure_sort_buffers(grid)
        _, ws, _ = self._sort_ws
        gs = grid._grid()
        d = self.DataDev
        if self._keys_for is grid:
            _lib.call("chm_sort_from_keys", self.Args["Np"], gs, ptr(d["indx_in_cell"]),
                      ptr(d["cell_offset"]), ptr(d["sort_indx"]), ptr(d["Np_stay"]), ptr(ws),
                      ws.numel(), self.comm.stream())
        else:
            self.call("chm_index_sort", ptr(d["x"]), ptr(d["y"]), ptr(d["z"]), self.Args["Np"],
                      gs, ptr(d["indx_in_cell"]), ptr(d["cell_offset"]), ptr(d["sort_indx"]),
                      ptr(d["Np_stay"]), ptr(ws), ws.numel())
        self._keys_for = None
        self._sorted_on = grid
        self.version += 1
        dev_np_stay = d["Np_stay"]
        self.Args.set_lazy("Np_stay", lambda: int(dev_np_stay.item()))

    def sorted_keys_ptr(self, grid):
        """Device address of the keys of the current sort's slots, which the
        sort left in its workspace (for the gather), or None when the species
        is not sorted on this grid."""
        if (not self.flag_sorted or self._sort_ws is None or self.Args["Np"] <= 0
                or getattr(self, "_sorted_on", None) is not grid):
            return None
        key, ws, _ = self._sort_ws
        if key != (self.Args["Np"], grid.Args["Nxm1Nrm1"] + 1):
            return None
        return _lib.lib.chm_sort_sorted_keys(ptr(ws), self.Args["Np"], grid._grid()) or None

    # ---------------------------------------------------------------- align
    def align_and_damp(self, comps_align):
        """particles_methods_cl.py:263-293: keep the Np_stay particles of the
        valid cells, in sorted order (used by the moving window)."""
        Np_stay = self.Args["Np_stay"]
        srt = self.DataDev["sort_indx"]
        for comp in comps_align:
            src = self.DataDev[comp]
            dst = torch.empty(Np_stay, dtype=src.dtype, device=src.device)
            self.call_raw("chm_align_elems", ptr(srt), Np_stay, ptr(src), ptr(dst), src.element_size())
            self.DataDev[comp] = dst
        self.DataDev.pop("pid", None)  # the reference renumbers the particles here
        self.reset_num_parts()
        self.flag_sorted = False
        self.version += 1

    def ensure_prepush_buffers(self, grid):
        """DataDev 'xh', 'yh', 'zh' (positions of the next first half push,
        written by a pre-pushing gather) and the sort buffers of grid."""
        self._ensure_sort_buffers(grid)
        d = self.DataDev
        for k in ("xh", "yh", "zh"):
            if k not in d or d[k].numel() != self.Args["Np"]:
                d[k] = torch.empty_like(d["x"])

    def use_prepushed(self, grid):
        """Takes the positions and keys a pre-pushing gather left (xh, yh,
        zh, indx_in_cell) as this step's push_coords('half', grid): the next
        sort_parts(grid) sorts those keys and the J deposit reads xh, yh, zh
        (Grid.depose_currents(half_in=True))."""
        if self.Args["Np"] == 0 or "Immobile" in self.Args:
            return
        self._keys_for = grid
        self.flag_sorted = False
        self.version += 1

    def ensure_pid(self):
        """DataDev['pid'] (the index of each particle without reordering),
        the identity if the storage has not been reordered yet."""
        d = self.DataDev
        if "pid" not in d and self.Args["Np"] > 0:
            d["pid"] = torch.arange(self.Args["Np"], dtype=torch.int32, device=d["x"].device)

    def reorder_parts(self, extra=(), dead=()):
        """Storage order only, no counterpart in the reference: every
        component of all Np particles is permuted into the current sorted
        order (none dropped), and sort_indx becomes the identity, so the
        particles' memory order follows their cells again.  Values, cell
        offsets and Np_stay are unchanged; particle indices change, and
        DataDev['pid'] keeps each particle's index in the numbering without
        reordering (the gather's skip of particles with index >= Np_stay uses
        it, so the same particles are skipped).  extra: further per-particle
        DataDev arrays to permute along (the pre-pushed positions and their
        keys); dead: arrays the step overwrites before it reads them (x, y, z
        when it consumes pre-pushed positions), left as they are.  Needs a
        current sort (flag_sorted).  One launch permutes up to 16 arrays into
        scratch tensors (chm_align_multi, sort_indx read once), which are
        then copied back: the arrays keep their device addresses, which
        captured step graphs hold."""
        Np = self.Args["Np"]
        if Np == 0 or not self.flag_sorted:
            return
        d = self.DataDev
        self.ensure_pid()
        if self._iota is None or self._iota.numel() != Np:
            self._iota = torch.arange(Np, dtype=torch.int32, device=d["x"].device)
        names = [c for c in self._arg_names() + ["pid"] + list(extra) if c not in dead]
        sp = self._spare
        for c in names:
            t = sp.get(c)
            if t is None or t.shape != d[c].shape or t.dtype != d[c].dtype:
                sp[c] = torch.empty_like(d[c])
        srt = d["sort_indx"]
        for i in range(0, len(names), ALIGN_MAXARR):
            grp = names[i:i + ALIGN_MAXARR]
            n = len(grp)
            src = (ctypes.c_void_p * n)(*(d[c].data_ptr() for c in grp))
            dst = (ctypes.c_void_p * n)(*(sp[c].data_ptr() for c in grp))
            eb = (ctypes.c_int32 * n)(*(d[c].element_size() for c in grp))
            self.call_raw("chm_align_multi", ptr(srt), Np, n, ctypes.addressof(src),
                          ctypes.addressof(dst), ctypes.addressof(eb))
        for c in names:
            d[c].copy_(sp[c])
        srt.copy_(self._iota)
        if not extra:
            self._keys_for = None
        self.version += 1

    def free_mp(self):
        pass

    def free_added(self):
        for arg in self._arg_names():
            self.DataDev.pop(arg + "_new", None)

    def gather_skip_ids(self):
        """The ids the gather compares with Np_stay (the reference's skip of
        particles with index >= Np_stay, grid_deposit_m1.cl:367-368): the
        unreordered index DataDev['pid'] if the storage was reordered, else
        None (the storage index itself).  With ``self.gather_skip = "cell"``
        (an option of the differentiable form, off by default) every visited
        particle -- every particle in a valid cell -- is pushed: the ids are
        all zero, so the index test never skips one."""
        if getattr(self, "gather_skip", "index") != "cell":
            return self.DataDev.get("pid")
        z = getattr(self, "_zero_ids", None)
        if z is None or z.numel() != self.Args["Np"]:
            z = self._zero_ids = torch.zeros(max(self.Args["Np"], 1), dtype=torch.int32,
                                             device=self.comm.device)
        return z

    def host(self, comp):
        """Host copy of a particle component (float64 numpy)."""
        return self.DataDev[comp].detach().double().cpu().numpy()

