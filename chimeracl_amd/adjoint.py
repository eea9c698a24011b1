"""Differentiable form of the PIC step: reverse-mode gradients through the
set-up (laser amplitude a0 and waist R = w0, laser.py:3-37; density-profile
knot values, particles_methods_cl.py:179-204) and any number of steps of
PIC_loop.step (pic_loop.py:57-142), behind torch.autograd.Function.

Forward: the ordinary step (graph-replayed) with the state saved every
``ckpt`` steps.  Backward: per checkpoint segment, last segment first, the
steps are recomputed from the checkpoint while the data each step's adjoint
needs are recorded (positions and momenta at the step's start, the gathered
E and B planes, the sorted order the gather used); then the recorded steps
are reversed with the hand-written adjoint kernels of libchimera_hip.so:

  forward step                              adjoint, in reverse order
  x1 = x0 + u0 g0 dt/2 ; sort               push_x adjoint
  J  = deposit_J(x1, u0, g0, w)             J deposit adjoint  (per particle)
  x2 = x1 + u0 g0 dt/2 ; sort               push_x adjoint
  rho = deposit_rho(x2, w) + rho_ions(w_i)  rho deposit adjoint (ions: summed)
  rho^, J^ = S T_f(rho, J)                  T_f^H S  (backward transform with
                                             the transposed-and-scaled tables)
  dN0 = dN1 ; dN1 = grad(rho^)              grad^H ; the dN1 -> dN0 carry
  E, G = Damp(PSATD(E, G, J^, dN0, dN1))    PSATD^T Damp   (Damp is self-adjoint)
  B^ = rot(G)/w^2                           rot^H(Bbar/w^2)
  E, B real = T_b(E^, B^), ghost rows       warp^H, T_b^H (forward transform,
                                             transposed tables)
  u, g = Boris(u0, gather(E, B; x2))        gather + push adjoint, then the
                                             field adjoints scattered by the
                                             sorted two-pass deposit

The state carried from step to step is (x, y, z, px, py, pz, g_inv) of the
mobile species and the spectral E, G, dN1; its adjoint is carried the same
way.  Every kernel adds without float atomics, so the gradient is
deterministic.  Parameters enter only at set-up: the gradients with
respect to a0 and w0 are the inner products of the step-0 field adjoints
with the tangents of the laser initialisation, and those with respect to
the knot values come from the accumulated weight adjoints.
"""
from __future__ import annotations

import ctypes
import math
import time

import numpy as np
import torch

from . import _lib
from .methods.generic_methods_hip import Communicator, ptr

_COMPS = ("x", "y", "z")


def _planes(tensors):
    """chm_planes_t from [m0, m1, ...] real-space tensors."""
    st = _lib.Planes()
    st.m0 = ptr(tensors[0])
    for m, t in enumerate(tensors[1:]):
        st.mc[m] = ptr(t)
    return st


def _planes_arr(list_of_lists):
    return (_lib.Planes * len(list_of_lists))(*[_planes(l) for l in list_of_lists])


def _fb_arr(list_of_lists):
    arr = (_lib.Fb * len(list_of_lists))()
    for i, l in enumerate(list_of_lists):
        for m, t in enumerate(l):
            arr[i].m[m] = ptr(t)
    return arr


class AdjointSpectral:
    """chm_spectral_t for the adjoint operators of a Solver: the forward
    struct with the DHT-type tables replaced by their transposes, scaled so
    that chm_transform_* dir 0 applies T_b^H and dir 1 applies T_f^H:
    DHT <- DHT_inv^T / Nx, DHT_inv <- Nx DHT^T, dDp <- dDp^T, dDm <- dDm^T
    (transformer.py:65-97 for the forward tables)."""

    def __init__(self, solver):
        self.solver = solver
        a, D, comm = solver.Args, solver.DataDev, solver.comm
        Nx, Nk, M = a["Nx"], a["Nr"] - 1, a["M"]
        self.sp = _lib.Spectral.from_buffer_copy(solver._sp)
        nel = _lib.lib.chm_dht_table_elems(Nk, Nk)
        st, dev = comm.stream(), comm.device
        self._keep = []
        for m in range(M + 1):
            tabs = {"DHT": D[f"DHT_inv_m{m}"].t() / Nx, "DHT_inv": D[f"DHT_m{m}"].t() * Nx,
                    "dDp": D[f"dDHT_plus_m{m}"].t(), "dDm": D[f"dDHT_minus_m{m}"].t()}
            for slot, T in tabs.items():
                T = T.contiguous()
                self._keep.append(T)
                if not comm.mixed:
                    pad = torch.empty(nel, dtype=torch.float64, device=dev)
                    _lib.call("chm_pad_table", ptr(T), Nk, Nk, ptr(pad), None, None, None, st)
                    getattr(self.sp, slot)[m] = ptr(pad)
                    self._keep.append(pad)
                elif comm.dht_kind == _lib.CHM_DHT_BF16X3:
                    hi = torch.empty(nel, dtype=torch.int16, device=dev)
                    lo = torch.empty(nel, dtype=torch.int16, device=dev)
                    _lib.call("chm_pad_table", ptr(T), Nk, Nk, None, None, ptr(hi), ptr(lo), st)
                    getattr(self.sp, slot)[m] = ptr(hi)
                    getattr(self.sp, slot + "_lo")[m] = ptr(lo)
                    self._keep += [hi, lo]
                else:
                    pad = torch.empty(nel, dtype=torch.float32, device=dev)
                    _lib.call("chm_pad_table", ptr(T), Nk, Nk, None, ptr(pad), None, None, st)
                    getattr(self.sp, slot)[m] = ptr(pad)
                    self._keep.append(pad)


class DiffPIC:
    """A differentiable laser-plasma run: a Solver, one mobile electron
    species and its immobile ions at identical positions, a Gaussian laser
    and a piecewise-linear density profile through fixed knot positions.

    params: a0, R (= w0, the waist) and the knot values.  ``forward`` runs
    the set-up and ``nsteps`` steps, saving the state every ``ckpt`` steps;
    ``backward`` returns the gradients for adjoints of the final electron
    momenta and of the final real-space Ex planes."""

    def __init__(self, grid_cfg, laser_cfg, species_cfgs, plasma, knot_x, comm, ckpt=10,
                 graph=True, gather_skip="index"):
        """gather_skip: "index" (default) keeps the reference's gather, which
        skips particles whose index is >= Np_stay (grid_deposit_m1.cl:367-368:
        whether a particle is pushed then depends on how many others have
        left the valid cells, a discontinuity of the objective); "cell" (a
        deviation from the reference, for gradients) pushes every particle
        of a valid cell."""
        import chimeracl_amd as CA
        self.comm = comm
        self.ckpt = int(ckpt)
        self.laser = dict(laser_cfg)
        self.knot_x = np.asarray(knot_x, dtype=np.float64)
        self.solver = CA.Solver(dict(grid_cfg), comm)
        eo, io = species_cfgs
        self.eons, self.ions = CA.Particles(dict(eo), comm), CA.Particles(dict(io), comm)
        from .config import ParticleSource
        self.eons.add_particles(source=ParticleSource(plasma))
        self.ions.add_particles(source=ParticleSource(plasma))
        if gather_skip not in ("index", "cell"):
            raise ValueError(f"gather_skip must be 'index' or 'cell', not {gather_skip!r}")
        self.gather_skip = self.eons.gather_skip = gather_skip
        # the adjoint gather's activity test compares the index with this
        # value: the recorded Np_stay, or (rule "cell") the largest uint32
        self._np_all = torch.full((1,), -1, dtype=torch.int32, device=comm.device)
        dev, real = comm.device, comm.real
        self.base = {k: self.eons.DataDev[k].clone() for k in self.eons._arg_names()}
        self.base_ions = {k: self.ions.DataDev[k].clone() for k in self.ions._arg_names()}
        self.loop = CA.PIC_loop(solvers=[self.solver], species=[self.eons, self.ions], graph=graph)
        self.adj_sp = AdjointSpectral(self.solver)
        self._twin = None
        a = self.solver.Args
        self.M, self.Nx, self.Nr = a["M"], a["Nx"], a["Nr"]
        self.Np = self.eons.Args["Np"]
        self.Np_i = self.ions.Args["Np"]
        self._alloc_adjoint()
        self.timings = {}

    # ------------------------------------------------------------------ buffers
    def _spec(self):
        return torch.zeros((self.Nr - 1, self.Nx), dtype=self.comm.cplx, device=self.comm.device)

    def _real_planes(self):
        dev, shp = self.comm.device, (self.Nr, self.Nx)
        return [torch.zeros(shp, dtype=self.comm.real, device=dev)] + \
               [torch.zeros(shp, dtype=self.comm.cplx, device=dev) for _ in range(self.M)]

    def _alloc_adjoint(self):
        M = self.M
        spec = lambda: [[self._spec() for _ in range(M + 1)] for _ in range(3)]
        self.Eb, self.Gb, self.carry = spec(), spec(), spec()
        self.Et, self.Bb, self.Jb, self.N1b = spec(), spec(), spec(), spec()
        self.rhob = [[self._spec() for _ in range(M + 1)]]
        self.Er_b = [self._real_planes() for _ in range(3)]
        self.Br_b = [self._real_planes() for _ in range(3)]
        self.Jr_b = [self._real_planes() for _ in range(3)]
        self.rhor_b = [self._real_planes()]
        self.rho_ion_b = [self._real_planes()]
        f = lambda: torch.zeros(self.Np, dtype=self.comm.real, device=self.comm.device)
        self.pb = {k: f() for k in ("x", "y", "z", "px", "py", "pz", "g_inv", "w")}
        self.fb = [f() for _ in range(6)]
        self.tmp = {k: f() for k in ("x1", "y1", "z1", "x2", "y2", "z2")}
        self.wb_ion = torch.zeros(self.Np_i, dtype=self.comm.real, device=self.comm.device)
        self.prof_ws = torch.empty(_lib.lib.chm_profile_adj_workspace_bytes(len(self.knot_x)),
                                   dtype=torch.uint8, device=self.comm.device)
        # structs
        self.s_Eb, self.s_Gb, self.s_carry = _fb_arr(self.Eb), _fb_arr(self.Gb), _fb_arr(self.carry)
        self.s_Et, self.s_Bb = _fb_arr(self.Et), _fb_arr(self.Bb)
        self.s_Jb, self.s_N1b, self.s_rhob = _fb_arr(self.Jb), _fb_arr(self.N1b), _fb_arr(self.rhob)
        self.s_Er_b, self.s_Br_b = _planes_arr(self.Er_b), _planes_arr(self.Br_b)
        self.s_Jr_b, self.s_rhor_b = _planes_arr(self.Jr_b), _planes_arr(self.rhor_b)
        self.s_rho_ion_b = _planes_arr(self.rho_ion_b)

    def _state_tensors(self):
        d, D, M = self.eons.DataDev, self.solver.DataDev, self.M
        parts = [d[k] for k in ("x", "y", "z", "px", "py", "pz", "g_inv")]
        flds = [D[f"{f}{c}_fb_m{m}"] for f in ("E", "G", "dN1") for c in _COMPS for m in range(M + 1)]
        return parts + flds

    def _save_state(self):
        return [t.clone() for t in self._state_tensors()]

    def _load_state(self, saved):
        for t, s in zip(self._state_tensors(), saved):
            t.copy_(s)

    # ------------------------------------------------------------------ set-up
    def _twin_solver(self):
        if self._twin is None:
            from .solver import Solver
            a = self.solver.Args
            cfg = {k: a[k] for k in ("Xmin", "Xmax", "Nx", "Nr", "M", "dt")}
            cfg["Rmax"] = a["dr"] * (a["Nr"] - 1.5)
            cfg["DampCells"] = a["DampCells"]
            self._twin = Solver(cfg, Communicator(device=self.comm.device.index, precision="f64"))
        return self._twin

    def reset(self, a0, R, knot_vals):
        """Set-up of a run: zero fields, laser (laser.py:3-37) and the plasma
        with weights w_base * f(x) (dens_profile)."""
        from .laser import add_gausian_pulse
        self.params = (float(a0), float(R), np.asarray(knot_vals, dtype=np.float64).copy())
        D = self.solver.DataDev
        for n in [f + c for f in ("E", "B", "J", "G") for c in _COMPS] + ["rho"]:
            for m in range(self.M + 1):
                D[f"{n}_m{m}"].zero_()
        for n in ["rho"] + [f + c for f in self.solver.FB_FIELDS for c in _COMPS]:
            for m in range(self.M + 1):
                D[f"{n}_fb_m{m}"].zero_()
        las = dict(self.laser, a0=float(a0), R=float(R))
        add_gausian_pulse(self.solver, las)
        for k, t in self.base.items():
            self.eons.DataDev[k].copy_(t)
        for k, t in self.base_ions.items():
            self.ions.DataDev[k].copy_(t)
        self.eons.dens_profile(self.knot_x, self.params[2])
        self.ions.dens_profile(self.knot_x, self.params[2])
        if "cell_offset" in self.ions.DataDev:
            # the captured step adds the ions' staged per-cell rho sums:
            # restage them for the new weights (the ions stay sorted)
            self.solver.depose_charge(species=[self.ions])
            D["rho_m0"].zero_()
        self.loop.it = 0

    # ------------------------------------------------------------------ forward
    def forward(self, a0, R, knot_vals, nsteps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        self.reset(a0, R, knot_vals)
        self.nsteps = nsteps
        self.ckpts = []
        for it in range(nsteps):
            if it % self.ckpt == 0:
                self.ckpts.append(self._save_state())
            self.loop.step()
        torch.cuda.synchronize()
        self.timings["forward"] = time.perf_counter() - t0
        d, D = self.eons.DataDev, self.solver.DataDev
        return ([d[k].clone() for k in ("px", "py", "pz")] +
                [D[f"Ex_m{m}"].clone() for m in range(self.M + 1)])

    def _run_step_recording(self):
        d, D, M = self.eons.DataDev, self.solver.DataDev, self.M
        start = [d[k].clone() for k in ("x", "y", "z", "px", "py", "pz", "g_inv")]
        self.loop.step()
        rec = {"start": start,
               "E": [[D[f"E{c}_m{m}"].clone() for m in range(M + 1)] for c in _COMPS],
               "B": [[D[f"B{c}_m{m}"].clone() for m in range(M + 1)] for c in _COMPS],
               "sort_indx": d["sort_indx"].clone(), "cell_offset": d["cell_offset"].clone(),
               "np_stay": d["Np_stay"].clone()}
        return rec

    # ------------------------------------------------------------------ backward
    def _call(self, name, *args):
        _lib.call(name + self.comm.sfx, *args, self.comm.stream())

    def backward(self, grad_u, grad_Ex):
        """grad_u: adjoints of the final (px, py, pz) (None = 0); grad_Ex:
        adjoints of the final Ex_m planes (None = 0).  Returns (dL/da0,
        dL/dR, dL/dknots) as float, float, float64 ndarray."""
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        pb = self.pb
        for t in list(pb.values()) + [self.wb_ion]:
            t.zero_()
        for grp in (self.Eb, self.Gb, self.carry):
            for l in grp:
                for t in l:
                    t.zero_()
        for t in self.rho_ion_b[0]:
            t.zero_()
        if grad_u is not None:
            for k, g in zip(("px", "py", "pz"), grad_u):
                if g is not None:
                    pb[k].copy_(g)
        seed_E = grad_Ex
        nseg = len(self.ckpts)
        for seg in reversed(range(nseg)):
            self._load_state(self.ckpts[seg])
            first = seg * self.ckpt
            last = min(self.nsteps, first + self.ckpt)
            recs = []
            for it in range(first, last):
                self.loop.it = it
                recs.append(self._run_step_recording())
            for k in reversed(range(len(recs))):
                self._adjoint_step(recs[k], seed_E if (first + k == self.nsteps - 1) else None)
                recs[k] = None
        grads = self._param_grads()
        torch.cuda.synchronize()
        self.timings["backward"] = time.perf_counter() - t0
        return grads

    def _adjoint_step(self, rec, seed_E):
        s, e, a, M = self.solver, self.eons, self.solver.Args, self.M
        d, pb, tmp = e.DataDev, self.pb, self.tmp
        grid = s._grid()
        x0, y0, z0, ux, uy, uz, g0 = rec["start"]
        Np, dt2 = self.Np, e.Args["dt_2"]
        # positions after the first and the second half push
        x1, y1, z1 = tmp["x1"], tmp["y1"], tmp["z1"]
        x2, y2, z2 = tmp["x2"], tmp["y2"], tmp["z2"]
        x1.copy_(x0); y1.copy_(y0); z1.copy_(z0)
        self._call("chm_push_x", ptr(x1), ptr(y1), ptr(z1), ptr(ux), ptr(uy), ptr(uz), ptr(g0), dt2, Np)
        x2.copy_(x1); y2.copy_(y1); z2.copy_(z1)
        self._call("chm_push_x", ptr(x2), ptr(y2), ptr(z2), ptr(ux), ptr(uy), ptr(uz), ptr(g0), dt2, Np)
        E_r, B_r = _planes_arr(rec["E"]), _planes_arr(rec["B"])
        # gather + push
        fb_arr = (ctypes.c_void_p * 6)(*[ptr(t) for t in self.fb])
        fbp = ctypes.cast(fb_arr, ctypes.c_void_p)
        self._call("chm_gather_push_adj", ptr(x2), ptr(y2), ptr(z2), ptr(ux), ptr(uy), ptr(uz), Np,
                   ptr(self._np_all if self.gather_skip == "cell" else rec["np_stay"]),
                   e.Args["FactorPush"], grid, E_r, B_r, ptr(pb["px"]),
                   ptr(pb["py"]), ptr(pb["pz"]), ptr(pb["g_inv"]), ptr(pb["x"]), ptr(pb["y"]),
                   ptr(pb["z"]), fbp)
        for grp in (self.Er_b, self.Br_b):
            for l in grp:
                for t in l:
                    t.zero_()
        ws = s._dep_ws()
        for k, out in ((0, self.Er_b), (3, self.Br_b)):
            self._call("chm_gather_adj_scatter", ptr(rec["sort_indx"]), ptr(rec["cell_offset"]),
                       ptr(x2), ptr(y2), ptr(z2), ptr(self.fb[k]), ptr(self.fb[k + 1]),
                       ptr(self.fb[k + 2]), grid, _planes_arr(out), ptr(ws), ws.numel())
        if seed_E is not None:
            for m, g in enumerate(seed_E):
                if g is not None:
                    self.Er_b[0][m].add_(g)
        self._call("chm_warp_axis_adj", self.s_Er_b, 3, a["Nx"], a["Nr"], M)
        self._call("chm_warp_axis_adj", self.s_Br_b, 3, a["Nx"], a["Nr"], M)
        # T_b^H: forward transform with the adjoint tables
        asp, tws = self.adj_sp.sp, s._tws
        self._call("chm_transform", asp, self.s_Er_b, self.s_Et, 3, 0, 0, 0, ptr(tws), tws.numel())
        self._call("chm_transform", asp, self.s_Br_b, self.s_Bb, 3, 0, 0, 0, ptr(tws), tws.numel())
        # B^ = rot(G)/w^2
        self._call("chm_field_rot_adj", asp, self.s_Bb, self.s_Gb, 1, 1, ptr(tws), tws.numel())
        # PSATD + damping
        self._call("chm_damp_fields_adj", s._sp, self.s_Eb, self.s_Gb, self.s_Et, self.s_Jb,
                   self.s_carry, self.s_N1b)
        # dN1 = grad(rho^)
        self._call("chm_field_grad_adj", asp, self.s_N1b, self.s_rhob, 0, ptr(tws), tws.numel())
        # smoothing, then T_f^H (backward transform with the adjoint tables)
        self._call("chm_mult_table", s._sp, self.s_rhob, 1, 0)
        self._call("chm_mult_table", s._sp, self.s_Jb, 3, 0)
        self._call("chm_transform", asp, self.s_rhor_b, self.s_rhob, 1, 1, 0, 0, ptr(tws), tws.numel())
        self._call("chm_transform", asp, self.s_Jr_b, self.s_Jb, 3, 1, 0, 0, ptr(tws), tws.numel())
        dvi = s.DataDev["dV_inv"]
        self._call("chm_postproc_deposit_adj", self.s_rhor_b, 1, ptr(dvi), a["Nx"], a["Nr"], M)
        self._call("chm_postproc_deposit_adj", self.s_Jr_b, 3, ptr(dvi), a["Nx"], a["Nr"], M)
        # rho deposit at x2 (electrons); ions: summed over the steps
        for mm, t in enumerate(self.rho_ion_b[0]):
            t.add_(self.rhor_b[0][mm])
        self._call("chm_deposit_scalar_adj", ptr(x2), ptr(y2), ptr(z2), ptr(d["w"]),
                   int(e.Args["charge"]), Np, grid, self.s_rhor_b, ptr(pb["x"]), ptr(pb["y"]),
                   ptr(pb["z"]), ptr(pb["w"]))
        # second half push
        self._call("chm_push_x_adj", ptr(pb["x"]), ptr(pb["y"]), ptr(pb["z"]), ptr(ux), ptr(uy),
                   ptr(uz), ptr(g0), ptr(pb["px"]), ptr(pb["py"]), ptr(pb["pz"]), ptr(pb["g_inv"]),
                   dt2, Np)
        # J deposit at x1
        self._call("chm_deposit_vector_adj", ptr(x1), ptr(y1), ptr(z1), ptr(ux), ptr(uy), ptr(uz),
                   ptr(g0), ptr(d["w"]), int(e.Args["charge"]), Np, grid, self.s_Jr_b, ptr(pb["x"]),
                   ptr(pb["y"]), ptr(pb["z"]), ptr(pb["px"]), ptr(pb["py"]), ptr(pb["pz"]),
                   ptr(pb["g_inv"]), ptr(pb["w"]))
        # first half push
        self._call("chm_push_x_adj", ptr(pb["x"]), ptr(pb["y"]), ptr(pb["z"]), ptr(ux), ptr(uy),
                   ptr(uz), ptr(g0), ptr(pb["px"]), ptr(pb["py"]), ptr(pb["pz"]), ptr(pb["g_inv"]),
                   dt2, Np)

    def _param_grads(self):
        s, a, M = self.solver, self.solver.Args, self.M
        grid = s._grid()
        di = self.ions.DataDev
        # ions: rho deposit adjoint with the summed plane adjoints
        self._call("chm_deposit_scalar_adj", ptr(di["x"]), ptr(di["y"]), ptr(di["z"]), ptr(di["w"]),
                   int(self.ions.Args["charge"]), self.Np_i, grid, self.s_rho_ion_b,
                   ptr(self._scratch_ion("x")), ptr(self._scratch_ion("y")),
                   ptr(self._scratch_ion("z")), ptr(self.wb_ion))
        # knots
        xx = torch.from_numpy(self.knot_x).to(self.comm.device)
        dxm1 = 1.0 / (xx[1:] - xx[:-1])
        kg = torch.zeros(len(self.knot_x), dtype=torch.float64, device=self.comm.device)
        for parts, base, wb in ((self.eons, self.base, self.pb["w"]), (self.ions, self.base_ions, self.wb_ion)):
            self._call("chm_dens_profile_adj", ptr(base["x"]), ptr(base["w"]), ptr(wb), parts.Args["Np"],
                       ptr(xx), ptr(dxm1), len(self.knot_x), ptr(kg), ptr(self.prof_ws),
                       self.prof_ws.numel())
        # laser: tangents of the step-0 spectral Ez, Gz (mode 0)
        a0, R, _ = self.params
        dE_a, dG_a = self._laser_tangent(a0, R, "a0")
        dE_R, dG_R = self._laser_tangent(a0, R, "R")
        lam_E = self.Eb[2][0].to(torch.complex128)
        lam_G = self.Gb[2][0].to(torch.complex128)
        ip = lambda lam, v: float((lam.real * v.real + lam.imag * v.imag).sum())
        g_a0 = ip(lam_E, dE_a) + ip(lam_G, dG_a)
        g_R = ip(lam_E, dE_R) + ip(lam_G, dG_R)
        return g_a0, g_R, kg.cpu().numpy()

    def _scratch_ion(self, k):
        key = "ion_" + k
        t = self.tmp.get(key)
        if t is None:
            t = self.tmp[key] = torch.zeros(self.Np_i, dtype=self.comm.real, device=self.comm.device)
        return t

    def _laser_tangent(self, a0, R, which):
        """d(Ez^, Gz^ mode 0 after the laser set-up)/d(param), in FP64: the
        set-up is linear in the real-space pulse, so the tangent is the
        set-up applied to the pulse's derivative."""
        tw = self._twin_solver()
        a, D = tw.Args, tw.DataDev
        las = self.laser
        k0 = 2 * math.pi * las["k0"]
        st = tw.comm.stream()
        Ez = D["Ez_m0"]
        Ez.zero_()
        if which == "a0":
            _lib.call("chm_laser_pulse_f64", ptr(Ez), a["Nx"], a["Nr"], a["Xmin"], a["dx"], a["Rmin"],
                      a["dr"], k0, 1.0, las["x0"], las["Lx"], R, st)
        else:
            _lib.call("chm_laser_pulse_dR_f64", ptr(Ez), a["Nx"], a["Nr"], a["Xmin"], a["dx"],
                      a["Rmin"], a["dr"], k0, a0, las["x0"], las["Lx"], R, st)
        tw.fb_transform(scals=["Ez"], dir=0)
        D["Gz_fb_m0"].zero_()
        _lib.call("chm_laser_spectral_f64", tw._sp, ptr(D["w_m0"]), ptr(D["Ez_fb_m0"]), ptr(D["Gz_fb_m0"]),
                  las["x0"] - las["x_foc"], st)
        return D["Ez_fb_m0"].clone(), D["Gz_fb_m0"].clone()


# ---------------------------------------------------------------------- autograd
class PICFunction(torch.autograd.Function):
    """(a0, R, knots) -> (px, py, pz of the electrons, Ex_m planes) after
    ``sim.nsteps`` steps of a DiffPIC run.  Gradients come from
    DiffPIC.backward (hand-written adjoint kernels, checkpointing)."""

    @staticmethod
    def forward(ctx, a0, R, knots, sim, nsteps):
        outs = sim.forward(float(a0), float(R), knots.detach().cpu().numpy(), nsteps)
        ctx.sim = sim
        ctx.dtype = knots.dtype
        ctx.n_u = 3
        return tuple(outs)

    @staticmethod
    def backward(ctx, *grads):
        sim = ctx.sim
        gu = [g if g is not None else None for g in grads[:3]]
        gE = [g for g in grads[3:]]
        if all(g is None for g in gu):
            gu = None
        if all(g is None for g in gE):
            gE = None
        g_a0, g_R, g_k = sim.backward(gu, gE)
        return (torch.tensor(g_a0, dtype=torch.float64), torch.tensor(g_R, dtype=torch.float64),
                torch.from_numpy(g_k).to(ctx.dtype), None, None)


def run(sim, a0, R, knots, nsteps):
    """Differentiable run: a0, R, knots are CPU float64 tensors (requires_grad
    as wanted); returns (px, py, pz, Ex_m0, Ex_m1, ...) on the GPU."""
    return PICFunction.apply(a0, R, knots, sim, nsteps)


# ---------------------------------------------------------------------- objectives
def objective_gamma2(px, py, pz, threshold=5.0, tau=0.25):
    """Smooth sum of gamma^2 over electrons with gamma > threshold:
    sum gamma^2 sigmoid((gamma - threshold)/tau), gamma = sqrt(1 + u^2)."""
    u2 = px.double() ** 2 + py.double() ** 2 + pz.double() ** 2
    gam = torch.sqrt(1.0 + u2)
    return (gam ** 2 * torch.sigmoid((gam - threshold) / tau)).sum()


def objective_wake_energy(Ex_planes, rgrid, ncols=256):
    """Longitudinal wake-field energy in the last ``ncols`` x-columns (the
    highest x indices), rows 1..Nr-1: sum r (Ex_m0^2 + 2 sum_m>=1 |Ex_m|^2)."""
    r = torch.as_tensor(rgrid[1:], dtype=torch.float64, device=Ex_planes[0].device)[:, None]
    tot = (r * Ex_planes[0][1:, -ncols:].double() ** 2).sum()
    for Em in Ex_planes[1:]:
        v = Em[1:, -ncols:]
        tot = tot + 2.0 * (r * (v.real.double() ** 2 + v.imag.double() ** 2)).sum()
    return tot


def from_config(name, precision="f64", dht="bf16x3", max_particles=None, ckpt=None,
                gather_skip="index", **over):
    """DiffPIC for one of the workloads of chimeracl_amd.config (C4: Nz=1024,
    Nr=128, 2x2x4 ppc, 32 knots, checkpoints every 10 steps; C5: Nz=4096,
    Nr=256, 2x2x8 ppc, 64 knots, checkpoints every 20 steps).  Returns
    (sim, a0, R, knot values, number of steps)."""
    from . import config as C
    cfg = dict(C.CONFIGS[name], **over)
    g = C.grid_config(name, **over)
    las = C.laser_config(name, **over)
    comm = Communicator(precision=precision, dht=dht)
    dx = (g["Xmax"] - g["Xmin"]) / (g["Nx"] - 1)
    dr = g["Rmax"] / (g["Nr"] - 1.5)
    eo, io = C.species_config(name, {"dx": dx, "dr": dr, "dt": g["dt"]}, **over)
    plasma = C.load_plasma(g, cfg["ppc"], cfg["dens"], seed=0, max_particles=max_particles,
                           order="cell")
    kx, kv = C.density_profile(name, g, **over)
    sim = DiffPIC(g, las, (eo, io), plasma, kx, comm, ckpt=ckpt or cfg.get("ckpt", 10),
                  gather_skip=gather_skip)
    return sim, las["a0"], las["R"], np.asarray(kv, dtype=np.float64), cfg["steps"]
