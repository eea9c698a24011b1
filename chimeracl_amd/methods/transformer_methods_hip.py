"""Fourier-Bessel transforms and spectral operators (replaces chimeraCL/
methods/transformer_methods_cl.py and the table construction of
transformer.py / solver.py, which here runs in FP64 on the GPU)."""
from __future__ import annotations

import torch

from .. import _lib
from .generic_methods_hip import GenericMethodsHIP, ptr


class TransformerMethodsHIP(GenericMethodsHIP):
    FB_FIELDS = ["E", "B", "G", "J", "dN0", "dN1"]

    def init_transformer_methods(self):
        self._fcache = {}

    # ---------------------------------------------------------------- data
    def _init_transformer_data_on_dev(self):
        """transformer.py:100-132 (spectral planes [Nr-1, Nx] per mode)."""
        a = self.Args
        names = ["rho"] + [f + c for f in self.FB_FIELDS for c in a["vec_comps"]]
        for n in names:
            for m in range(a["M"] + 1):
                self.DataDev[f"{n}_fb_m{m}"] = torch.zeros((a["Nr"] - 1, a["Nx"]), dtype=self.comm.cplx,
                                                           device=self.comm.device)

    def _build_tables(self):
        """_make_spectral_axes + _make_DHT (transformer.py:28-97),
        _make_ms_coefficients (solver.py:41-52), _init_field_damping
        (solver_methods_cl.py:32-41): all built in FP64 by the GPU."""
        a, dev, D = self.Args, self.comm.device, self.DataDev
        Nx, Nr, M = a["Nx"], a["Nr"], a["M"]
        Nk = Nr - 1
        f64, c128 = torch.float64, torch.complex128
        nd = int(a["DampCells"]) if "DampCells" in a else 0
        t = {"kx": torch.empty(Nx, dtype=f64, device=dev),
             "kr": torch.empty((M + 2, Nk), dtype=f64, device=dev),
             "phase_minus": torch.empty(Nx, dtype=c128, device=dev),
             "phase_plus": torch.empty(Nx, dtype=c128, device=dev),
             "fft_twiddle": torch.empty(Nx, dtype=c128, device=dev),
             "DampProfile": torch.empty(max(2 * nd, 1), dtype=f64, device=dev)}
        out = _lib.TablesOut()
        out.kx, out.kr = ptr(t["kx"]), ptr(t["kr"])
        out.phase_minus, out.phase_plus = ptr(t["phase_minus"]), ptr(t["phase_plus"])
        out.twiddle = ptr(t["fft_twiddle"])
        out.damp = ptr(t["DampProfile"]) if nd > 0 else None
        per_m = {"w": "w_m", "c1": "MxSlv_cos(wdt)_m", "c2": "MxSlv_sin(wdt)*w_m",
                 "c3": "MxSlv_1/w**2_m", "smooth": "SmoothingFilter_m"}
        per_k = {"DHT": "DHT_m", "DHT_inv": "DHT_inv_m", "dDp": "dDHT_plus_m", "dDm": "dDHT_minus_m"}
        for m in range(M + 1):
            for fld, key in per_m.items():
                t[key + str(m)] = torch.empty((Nk, Nx), dtype=f64, device=dev)
                getattr(out, fld)[m] = ptr(t[key + str(m)])
            for fld, key in per_k.items():
                t[key + str(m)] = torch.empty((Nk, Nk), dtype=f64, device=dev)
                getattr(out, fld)[m] = ptr(t[key + str(m)])
        ws = torch.empty(_lib.lib.chm_tables_workspace_bytes(Nx, Nr, M), dtype=torch.uint8, device=dev)
        _lib.call("chm_build_tables_f64", Nx, Nr, M, a["dx"], a["dr"], a["dt"], a["Xmin"], nd, out,
                  ptr(ws), ws.numel(), self.comm.stream())
        for m in range(M + 2):
            t["kr_m" + str(m)] = t["kr"][m]
        for m in range(M + 1):
            t["Poiss_m" + str(m)] = t["MxSlv_1/w**2_m" + str(m)]
        D.update(t)
        # host copies the reference keeps in Args (laser.py reads kx, w_m0)
        a["kx"] = t["kx"].cpu().numpy()
        for m in range(M + 2):
            a["kr_m" + str(m)] = t["kr"][m].cpu().numpy()
        for m in range(M + 1):
            a["w_m" + str(m)] = t["w_m" + str(m)].cpu().numpy()
        if nd > 0:
            a["DampProfile"] = t["DampProfile"].cpu().numpy()
        self._make_spectral_struct(nd)

    def _make_spectral_struct(self, nd):
        a, D, dev = self.Args, self.DataDev, self.comm.device
        M = a["M"]
        sp = _lib.Spectral(Nx=a["Nx"], Nr=a["Nr"], M=M, ndamp=2 * nd, dt_inv=a["dt_inv"])
        Nk = a["Nr"] - 1
        nel = _lib.lib.chm_dht_table_elems(Nk, Nk)
        if not self.comm.mixed:
            src = dict(D)
        else:
            # FP32 copies of every table; DHT-type matrices also split in BF16
            src = {}
            sp.dht_kind = self.comm.dht_kind
            for key in ["kx", "DampProfile"] + [k + str(m) for m in range(M + 1) for k in
                                                 ("MxSlv_cos(wdt)_m", "MxSlv_sin(wdt)*w_m",
                                                  "MxSlv_1/w**2_m", "SmoothingFilter_m")]:
                src[key] = D[key].float()
            for key in ("phase_minus", "phase_plus", "fft_twiddle"):
                src[key] = D[key].to(torch.complex64)
        # DHT-type matrices in the zero-padded layout the MFMA products read
        st = self.comm.stream()
        for m in range(M + 1):
            for key in ("DHT_m", "DHT_inv_m", "dDHT_plus_m", "dDHT_minus_m"):
                T = D[key + str(m)]
                if not self.comm.mixed:
                    pad = torch.empty(nel, dtype=torch.float64, device=dev)
                    _lib.call("chm_pad_table", ptr(T), Nk, Nk, ptr(pad), None, None, None, st)
                    src[key + str(m)] = pad
                elif self.comm.dht_kind == _lib.CHM_DHT_BF16X3:
                    hi = torch.empty(nel, dtype=torch.int16, device=dev)
                    lo = torch.empty(nel, dtype=torch.int16, device=dev)
                    _lib.call("chm_pad_table", ptr(T), Nk, Nk, None, None, ptr(hi), ptr(lo), st)
                    src[key + str(m)], src[key + str(m) + "_lo"] = hi, lo
                else:
                    pad = torch.empty(nel, dtype=torch.float32, device=dev)
                    _lib.call("chm_pad_table", ptr(T), Nk, Nk, None, ptr(pad), None, None, st)
                    src[key + str(m)] = pad
        self._kernel_tables = src
        sp.kx = ptr(src["kx"])
        sp.phase_minus, sp.phase_plus = ptr(src["phase_minus"]), ptr(src["phase_plus"])
        sp.damp = ptr(src["DampProfile"]) if nd > 0 else None
        sp.twiddle = ptr(src["fft_twiddle"])
        for m in range(M + 1):
            sp.DHT[m], sp.DHT_inv[m] = ptr(src[f"DHT_m{m}"]), ptr(src[f"DHT_inv_m{m}"])
            sp.dDp[m], sp.dDm[m] = ptr(src[f"dDHT_plus_m{m}"]), ptr(src[f"dDHT_minus_m{m}"])
            if f"DHT_m{m}_lo" in src:
                sp.DHT_lo[m], sp.DHT_inv_lo[m] = ptr(src[f"DHT_m{m}_lo"]), ptr(src[f"DHT_inv_m{m}_lo"])
                sp.dDp_lo[m], sp.dDm_lo[m] = ptr(src[f"dDHT_plus_m{m}_lo"]), ptr(src[f"dDHT_minus_m{m}_lo"])
            sp.c1[m], sp.c2[m] = ptr(src[f"MxSlv_cos(wdt)_m{m}"]), ptr(src[f"MxSlv_sin(wdt)*w_m{m}"])
            sp.c3[m], sp.smooth[m] = ptr(src[f"MxSlv_1/w**2_m{m}"]), ptr(src[f"SmoothingFilter_m{m}"])
            if self.comm.mixed:
                # the FP64 PSATD coefficients (the update of the FP32 fields
                # runs in FP64 arithmetic from them: no systematic rotation
                # error from rounding the tables)
                sp.c1d[m] = ptr(D[f"MxSlv_cos(wdt)_m{m}"])
                sp.c2d[m] = ptr(D[f"MxSlv_sin(wdt)*w_m{m}"])
                sp.c3d[m] = ptr(D[f"MxSlv_1/w**2_m{m}"])
        self._sp = sp
        Nx, Nr, es = a["Nx"], a["Nr"], 4 if self.comm.mixed else 8
        nws = max(_lib.lib.chm_transform_workspace_bytes(Nx, Nr, M, 10, es),
                  _lib.lib.chm_spectral_workspace_bytes(Nx, Nr, M, es))
        self._tws = torch.empty(nws, dtype=torch.uint8, device=dev)

    def _refresh_phases(self):
        """get_phase_minus/plus (transformer_generic.cl:28-55) for the current
        Xmin; the reference recomputes them in every transform call."""
        a, D = self.Args, self.DataDev
        _lib.call("chm_phase_tables_f64", ptr(D["kx"]), a["Nx"], float(a["Xmin"]), ptr(D["phase_minus"]),
                  ptr(D["phase_plus"]), self.comm.stream())
        if self.comm.mixed:
            for k in ("phase_minus", "phase_plus"):
                self._kernel_tables[k].copy_(D[k].to(torch.complex64))

    # ---------------------------------------------------------------- structs
    def _fb(self, name):
        st = self._fcache.get(name)
        if st is None:
            st = _lib.Fb()
            for m in range(self.Args["M"] + 1):
                st.m[m] = ptr(self.DataDev[f"{name}_fb_m{m}"])
            self._fcache[name] = st
        return st

    def _fb_arr(self, names):
        key = ("arr",) + tuple(names)
        arr = self._fcache.get(key)
        if arr is None:
            arr = (_lib.Fb * len(names))(*[self._fb(n) for n in names])
            self._fcache[key] = arr
        return arr

    def _vec(self, v):
        return [v + c for c in self.Args["vec_comps"]]

    # ---------------------------------------------------------------- transforms
    def _transform(self, names, dir, mode, smooth=False):
        self.call("chm_transform", self._sp, self._planes_arr(names), self._fb_arr(names),
                  len(names), dir, 0 if mode == "full" else 1, 1 if smooth else 0,
                  ptr(self._tws), self._tws.numel())

    def transform_field(self, arg_cmp, dir, mode="full"):
        """transformer_methods_cl.py:38-64."""
        self._transform([arg_cmp], dir, mode)

    def fb_transform(self, scals=[], vects=[], dir=0, mode="full"):
        """transformer.py:17-26, all fields in one batched call."""
        names = list(scals) + [n for v in vects for n in self._vec(v)]
        self._transform(names, dir, mode)

    def fb_transform_smooth(self, scals=[], vects=[]):
        """fb_transform(dir=0) followed by fields_smooth of the same fields,
        with the filter applied in the transform's last pass."""
        names = list(scals) + [n for v in vects for n in self._vec(v)]
        self._transform(names, 0, "full", smooth=True)

    # ---------------------------------------------------------------- operators
    def fields_smooth(self, flds):
        """transformer_methods_cl.py:79-85."""
        self.call("chm_mult_table", self._sp, self._fb_arr(flds), len(flds), 0)

    def field_poiss_vec(self, fld):
        """transformer_methods_cl.py:66-71."""
        names = self._vec(fld)
        self.call("chm_mult_table", self._sp, self._fb_arr(names), len(names), 1)

    def field_poiss_scl(self, fld):
        self.call("chm_mult_table", self._sp, self._fb_arr([fld]), 1, 1)

    def field_grad(self, scl_in, vec_out, copy_to=None):
        """transformer_methods_cl.py:87-133.  copy_to: spectral vector that
        first receives the old vec_out (the dN0 <- dN1 copy of the loop)."""
        prev = self._fb_arr(self._vec(copy_to)) if copy_to else None
        self.call("chm_field_grad", self._sp, self._fb_arr([scl_in]), self._fb_arr(self._vec(vec_out)),
                  prev, ptr(self._tws), self._tws.numel())

    def field_div(self, vec_in, scl_out):
        """transformer_methods_cl.py:135-183."""
        self.call("chm_field_div", self._sp, self._fb_arr(self._vec(vec_in)), self._fb_arr([scl_out]),
                  ptr(self._tws), self._tws.numel())

    def field_rot(self, fld_in, fld_out, poisson=False):
        """transformer_methods_cl.py:185-263 (poisson: also *1/w^2)."""
        self.call("chm_field_rot", self._sp, self._fb_arr(self._vec(fld_in)),
                  self._fb_arr(self._vec(fld_out)), 1 if poisson else 0, ptr(self._tws),
                  self._tws.numel())
