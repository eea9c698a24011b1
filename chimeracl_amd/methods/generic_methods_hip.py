"""Device context and generic helpers (replaces chimeraCL/methods/
generic_methods_cl.py: GenericMethodsCL and Communicator).

Device memory is PyTorch-ROCm tensors; every computation is a call into
libchimera_hip.so on the current HIP stream.
"""
from __future__ import annotations

import numpy as np
import torch

from .. import _lib

_DTYPES = {
    np.dtype(np.float64): torch.float64, np.dtype(np.float32): torch.float32,
    np.dtype(np.complex128): torch.complex128, np.dtype(np.complex64): torch.complex64,
    np.dtype(np.uint32): torch.int32, np.dtype(np.int32): torch.int32,
    np.dtype(np.int64): torch.int64, np.dtype(np.bool_): torch.bool,
}


class Communicator:
    """One GPU and the precision of the step (generic_methods_cl.py:152-195).

    precision: 'f64' (plain form) or 'mixed' (FP32 particles, fields and FFT;
    DHT products as split-BF16 3-term ('bf16x3') or FP32 ('f32') MFMA).
    """

    def __init__(self, device=0, precision="f64", dht="bf16x3", **ctx_kw_args):
        if not torch.cuda.is_available():
            raise RuntimeError("chimeracl_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
        if precision not in ("f64", "mixed"):
            raise ValueError(f"precision must be 'f64' or 'mixed', not {precision!r}")
        self.device = torch.device("cuda", device)
        self.precision = precision
        self.mixed = precision == "mixed"
        self.real = torch.float32 if self.mixed else torch.float64
        self.cplx = torch.complex64 if self.mixed else torch.complex128
        self.np_real = np.float32 if self.mixed else np.float64
        self.sfx = "_f32" if self.mixed else "_f64"
        self.dht = dht
        self.dht_kind = _lib.CHM_DHT_BF16X3 if dht == "bf16x3" else _lib.CHM_DHT_F32
        self.dev_type = "GPU"
        self.dev_name = torch.cuda.get_device_name(self.device)
        self.plat_name = "ROCm"

    def stream(self):
        return torch.cuda.current_stream(self.device).cuda_stream

    def synchronize(self):
        torch.cuda.synchronize(self.device)


_default = {}


def default_comm(precision="f64"):
    if precision not in _default:
        _default[precision] = Communicator(precision=precision)
    return _default[precision]


class LazyArgs(dict):
    """Args dict whose entries may be resolved from the device on first read
    (Np_stay stays on the GPU during a step; reading it synchronises)."""

    def __init__(self, *a, **k):
        super().__init__(*a, **k)
        self._lazy = {}

    def set_lazy(self, key, fn):
        dict.__setitem__(self, key, None)
        self._lazy[key] = fn

    def __getitem__(self, key):
        fn = self._lazy.pop(key, None)
        if fn is not None:
            dict.__setitem__(self, key, fn())
        return dict.__getitem__(self, key)

    def get(self, key, default=None):
        return self[key] if key in self else default

    def __setitem__(self, key, val):
        self._lazy.pop(key, None)
        dict.__setitem__(self, key, val)


def ptr(t):
    return t.data_ptr() if t is not None else None


class GenericMethodsHIP:
    def import_comm(self, comm):
        self.comm = comm
        self.dev_type = comm.dev_type
        self.plat_name = comm.plat_name

    def set_global_working_group_size(self):
        self.WGS = 256

    def call(self, name, *args):
        """Calls the precision-suffixed entry point on the current stream."""
        _lib.call(name + self.comm.sfx, *args, self.comm.stream())

    def call_raw(self, name, *args):
        """Calls an entry point without a precision suffix on the current stream."""
        _lib.call(name, *args, self.comm.stream())

    def dev_arr(self, val=None, shape=(1,), dtype=np.double):
        """generic_methods_cl.py:78-87: host ndarray -> device, 0 -> zeros,
        None -> uninitialised, scalar -> filled."""
        dev = self.comm.device
        if isinstance(val, torch.Tensor):
            return val.to(dev)
        if isinstance(val, np.ndarray):
            tdt = _DTYPES.get(val.dtype, None)
            a = val.view(np.int32) if val.dtype == np.uint32 else val
            return torch.from_numpy(np.ascontiguousarray(a)).to(dev, tdt)
        tdt = _DTYPES.get(np.dtype(dtype), dtype)
        if val is None:
            return torch.empty(shape, dtype=tdt, device=dev)
        arr = torch.zeros(shape, dtype=tdt, device=dev)
        if val != 0:
            arr.fill_(val)
        return arr

    def set_to(self, arr, val):
        if val == 0:
            arr.zero_()
        else:
            arr.fill_(val)

    def send_args_to_dev(self):
        """generic_methods_cl.py:56-76 (ints, floats and ndarrays mirrored)."""
        dont_send = self.Args.get("dont_send", [])
        for key in list(self.Args.keys()):
            if key in dont_send or key in ("dont_send", "dont_keep"):
                continue
            v = dict.__getitem__(self.Args, key)
            if type(v) is int:
                self.DataDev[key] = self.dev_arr(np.array([v], dtype=np.int64))
            elif type(v) is float:
                self.DataDev[key] = self.dev_arr(np.array([v], dtype=np.float64))
            elif isinstance(v, np.ndarray) and v.dtype != object:
                self.DataDev[key] = self.dev_arr(v)
