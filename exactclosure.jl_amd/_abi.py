"""ctypes binding of libexactclosure (include/exactclosure.h).

The library is loaded from this package directory.  There is no fallback:
if the shared library is missing or no GPU is present the calls raise.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libexactclosure.so")

EC_OK = 0
EC_E_INVALID = -1
EC_E_HIP = -2
EC_E_OOM = -3
EC_E_UNSUPPORTED = -4
EC_F32 = 0
EC_F64 = 1

SRC = {"u1": 0, "u2": 1, "u3": 2, "s11": 3, "s22": 4, "s33": 5, "s12": 6, "s23": 7,
       "s31": 8, "r11": 9, "r22": 10, "r33": 11, "p": 12, "k1": 13, "k2": 14, "k3": 15}

c_int, c_double, c_size_t, c_void_p = ctypes.c_int, ctypes.c_double, ctypes.c_size_t, ctypes.c_void_p
c_u64 = ctypes.c_uint64
P3 = c_void_p * 3
P6 = c_void_p * 6
P9 = c_void_p * 9


class ECError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libexactclosure error {code}: {msg}")
        self.code = code


class FilterQuantity(ctypes.Structure):
    _fields_ = [("source", c_int), ("len", c_int * 3), ("off", c_int * 3),
                ("w", ctypes.POINTER(c_double) * 3), ("out", c_void_p)]


class VelSet(ctypes.Structure):
    _fields_ = [("nbar", c_int), ("kind", c_int), ("out", c_void_p * 3)]


class ClosureSet(ctypes.Structure):
    _fields_ = [("nbar", c_int), ("kind", c_int), ("ubar", c_void_p * 3), ("flux", c_void_p * 9)]


# name -> (restype, argtypes)
_SIGS = {
    "ec_create": (c_int, [ctypes.POINTER(c_void_p), c_int, c_void_p]),
    "ec_destroy": (c_int, [c_void_p]),
    "ec_last_error": (ctypes.c_char_p, [c_void_p]),
    "ec_sync": (c_int, [c_void_p]),
    "ec_version": (c_int, []),
    "ec_capture_begin": (c_int, [c_void_p]),
    "ec_capture_end": (c_int, [c_void_p, ctypes.POINTER(c_void_p)]),
    "ec_graph_launch": (c_int, [c_void_p, c_void_p]),
    "ec_graph_destroy": (c_int, [c_void_p, c_void_p]),
    "ec_stream_after": (c_int, [c_void_p, c_void_p]),
    "ec_alloc": (c_int, [c_void_p, c_size_t, ctypes.POINTER(c_void_p)]),
    "ec_free": (c_int, [c_void_p, c_void_p]),
    "ec_h2d": (c_int, [c_void_p, c_void_p, c_void_p, c_size_t]),
    "ec_d2h": (c_int, [c_void_p, c_void_p, c_void_p, c_size_t]),
    "ec_d2d": (c_int, [c_void_p, c_void_p, c_void_p, c_size_t]),
    "ec_zero": (c_int, [c_void_p, c_void_p, c_size_t]),
    "ec_axpy": (c_int, [c_void_p, c_int, c_size_t, c_double, c_void_p, c_void_p]),
    "ec_convert": (c_int, [c_void_p, c_int, c_void_p, c_int, c_void_p, c_size_t]),
    "ec_axpy2": (c_int, [c_void_p, c_int, c_size_t, c_double, c_int, c_void_p, c_void_p]),
    "ec_axpby": (c_int, [c_void_p, c_int, c_size_t, c_double, c_void_p, c_double, c_void_p]),
    "ec_profile_begin": (c_int, [c_void_p]),
    "ec_profile_end": (c_int, [c_void_p, ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(c_double),
                              ctypes.POINTER(c_double), c_int, ctypes.POINTER(c_int)]),
    "ec_ns_stresstensor": (c_int, [c_void_p, c_int, c_int, c_double, c_double, P3, P6]),
    "ec_ns_tensordivergence": (c_int, [c_void_p, c_int, c_int, c_double, P6, P3, c_int]),
    "ec_ns_tensordivergence_nonsym": (c_int, [c_void_p, c_int, c_int, c_double, c_int, P9, P3, c_int]),
    "ec_ns_rhs": (c_int, [c_void_p, c_int, c_int, c_double, c_double, P3, P3, c_int]),
    "ec_ns_rhs_add": (c_int, [c_void_p, c_int, c_int, c_double, c_double, P3, P3, c_double, P3]),
    "ec_ns_rhs_div": (c_int, [c_void_p, c_int, c_int, c_double, c_double, P3, P3, c_void_p, c_double]),
    "ec_poisson_solve": (c_int, [c_void_p, c_void_p]),
    "ec_ns_divergence": (c_int, [c_void_p, c_int, c_int, c_double, P3, c_void_p]),
    "ec_ns_pressuregradient": (c_int, [c_void_p, c_int, c_int, c_double, P3, c_void_p]),
    "ec_ns_pressuregradient_into": (c_int, [c_void_p, c_int, c_int, c_double, P3, c_void_p, P3]),
    "ec_poisson_create": (c_int, [c_void_p, c_int, c_int, c_double, ctypes.POINTER(c_void_p)]),
    "ec_poisson_destroy": (c_int, [c_void_p, c_void_p]),
    "ec_poisson_pressure": (c_void_p, [c_void_p]),
    "ec_ns_project": (c_int, [c_void_p, c_void_p, P3]),
    "ec_ns_project_into": (c_int, [c_void_p, c_void_p, P3, P3]),
    "ec_ns_pressure_solve": (c_int, [c_void_p, c_void_p, P3]),
    "ec_fft_r2c": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "ec_fft_c2r": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "ec_reduce_maxabs": (c_int, [c_void_p, c_int, c_size_t, c_void_p, ctypes.POINTER(c_double)]),
    "ec_reduce_sumsq": (c_int, [c_void_p, c_int, c_size_t, c_void_p, ctypes.POINTER(c_double)]),
    "ec_reduce_diffsq": (c_int, [c_void_p, c_int, c_size_t, c_void_p, c_void_p,
                                 ctypes.POINTER(c_double)]),
    "ec_ns_randomfield": (c_int, [c_void_p, c_void_p, c_int, c_double, c_double, c_u64, P3]),
    "ec_ns_spectrum": (c_int, [c_void_p, c_void_p, P3, c_int, ctypes.POINTER(c_double)]),
    "ec_randn": (c_int, [c_void_p, c_int, c_size_t, c_u64, c_u64, c_void_p]),
    "ec_filter_apply": (c_int, [c_void_p, c_int, c_int, ctypes.POINTER(c_int), ctypes.POINTER(ctypes.POINTER(c_double)),
                                c_void_p, c_void_p]),
    "ec_filter_coarse": (c_int, [c_void_p, c_int, c_int, c_int, c_double, c_double, P3, c_void_p,
                                 c_int, ctypes.POINTER(FilterQuantity)]),
    "ec_filter_coarse_noalloc": (c_int, [c_void_p, c_int, c_int, c_int, c_double, c_double, P3, c_void_p,
                                 c_int, ctypes.POINTER(FilterQuantity)]),
    "ec_ns_closure": (c_int, [c_void_p, c_int, c_int, c_double, c_double, P3, c_void_p, c_int,
                              ctypes.POINTER(ClosureSet)]),
    "ec_closure_set_tiles": (c_int, [c_void_p, c_int, c_int, c_int]),
    "ec_ns_filter_vel": (c_int, [c_void_p, c_int, c_int, c_double, P3, c_void_p, c_int, ctypes.POINTER(VelSet)]),
    "ec_ns_filter_vel_rk4": (c_int, [c_void_p, c_int, c_int, c_double, P3, c_void_p, c_int, ctypes.POINTER(VelSet),
                                     c_int, c_double, P3, P3, P3]),
    "ec_ns_filter_vel_rk4_states": (c_int, [c_void_p, c_int, c_int, c_double, P3, c_void_p, c_int,
                                            ctypes.POINTER(VelSet), c_int, c_double, P3, P9]),
    "ec_filter_vel_set_tiles": (c_int, [c_void_p, c_int, c_int]),
    "ec_fft_set_cols_persistent": (c_int, [c_void_p, c_int]),
    "ec_poisson_set_z_tridiagonal": (c_int, [c_void_p, c_int]),
    "ec_ns_commutator": (c_int, [c_void_p, c_int, c_double, c_double, P3, P9, P9]),
    "ec_ns_sfs_filter": (c_int, [c_void_p, c_int, c_int, c_int, c_double, c_double, P3, c_void_p, c_int, P3, P9, P9]),
    "ec_ns_tensordissipation": (c_int, [c_void_p, c_int, c_double, P9, P3, c_void_p]),
    "ec_ns_rk4_update": (c_int, [c_void_p, c_int, c_int, c_double, c_int, c_double, P3, P3, P3,
                                 c_void_p, P3]),
    "ec_ns_rk4_update_states": (c_int, [c_void_p, c_int, c_int, c_double, c_int, c_double, P3, P9, P3,
                                        c_void_p]),
    "ec_burgers_batch_create": (c_int, [c_void_p, c_int, c_int, ctypes.POINTER(c_int),
                                        ctypes.POINTER(c_int), c_double, c_double, c_double, c_int,
                                        c_u64, ctypes.POINTER(c_void_p)]),
    "ec_burgers_batch_destroy": (c_int, [c_void_p, c_void_p]),
    "ec_burgers_batch_set_threads": (c_int, [c_void_p, c_void_p, c_int]),
    "ec_burgers_batch_run": (c_int, [c_void_p, c_void_p, c_int, c_int]),
    "ec_burgers_batch_get": (c_int, [c_void_p, c_void_p, c_int, c_int, ctypes.POINTER(c_double), c_int]),
    "ec_burgers_randomfield": (c_int, [c_void_p, c_int, c_double, c_double, c_double, c_int, c_u64, c_void_p]),
    "ec_burgers_dns_aided_les": (c_int, [c_void_p, c_int, c_int, c_double, c_double, c_double, c_double, c_int,
                                         c_void_p, c_int, ctypes.POINTER(c_double), c_int, ctypes.POINTER(c_double),
                                         c_void_p, ctypes.POINTER(c_int)]),
    "ec_burgers_batch_dt": (c_int, [c_void_p, c_void_p, c_int, ctypes.POINTER(c_double)]),
    "ec_burgers_rhs": (c_int, [c_void_p, c_int, c_int, c_double, c_double, c_void_p, c_void_p]),
    "ec_burgers_fractions": (c_int, [c_void_p, c_int, c_int, c_double, c_double, c_int, c_void_p, c_int,
                                     ctypes.POINTER(c_double), c_int, ctypes.POINTER(c_double),
                                     ctypes.POINTER(c_double)]),
    "ec_burgers_dissipation": (c_int, [c_void_p, c_int, c_int, c_double, c_double, c_int, c_void_p, c_int,
                                       ctypes.POINTER(c_double), c_int, ctypes.POINTER(c_double), c_void_p]),
    "ec_burgers_smag_fit": (c_int, [c_void_p, c_int, c_int, c_double, c_double, c_double, c_int, c_void_p, c_int,
                                    ctypes.POINTER(c_double), c_int, ctypes.POINTER(c_double),
                                    ctypes.POINTER(c_double)]),
    "ec_burgers_smag_solve": (c_int, [c_void_p, c_int, c_int, c_double, c_double, c_double, c_double, c_double,
                                      c_int, c_void_p, c_int, ctypes.POINTER(c_double), c_void_p]),
    "ec_cnn_create": (c_int, [c_void_p, c_u64, ctypes.POINTER(c_void_p)]),
    "ec_cnn_destroy": (c_int, [c_void_p, c_void_p]),
    "ec_cnn_num_params": (c_int, [c_void_p]),
    "ec_cnn_params": (c_int, [c_void_p, c_void_p, c_int, c_int, ctypes.POINTER(ctypes.c_float), c_int]),
    "ec_cnn_forward": (c_int, [c_void_p, c_void_p, c_int, c_int, ctypes.POINTER(c_void_p), c_double,
                               ctypes.POINTER(c_void_p), c_double]),
    "ec_cnn_backward": (c_int, [c_void_p, c_void_p, ctypes.POINTER(c_void_p), c_double, ctypes.POINTER(c_double)]),
    "ec_cnn_adam": (c_int, [c_void_p, c_void_p, c_double, c_double, c_double, c_double]),
}

_lib = None


def lib():
    """Load libexactclosure.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run `python exactclosure.jl_amd/build.py`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name, None)
            if f is None:
                continue
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def declared_symbols(header=None):
    """Function names declared in include/exactclosure.h."""
    import re
    header = header or os.path.join(os.path.dirname(HERE), "include", "exactclosure.h")
    txt = open(header).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*|void \*)\s*(ec_\w+)\s*\(", txt, re.M)))


class Context:
    """One device + one stream (the current torch stream at creation)."""

    def __init__(self, device=0, stream=None):
        self.lib = lib()
        h = c_void_p()
        rc = self.lib.ec_create(ctypes.byref(h), device, stream)
        if rc != EC_OK:
            raise ECError(rc, f"ec_create(device={device}) failed (no usable GPU?)")
        self.h = h
        self.device = device
        self.stream = stream  # the HIP stream handle (None: the device's default stream)

    def check(self, rc, what=""):
        if rc != EC_OK:
            msg = self.lib.ec_last_error(self.h)
            raise ECError(rc, f"{what}: {msg.decode() if msg else ''}")

    def call(self, name, *args):
        self.check(getattr(self.lib, name)(self.h, *args), name)

    def sync(self):
        self.call("ec_sync")

    def close(self):
        if self.h:
            self.lib.ec_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_ctx = {}


def context(device=0):
    """The process-wide context of `device`, ordered on torch's current stream."""
    import torch

    if device not in _ctx:
        if not torch.cuda.is_available():
            raise RuntimeError("no GPU: libexactclosure needs a HIP device")
        stream = torch.cuda.current_stream(device).cuda_stream
        _ctx[device] = Context(device, stream or None)
    return _ctx[device]


def ptr(t):
    return c_void_p(t.data_ptr())


def ptrs(ts, kind):
    return kind(*[t.data_ptr() if t is not None else None for t in ts])


def dtype_code(t):
    import torch

    if t.dtype == torch.float64:
        return EC_F64
    if t.dtype == torch.float32:
        return EC_F32
    raise TypeError(f"unsupported dtype {t.dtype}")
