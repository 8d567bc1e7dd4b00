"""ctypes binding of libldga_hip.so (declared in include/ldga_hip.h).

Loading fails loudly when the library has not been built: there is no CPU
fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parents[1] / "lib" / "libldga_hip.so"

LDGA_OK = 0
LDGA_ERR_SHAPE = -1
LDGA_ERR_HIP = -2
LDGA_ERR_ZERO_PIVOT = -3
LDGA_ERR_ARG = -4
LDGA_ERR_STATE = -5

PREC_FP64 = 0
PREC_MIXED = 1

EOM_KINDS = {"full": 0, "χ_m": 1, "γ_m": 2, "χ_d": 3, "γ_d": 4, "rest_01": 5, "rest": 6,
             "sp_01": 7, "sp_02": 8, "sp": 9, "ch": 10}

# every symbol include/ldga_hip.h declares
EXPORTED = [
    "ldga_version", "ldga_create", "ldga_destroy", "ldga_last_error", "ldga_sync",
    "ldga_set_kgrid", "ldga_default_wedge",
    "ldga_g_tables", "ldga_g_tables_dev", "ldga_bubble", "ldga_bubble_dev",
    "ldga_chi_gamma", "ldga_chi_gamma_dev", "ldga_chi_gamma_noalloc_dev",
    "ldga_lambda0", "ldga_lambda0_dev", "ldga_F_from_chi", "ldga_F_from_chi_dev", "ldga_gen_chi", "ldga_gen_chi_dev",
    "ldga_sum_kw", "ldga_sum_kw_dev", "ldga_sum_w", "ldga_sum_w_dev",
    "ldga_sigma", "ldga_sigma_dev",
    "ldga_malloc", "ldga_free", "ldga_memcpy_h2d", "ldga_memcpy_d2h",
    "ldga_G_shell_sum", "ldga_lambdadm_residual", "ldga_lambdadm_residual_dev", "ldga_g_from_sigma_ladder_dev",
]


class Freq(C.Structure):
    _fields_ = [("n_iw", C.c_int32), ("n_inu", C.c_int32), ("n_inu_shell", C.c_int32),
                ("shift", C.c_int32), ("fft_first", C.c_int32), ("n_fft", C.c_int32)]


class Model(C.Structure):
    _fields_ = [("U", C.c_double), ("mu", C.c_double), ("beta", C.c_double), ("n", C.c_double),
                ("Ekin_DMFT", C.c_double), ("chi_m_c2", C.c_double), ("chi_d_c2", C.c_double)]


class BseReport(C.Structure):
    _fields_ = [("n_matrices", C.c_int64), ("n_zero_pivot", C.c_int64), ("ir_steps_max", C.c_int32),
                ("n_fallback", C.c_int32), ("max_rel_resid", C.c_double), ("ir_steps_mean", C.c_double),
                ("n_fp64_resolved", C.c_int64)]


class LambdaDmOut(C.Structure):
    _fields_ = [("n", C.c_double), ("mu", C.c_double), ("Ekin", C.c_double), ("Epot1", C.c_double),
                ("Epot2", C.c_double), ("lhs_c1", C.c_double), ("res_pp", C.c_double), ("res_epot", C.c_double),
                ("jac", C.c_double * 4), ("mu_iters", C.c_int32), ("mu_converged", C.c_int32),
                ("n_full", C.c_double)]


class LdgaError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libldga_hip error {code}: {msg}")
        self.code = code


_lib = None


def lib():
    """The loaded library (raises if it is missing)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() "
                              "(there is no CPU fallback for the ladder kernels)")
        L = C.CDLL(str(LIB_PATH))
        vp, i32, dbl = C.c_void_p, C.c_int, C.c_double
        sig = {
            "ldga_version": (i32, []),
            "ldga_create": (i32, [i32, C.POINTER(vp)]),
            "ldga_destroy": (i32, [vp]),
            "ldga_last_error": (C.c_char_p, [vp]),
            "ldga_sync": (i32, [vp]),
            "ldga_set_kgrid": (i32, [vp, i32, i32, i32, vp, vp]),
            "ldga_default_wedge": (i32, [i32, i32, vp, vp]),
            "ldga_g_tables": (i32, [vp, C.POINTER(Freq), vp, vp, vp]),
            "ldga_g_tables_dev": (i32, [vp, C.POINTER(Freq), vp, vp, vp, vp]),
            "ldga_bubble": (i32, [vp, C.POINTER(Freq), dbl, vp, vp, vp]),
            "ldga_bubble_dev": (i32, [vp, C.POINTER(Freq), dbl, vp, vp, vp, vp]),
            "ldga_chi_gamma": (i32, [vp, i32, i32, C.POINTER(Freq), dbl, dbl, vp, vp, vp, vp, C.POINTER(BseReport)]),
            "ldga_chi_gamma_dev": (i32, [vp, i32, i32, C.POINTER(Freq), dbl, dbl, vp, vp, vp, vp,
                                         C.POINTER(BseReport), vp]),
            "ldga_chi_gamma_noalloc_dev": (i32, [vp, i32, i32, C.POINTER(Freq), dbl, dbl, vp, vp, vp, vp,
                                                 C.POINTER(BseReport), vp]),
            "ldga_lambda0": (i32, [vp, C.POINTER(Freq), dbl, vp, vp, vp]),
            "ldga_F_from_chi": (i32, [vp, C.POINTER(Freq), dbl, i32, vp, vp, vp]),
            "ldga_gen_chi": (i32, [vp, C.POINTER(Freq), vp, vp, vp]),
            "ldga_gen_chi_dev": (i32, [vp, C.POINTER(Freq), vp, vp, vp, vp]),
            "ldga_F_from_chi_dev": (i32, [vp, C.POINTER(Freq), dbl, i32, vp, vp, vp, vp]),
            "ldga_lambda0_dev": (i32, [vp, C.POINTER(Freq), dbl, vp, vp, vp, vp]),
            "ldga_sum_kw": (i32, [vp, vp, i32, i32, i32, dbl, dbl, dbl, vp]),
            "ldga_sum_kw_dev": (i32, [vp, vp, i32, i32, i32, dbl, dbl, dbl, vp, vp]),
            "ldga_sum_w": (i32, [vp, vp, i32, i32, i32, i32, dbl, dbl, dbl, vp, vp]),
            "ldga_sum_w_dev": (i32, [vp, vp, i32, i32, i32, i32, dbl, dbl, dbl, vp, vp, vp]),
            "ldga_sigma": (i32, [vp, C.POINTER(Freq), C.POINTER(Model), i32, dbl, dbl, i32, i32, dbl,
                                 vp, vp, vp, vp, vp, vp, vp]),
            "ldga_sigma_dev": (i32, [vp, C.POINTER(Freq), C.POINTER(Model), i32, dbl, dbl, i32, i32, dbl,
                                     vp, vp, vp, vp, vp, vp, vp, vp]),
            "ldga_malloc": (i32, [vp, C.c_uint64, C.POINTER(vp)]),
            "ldga_free": (i32, [vp, vp]),
            "ldga_memcpy_h2d": (i32, [vp, vp, vp, C.c_uint64]),
            "ldga_memcpy_d2h": (i32, [vp, vp, vp, C.c_uint64]),
            "ldga_G_shell_sum": (dbl, [i32, dbl]),
            "ldga_lambdadm_residual": (i32, [vp, C.POINTER(Freq), C.POINTER(Model), i32, dbl, dbl, i32, dbl, i32,
                                             dbl, vp, vp, vp, vp, vp, vp, vp, i32, C.POINTER(LambdaDmOut), vp, vp]),
            "ldga_lambdadm_residual_dev": (i32, [vp, C.POINTER(Freq), C.POINTER(Model), i32, dbl, dbl, i32, dbl,
                                                 i32, dbl, vp, vp, vp, vp, vp, vp, vp, i32, C.POINTER(LambdaDmOut),
                                                 vp, vp, vp]),
            "ldga_g_from_sigma_ladder_dev": (i32, [vp, C.POINTER(Freq), C.POINTER(Model), i32, dbl, dbl, i32, dbl,
                                                   vp, vp, vp, vp, dbl, vp, C.POINTER(LambdaDmOut), vp, vp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def ptr(a: np.ndarray, dtype, name: str) -> int:
    """Address of a Fortran-contiguous array of the given dtype (Julia layout)."""
    if not isinstance(a, np.ndarray) or a.dtype != dtype or not a.flags.f_contiguous:
        raise TypeError(f"{name} must be a Fortran-ordered {np.dtype(dtype)} array")
    return a.ctypes.data


class Context:
    """One libldga_hip context (device, stream, cached buffers, current k grid)."""

    def __init__(self, device: int = 0):
        self._h = C.c_void_p()
        rc = lib().ldga_create(device, C.byref(self._h))
        if rc != LDGA_OK:
            raise LdgaError(rc, f"ldga_create(device={device}) failed (no HIP device?)")
        self.device = device
        self.kgrid_key = None

    @property
    def handle(self):
        return self._h

    def check(self, rc: int, allow=()):
        if rc != LDGA_OK and rc not in allow:
            raise LdgaError(rc, lib().ldga_last_error(self._h).decode())
        return rc

    def set_kgrid(self, kG):
        key = (kG.D, kG.Ns, kG.wedge.tobytes())
        if key != self.kgrid_key:
            w = np.ascontiguousarray(kG.wedge, dtype=np.int32)
            e = np.ascontiguousarray(kG.eps_wedge, dtype=np.float64)
            self.check(lib().ldga_set_kgrid(self._h, kG.D, kG.Ns, kG.Nq, w.ctypes.data, e.ctypes.data))
            self.kgrid_key = key

    def sync(self):
        self.check(lib().ldga_sync(self._h))

    # resident device arrays ------------------------------------------------
    def to_device(self, a: np.ndarray) -> "DeviceArray":
        a = np.asfortranarray(a)
        d = DeviceArray(self, a.nbytes, a.shape, a.dtype)
        self.check(lib().ldga_memcpy_h2d(self._h, d.ptr, a.ctypes.data, a.nbytes))
        return d

    def empty(self, shape, dtype) -> "DeviceArray":
        return DeviceArray(self, int(np.prod(shape)) * np.dtype(dtype).itemsize, shape, dtype)

    def close(self):
        if self._h:
            lib().ldga_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceArray:
    """A device buffer of the context with a (Fortran-order) shape and dtype."""

    def __init__(self, ctx: Context, nbytes: int, shape, dtype):
        self.ctx, self.nbytes, self.shape, self.dtype = ctx, nbytes, tuple(shape), np.dtype(dtype)
        p = C.c_void_p()
        ctx.check(lib().ldga_malloc(ctx.handle, nbytes, C.byref(p)))
        self._p = p

    @property
    def ptr(self):
        return self._p.value

    def to_host(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype, order="F")
        self.ctx.check(lib().ldga_memcpy_d2h(self.ctx.handle, out.ctypes.data, self.ptr, self.nbytes))
        return out

    def free(self):
        if self._p and self._p.value:
            lib().ldga_free(self.ctx.handle, self._p)
            self._p = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


_default = None


def default_context() -> Context:
    global _default
    if _default is None:
        _default = Context(int(os.environ.get("LDGA_DEVICE", "0")))
    return _default


def default_wedge(D: int, Ns: int):
    """The library's wedge for (D, Ns): (coords (Nq, D) int32, kMult (Nq,))."""
    L = lib()
    Nq = L.ldga_default_wedge(D, Ns, None, None)
    if Nq <= 0:
        raise LdgaError(Nq, f"no wedge for D={D}, Ns={Ns}")
    w = np.empty((Nq, D), dtype=np.int32)
    m = np.empty(Nq, dtype=np.float64)
    L.ldga_default_wedge(D, Ns, w.ctypes.data, m.ctypes.data)
    return w, m
