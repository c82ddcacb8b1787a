"""ctypes binding of the native engine libm2.so (C ABI in include/m2.h).

There is no CPU fallback: if the library is missing this module raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "lib", "libm2.so")

M2_FP64, M2_MIXED = 0, 1
M2_ORBITAL, M2_AUXILIARY = 0, 1

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_lib = None


class M2Error(RuntimeError):
    """Mirrors the reference's Error(code, msg) (inc/error.hpp:18-35)."""

    def __init__(self, code, msg):
        super().__init__("M2[%d]: %s" % (code, msg))
        self.code = code
        self.msg = msg


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError("libm2.so is not built (%s); run `python -m molecular2_amd.build`" % LIBPATH)
        L = C.CDLL(LIBPATH)
        vp = C.c_void_p
        L.m2_ctx_create.argtypes = [C.POINTER(vp), C.c_int]
        L.m2_ctx_destroy.argtypes = [vp]
        L.m2_last_error.argtypes = [vp]
        L.m2_last_error.restype = C.c_char_p
        L.m2_ctx_stream.argtypes = [vp]
        L.m2_ctx_stream.restype = vp
        L.m2_basis_set.argtypes = [vp, C.c_int, C.c_int, _ip, _ip, _dp, _dp, _dp, _ip, C.c_int]
        L.m2_basis_info.argtypes = [vp, C.c_int, _ip, _ip]
        L.m2_shell_coefs.argtypes = [vp, C.c_int, C.c_int, _dp]
        L.m2_schwarz.argtypes = [vp, _dp]
        L.m2_shellblock_norm.argtypes = [vp, _dp, _dp]
        L.m2_one_body.argtypes = [vp, C.c_int, _dp, _dp, _dp, _dp, _dp]
        L.m2_jk_direct.argtypes = [vp, _dp, C.c_double, C.c_double, C.c_uint, _dp, _dp]
        L.m2_jk_direct_device.argtypes = [vp, vp, C.c_double, C.c_double, C.c_uint, vp, vp]
        L.m2_set_mixed_threshold.argtypes = [vp, C.c_double]
        L.m2_last_stats.argtypes = [vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
        for name in ("m2_ctx_create", "m2_basis_set", "m2_basis_info", "m2_shell_coefs", "m2_schwarz",
                     "m2_shellblock_norm", "m2_one_body", "m2_jk_direct", "m2_jk_direct_device",
                     "m2_set_mixed_threshold", "m2_last_stats"):
            getattr(L, name).restype = C.c_int
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_dp)


def shell_arrays(shells):
    l = np.array([s.l for s in shells], np.int32)
    npr = np.array([len(s.exps) for s in shells], np.int32)
    exps = np.ascontiguousarray(np.concatenate([np.asarray(s.exps, float) for s in shells]))
    coefs = np.ascontiguousarray(np.concatenate([np.asarray(s.coefs, float) for s in shells]))
    cen = np.ascontiguousarray(np.array([s.center for s in shells], float).reshape(-1))
    at = np.array([s.atom for s in shells], np.int32)
    return l, npr, exps, coefs, cen, at


class Engine:
    """One device context (libm2 m2_ctx)."""

    def __init__(self, device: int = 0):
        self.L = lib()
        h = C.c_void_p()
        st = self.L.m2_ctx_create(C.byref(h), device)
        if st != 0:
            raise M2Error(st, "m2_ctx_create failed (device %d)" % device)
        self.h = h
        self.nbf = 0
        self.nsh = 0

    def close(self):
        if getattr(self, "h", None):
            self.L.m2_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, st):
        if st != 0:
            raise M2Error(st, self.L.m2_last_error(self.h).decode())

    @property
    def stream(self) -> int:
        return self.L.m2_ctx_stream(self.h)

    def set_basis(self, shells, pure=True, which=M2_ORBITAL):
        l, npr, exps, coefs, cen, at = shell_arrays(shells)
        self._keep = (l, npr, exps, coefs, cen, at)
        self._chk(self.L.m2_basis_set(self.h, which, len(shells), l.ctypes.data_as(_ip), npr.ctypes.data_as(_ip),
                                      _p(exps), _p(coefs), _p(cen), at.ctypes.data_as(_ip), int(pure)))
        nbf, nsh = C.c_int(), C.c_int()
        self._chk(self.L.m2_basis_info(self.h, which, C.byref(nbf), C.byref(nsh)))
        if which == M2_ORBITAL:
            self.nbf, self.nsh = nbf.value, nsh.value
        return nbf.value

    def shell_coefs(self, shell, nprim, which=M2_ORBITAL):
        out = np.zeros(nprim)
        self._chk(self.L.m2_shell_coefs(self.h, which, shell, _p(out)))
        return out

    def schwarz(self):
        Q = np.zeros((self.nsh, self.nsh))
        self._chk(self.L.m2_schwarz(self.h, _p(Q)))
        return Q

    def shellblock_norm(self, A):
        A = np.ascontiguousarray(A, float)
        out = np.zeros((self.nsh, self.nsh))
        self._chk(self.L.m2_shellblock_norm(self.h, _p(A), _p(out)))
        return out

    def one_body(self, Z, xyz):
        n = self.nbf
        S, T, V = np.zeros((n, n)), np.zeros((n, n)), np.zeros((n, n))
        Zf = np.ascontiguousarray(Z, float)
        X = np.ascontiguousarray(xyz, float)
        self._chk(self.L.m2_one_body(self.h, len(Zf), _p(Zf), _p(X), _p(S), _p(T), _p(V)))
        return S, T, V

    def jk(self, D, multiplier=1.0, precision=1e-12, mixed=False):
        D = np.ascontiguousarray(D, float)
        J = np.zeros_like(D)
        K = np.zeros_like(D)
        self._chk(self.L.m2_jk_direct(self.h, _p(D), multiplier, precision, M2_MIXED if mixed else M2_FP64,
                                      _p(J), _p(K)))
        return J, K

    def jk_device(self, dD_ptr, dJ_ptr, dK_ptr, multiplier=1.0, precision=1e-12, mixed=False):
        self._chk(self.L.m2_jk_direct_device(self.h, C.c_void_p(dD_ptr), multiplier, precision,
                                             M2_MIXED if mixed else M2_FP64, C.c_void_p(dJ_ptr),
                                             C.c_void_p(dK_ptr)))

    def set_mixed_threshold(self, t):
        self._chk(self.L.m2_set_mixed_threshold(self.h, t))

    def last_stats(self):
        a, b, c = C.c_longlong(), C.c_longlong(), C.c_longlong()
        self._chk(self.L.m2_last_stats(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value
