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


class M2Fragment(C.Structure):
    """m2_fragment (include/m2.h): one fragment of the fragment-local DF-JK"""
    _fields_ = [("occ", C.c_int), ("nbfs", C.c_int), ("naux", C.c_int), ("start", C.c_int), ("auxstart", C.c_int),
                ("nshells", C.c_int), ("ndfshells", C.c_int), ("radius", C.c_double), ("com", C.c_double * 3)]


class SCFOptions(C.Structure):
    """Options of the host C++ SCF (the reference's RHF defaults,
    src/ProgramController.cpp:295-307, plus the precision policy)."""
    _fields_ = [("diis", C.c_int), ("maxdiis", C.c_int), ("maxiter", C.c_int),
                ("enconverge", C.c_double), ("densconverge", C.c_double), ("precision", C.c_double),
                ("df", C.c_int), ("mixed", C.c_int), ("mixed_threshold", C.c_double), ("fp64_final", C.c_int),
                ("guess", C.c_int), ("diis_start", C.c_int), ("incore", C.c_int)]


class SCFResult(C.Structure):
    _fields_ = [("energy", C.c_double), ("one_e", C.c_double), ("two_e", C.c_double), ("enuc", C.c_double),
                ("iterations", C.c_int), ("converged", C.c_int), ("n_full", C.c_int),
                ("n_incremental", C.c_int), ("n_fp32", C.c_int),
                ("jk_seconds", C.c_double), ("total_seconds", C.c_double), ("grad_seconds", C.c_double)]


class DeckInfo(C.Structure):
    _fields_ = [("natom", C.c_int), ("charge", C.c_int), ("multiplicity", C.c_int), ("nsh", C.c_int),
                ("nprim", C.c_int), ("nsh_aux", C.c_int), ("nprim_aux", C.c_int), ("nsh_min", C.c_int),
                ("nprim_min", C.c_int)]


class ShellsOut(C.Structure):
    _fields_ = [("l", C.POINTER(C.c_int)), ("nprim", C.POINTER(C.c_int)), ("exps", C.POINTER(C.c_double)),
                ("coefs", C.POINTER(C.c_double)), ("centers", C.POINTER(C.c_double)),
                ("atoms", C.POINTER(C.c_int))]


class DeckResult(C.Structure):
    _fields_ = [("command", C.c_int), ("molecule", C.c_int), ("method", C.c_int), ("energy", C.c_double),
                ("iterations", C.c_int), ("converged", C.c_int), ("seconds", C.c_double)]


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
        L.m2_jk_direct_device_noalloc.argtypes = [vp, vp, C.c_double, C.c_double, C.c_uint, vp, vp]
        L.m2_set_mixed_threshold.argtypes = [vp, C.c_double]
        L.m2_set_jk_families.argtypes = [vp, C.c_int]
        L.m2_set_jk_families.restype = C.c_int
        L.m2_last_stats.argtypes = [vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
        L.m2_set_profiling.argtypes = [vp, C.c_int]
        _ll = C.POINTER(C.c_longlong)
        L.m2_get_profile.argtypes = [vp, _dp, _ll, _ll, _ll]
        L.m2_reset_profile.argtypes = [vp]
        L.m2_get_profile_computed.argtypes = [vp, _ll]
        L.m2_set_prim_screening.argtypes = [vp, C.c_double]
        L.m2_set_prim_screening_fp32.argtypes = [vp, C.c_double]
        L.m2_df_setup.argtypes = [vp]
        L.m2_jk_df.argtypes = [vp, _dp, C.c_int, C.c_int, _dp, _dp]
        L.m2_jk_df_device.argtypes = [vp, vp, C.c_int, C.c_int, vp, vp]
        L.m2_df_timing.argtypes = [vp, _dp]
        L.m2_eigh.argtypes = [vp, C.c_int, _dp, _dp, _dp]
        L.m2_gemm.argtypes = [vp, C.c_char, C.c_char, C.c_int, C.c_int, C.c_int, C.c_double, _dp, C.c_int, _dp,
                              C.c_int, C.c_double, _dp, C.c_int]
        L.m2_rhf_last_error.restype = C.c_char_p
        L.m2_rhf.argtypes = [vp, C.c_int, _dp, _dp, C.c_int, C.c_int, C.c_int,
                             C.c_int, _ip, _ip, _dp, _dp, _dp, _ip,
                             C.c_int, _ip, _ip, _dp, _dp, _dp, _ip,
                             C.c_int, _ip, _ip, _dp, _dp, _dp, _ip,
                             C.POINTER(SCFOptions), C.POINTER(SCFResult), _dp, _dp, _dp, _dp, _dp]
        L.m2_uhf.argtypes = [vp, C.c_int, _dp, _dp, C.c_int, C.c_int, C.c_int,
                             C.c_int, _ip, _ip, _dp, _dp, _dp, _ip,
                             C.c_int, _ip, _ip, _dp, _dp, _dp, _ip,
                             C.c_int, _ip, _ip, _dp, _dp, _dp, _ip,
                             C.POINTER(SCFOptions), C.POINTER(SCFResult), _dp, _dp, _dp, _dp, _dp]
        L.m2_uhf.restype = C.c_int
        L.m2_jk_direct_uhf.argtypes = [vp, _dp, _dp, C.c_double, C.c_double, C.c_uint, _dp, _dp, _dp, _dp]
        L.m2_jk_direct_uhf.restype = C.c_int
        L.m2_soad_fock.argtypes = [C.c_int, C.c_int, _dp, _dp, C.c_int,
                                   C.c_int, _ip, _ip, _dp, _dp, _dp, _ip,
                                   C.c_int, _ip, _ip, _dp, _dp, _dp, _ip, _dp]
        L.m2_soad_fock.restype = C.c_int
        L.m2_rhf_batch.argtypes = [C.c_int, C.c_int, C.c_int, _ip, _ip, _dp, _dp, C.c_int, _ip, _ip, _ip, _dp, _dp,
                                   _dp, _ip, C.POINTER(SCFOptions), C.POINTER(SCFResult)]
        L.m2_rhf_batch.restype = C.c_int
        L.m2_rhf_batch_min.argtypes = [C.c_int, C.c_int, C.c_int, _ip, _ip, _dp, _dp, C.c_int, _ip, _ip, _ip, _dp,
                                       _dp, _dp, _ip, _ip, _ip, _ip, _dp, _dp, _dp, _ip, C.POINTER(SCFOptions),
                                       C.POINTER(SCFResult)]
        L.m2_rhf_batch_min.restype = C.c_int
        L.m2_rhf_hessian.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, C.c_int, C.c_int, C.c_int,
                                     C.c_int, _ip, _ip, _dp, _dp, _dp, _ip,
                                     C.c_int, _ip, _ip, _dp, _dp, _dp, _ip,
                                     C.c_int, _ip, _ip, _dp, _dp, _dp, _ip,
                                     C.POINTER(SCFOptions), C.c_double, _dp, C.POINTER(SCFResult)]
        L.m2_rhf_hessian.restype = C.c_int
        L.m2_deck_molecule.argtypes = [C.c_char_p, C.c_int, C.c_char_p, C.POINTER(DeckInfo), _dp, _dp,
                                       C.POINTER(ShellsOut), C.POINTER(ShellsOut), C.POINTER(ShellsOut)]
        L.m2_deck_molecule.restype = C.c_int
        L.m2_basis_file_shells.argtypes = [C.c_char_p, C.c_int, _ip, _dp, _ip, _ip, C.POINTER(ShellsOut)]
        L.m2_basis_file_shells.restype = C.c_int
        L.m2_print_eri.argtypes = [vp, C.c_char_p, C.c_double, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
        L.m2_print_eri.restype = C.c_int
        L.m2_eri_packed.argtypes = [vp, _dp]
        L.m2_eri_packed.restype = C.c_int
        L.m2_deck_run.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.c_int, C.POINTER(DeckResult), C.c_int,
                                  C.POINTER(C.c_int)]
        L.m2_deck_run.restype = C.c_int
        L.m2_ctx_device.argtypes = [vp]
        L.m2_ctx_device.restype = C.c_int
        L.m2_grad.argtypes = [vp, C.c_int, _dp, _dp, _dp, _dp, C.c_double, _dp]
        L.m2_grad_ex.argtypes = [vp, C.c_int, _dp, _dp, _dp, _dp, C.c_double, C.c_uint, _dp]
        for name in ("m2_grad", "m2_df_setup", "m2_jk_df", "m2_jk_df_device", "m2_df_timing", "m2_eigh", "m2_gemm", "m2_rhf"):
            getattr(L, name).restype = C.c_int
        L.m2_jk_df_local.argtypes = [vp, C.c_int, C.POINTER(M2Fragment), C.c_double, C.c_double, C.c_double,
                                     C.c_double, _dp, C.c_int, _dp, _dp, _dp, _dp]
        L.m2_df_local_reset.argtypes = [vp]
        L.m2_df_local_domains.argtypes = [vp, C.c_int, _ip, _ip]
        for name in ("m2_jk_df_local", "m2_df_local_reset", "m2_df_local_domains"):
            getattr(L, name).restype = C.c_int
        L.m2_rhf_hessian_cphf.argtypes = [vp, C.c_int, _dp, _dp, C.c_int, C.c_int, C.c_int,
                                          C.c_int, _ip, _ip, _dp, _dp, _dp, _ip,
                                          C.c_int, _ip, _ip, _dp, _dp, _dp, _ip,
                                          C.POINTER(SCFOptions), C.c_double, C.c_double, C.c_int, _dp,
                                          C.POINTER(SCFResult), _ip]
        L.m2_rhf_hessian_cphf.restype = C.c_int
        L.m2_set_jk_launch.argtypes = [vp, C.c_int]
        L.m2_set_jk_graphs.argtypes = [vp, C.c_int]
        L.m2_jk_graph_stats.argtypes = [vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong), _ip]
        L.m2_set_jk_block_tasks.argtypes = [vp, C.c_int, C.c_int, C.c_int]
        L.m2_get_jk_launch.argtypes = [vp, _ip, _ip, _ip]
        L.m2_jk_autotune.argtypes = [vp, vp, C.c_double, C.c_uint, C.c_int, _dp, _dp]
        for name in ("m2_set_jk_launch", "m2_set_jk_graphs", "m2_jk_graph_stats", "m2_set_jk_block_tasks",
                     "m2_get_jk_launch", "m2_jk_autotune"):
            getattr(L, name).restype = C.c_int
        for name in ("m2_set_profiling", "m2_get_profile", "m2_reset_profile", "m2_ctx_create", "m2_basis_set", "m2_basis_info", "m2_shell_coefs", "m2_schwarz",
                     "m2_shellblock_norm", "m2_one_body", "m2_jk_direct", "m2_jk_direct_device",
                     "m2_jk_direct_device_noalloc", "m2_get_profile_computed", "m2_set_prim_screening", "m2_set_prim_screening_fp32",
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


def _scf_shells(mol, basis, aux, guess):
    """The 21 shell arguments of m2_rhf / m2_uhf / m2_rhf_hessian (orbital,
    auxiliary, minimal STO-3G for the SOAD guess) for `mol`, and the orbital
    shells.  The argument list holds the numpy arrays it points into."""
    ip = lambda a: a.ctypes.data_as(_ip)
    empty = (np.zeros(1, np.int32), np.zeros(1, np.int32), np.zeros(1), np.zeros(1), np.zeros(1),
             np.zeros(1, np.int32))

    def group(shells):
        if not shells:
            return [0] + [ip(a) if a.dtype == np.int32 else _p(a) for a in empty], empty
        arrs = shell_arrays(shells)
        return [len(shells)] + [ip(a) if a.dtype == np.int32 else _p(a) for a in arrs], arrs

    sh = mol.shells(basis)
    args, keep = [], []
    for shells in (sh, mol.shells(aux) if aux else None, mol.shells("sto-3g") if guess == "soad" else None):
        a, k = group(shells)
        args += a
        keep.append(k)
    args.append(keep)        # not passed: only holds the arrays
    return sh, args


def _shells_buffers(nsh, nprim):
    b = {"l": np.zeros(nsh, np.int32), "nprim": np.zeros(nsh, np.int32), "exps": np.zeros(max(nprim, 1)),
         "coefs": np.zeros(max(nprim, 1)), "centers": np.zeros(max(3 * nsh, 1)), "atoms": np.zeros(nsh, np.int32)}
    o = ShellsOut(b["l"].ctypes.data_as(_ip), b["nprim"].ctypes.data_as(_ip), _p(b["exps"]), _p(b["coefs"]),
                  _p(b["centers"]), b["atoms"].ctypes.data_as(_ip))
    return b, o


def _shells_from(b):
    from .molecules import Shell
    out, p = [], 0
    for s in range(len(b["l"])):
        k = int(b["nprim"][s])
        out.append(Shell(int(b["l"][s]), list(b["exps"][p:p + k]), list(b["coefs"][p:p + k]),
                         tuple(b["centers"][3 * s:3 * s + 3]), int(b["atoms"][s])))
        p += k
    return out


def basis_file_shells(path: str, mol):
    """Shells of `mol` from a Molpro basis text file (m2_basis_file_shells,
    include/m2_input.h; host code, no GPU)."""
    L = lib()
    Z = np.ascontiguousarray(mol.Z, np.int32)
    X = np.ascontiguousarray(mol.xyz, float).reshape(-1)
    nsh, npr = C.c_int(0), C.c_int(0)
    if L.m2_basis_file_shells(path.encode(), mol.natom, Z.ctypes.data_as(_ip), _p(X), C.byref(nsh), C.byref(npr),
                              None) != 0:
        raise M2Error(1, L.m2_rhf_last_error().decode())
    b, o = _shells_buffers(nsh.value, npr.value)
    if L.m2_basis_file_shells(path.encode(), mol.natom, Z.ctypes.data_as(_ip), _p(X), C.byref(nsh), C.byref(npr),
                              C.byref(o)) != 0:
        raise M2Error(1, L.m2_rhf_last_error().decode())
    return _shells_from(b)


def deck_molecule(deck_text: str, imol: int = 0, basislib: str = ""):
    """Molecule construct `imol` of a reference input deck (text of a .mol
    file) with its shells from the XML basis library directory `basislib`
    (m2_deck_molecule, include/m2_input.h; host code, no GPU).  Returns
    (Molecule, orbital shells, jkbasis shells, STO-3G shells)."""
    from .molecules import Molecule
    L = lib()
    info = DeckInfo()
    t, bl = deck_text.encode(), basislib.encode()
    if L.m2_deck_molecule(t, imol, bl, C.byref(info), None, None, None, None, None) != 0:
        raise M2Error(1, L.m2_rhf_last_error().decode())
    Z = np.zeros(info.natom)
    X = np.zeros(3 * info.natom)
    bo, oo = _shells_buffers(info.nsh, info.nprim)
    ba, oa = _shells_buffers(info.nsh_aux, info.nprim_aux)
    bm, om = _shells_buffers(info.nsh_min, info.nprim_min)
    if L.m2_deck_molecule(t, imol, bl, C.byref(info), _p(Z), _p(X), C.byref(oo), C.byref(oa), C.byref(om)) != 0:
        raise M2Error(1, L.m2_rhf_last_error().decode())
    mol = Molecule([int(z) for z in Z], X.reshape(-1, 3), charge=info.charge, multiplicity=info.multiplicity)
    return mol, _shells_from(bo), _shells_from(ba), _shells_from(bm)


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

    def jk_device_noalloc(self, dD_ptr, dJ_ptr, dK_ptr, multiplier=1.0, precision=1e-12, mixed=False):
        """jk_device through m2_jk_direct_device_noalloc: the library allocates
        nothing (its work memory is one static device array per GPU)."""
        self._chk(self.L.m2_jk_direct_device_noalloc(self.h, C.c_void_p(dD_ptr), multiplier, precision,
                                                     M2_MIXED if mixed else M2_FP64, C.c_void_p(dJ_ptr),
                                                     C.c_void_p(dK_ptr)))

    def eri_packed(self):
        """(ij|kl) of the orbital basis, packed: [p(ij), p(kl)] with
        p(ij) = i(i+1)/2 + j, i >= j (m2_eri_packed)."""
        n = self.nbf
        np_ = n * (n + 1) // 2
        out = np.zeros((np_, np_))
        self._chk(self.L.m2_eri_packed(self.h, _p(out)))
        return out

    def print_eri(self, path, thrint=1e-12):
        """IntegralEngine::printERI format written to `path` (m2_print_eri);
        returns (integral count, insignificant count)."""
        a, b = C.c_longlong(), C.c_longlong()
        if self.L.m2_print_eri(self.h, path.encode(), thrint, C.byref(a), C.byref(b)) != 0:
            raise M2Error(1, self.L.m2_rhf_last_error().decode())
        return a.value, b.value

    def set_mixed_threshold(self, t):
        self._chk(self.L.m2_set_mixed_threshold(self.h, t))

    def set_jk_families(self, mode):
        """s-s ket families of the direct J/K build (0 off, 1 automatic, 2 all
        four class combinations); applies to the next set_basis."""
        self._chk(self.L.m2_set_jk_families(self.h, int(mode)))

    def set_profiling(self, on=True, counters=False):
        """HIP events around every class launch; counters=True also counts
        computed (primitive) quartets on the device, which slows the kernels."""
        self._chk(self.L.m2_set_profiling(self.h, (2 if counters else 1) if on else 0))

    def reset_profile(self):
        self._chk(self.L.m2_reset_profile(self.h))

    def profile(self):
        """dict combo (X, Y) -> (ms, launches, quartets, prim_quartets,
        computed_prim_quartets); the counts need set_profiling(counters=True)"""
        ms = np.zeros(36)
        la, q, pq, cpq = (np.zeros(36, np.int64) for _ in range(4))
        P = C.POINTER(C.c_longlong)
        self._chk(self.L.m2_get_profile(self.h, _p(ms), la.ctypes.data_as(P), q.ctypes.data_as(P),
                                        pq.ctypes.data_as(P)))
        self._chk(self.L.m2_get_profile_computed(self.h, cpq.ctypes.data_as(P)))
        return {(c // 6, c % 6): (ms[c], int(la[c]), int(q[c]), int(pq[c]), int(cpq[c])) for c in range(36)
                if la[c] > 0}

    def jk_df_local(self, frags, Cocc, sigmainv, Pt, mo_thresh=1e-6, fit_thresh=0.05, r_thresh=15.0,
                    thrint=1e-12):
        """Fock::compute_2body_fock_df_local through m2_jk_df_local.  frags:
        dicts with occ, nbfs, naux, start, auxstart, nshells, ndfshells,
        radius, com; the first call (or the first after df_local_reset)
        builds the fragment blocks and the domains.  Returns J, K."""
        arr = (M2Fragment * len(frags))()
        for k, f in enumerate(frags):
            for key in ("occ", "nbfs", "naux", "start", "auxstart", "nshells", "ndfshells"):
                setattr(arr[k], key, int(f[key]))
            arr[k].radius = float(f["radius"])
            for d in range(3):
                arr[k].com[d] = float(f["com"][d])
        Cf = np.asfortranarray(Cocc, float)
        n, nocc = Cf.shape
        sig = np.asfortranarray(sigmainv, float)
        P = np.ascontiguousarray(Pt, float)
        J, K = np.zeros((n, n)), np.zeros((n, n))
        self._chk(self.L.m2_jk_df_local(self.h, len(frags), arr, mo_thresh, fit_thresh, r_thresh, thrint,
                                        Cf.ctypes.data_as(_dp), nocc, sig.ctypes.data_as(_dp), _p(P), _p(J), _p(K)))
        return J, K

    def df_local_reset(self):
        self._chk(self.L.m2_df_local_reset(self.h))

    def df_local_domains(self, which, nocc, nfrag):
        """which: 0 LMO, 1 AO, 2 fit; list of fragment-centre lists per LMO"""
        cen = np.zeros(nocc * nfrag, np.int32)
        cnt = np.zeros(nocc, np.int32)
        self._chk(self.L.m2_df_local_domains(self.h, int(which), cen.ctypes.data_as(_ip), cnt.ctypes.data_as(_ip)))
        return [list(cen[i * nfrag:i * nfrag + cnt[i]]) for i in range(nocc)]

    def set_jk_launch(self, host_totals=False):
        """grids of the J/K class kernels: bounds from the basis with no host
        round trip (default), or the work-item totals read back first"""
        self._chk(self.L.m2_set_jk_launch(self.h, int(bool(host_totals))))

    def set_jk_graphs(self, on=True):
        """capture each J/K build key once into a HIP graph and replay it"""
        self._chk(self.L.m2_set_jk_graphs(self.h, int(bool(on))))

    def jk_graph_stats(self):
        """(captures, replays, graphs held)"""
        a, b, c = C.c_longlong(), C.c_longlong(), C.c_int()
        self._chk(self.L.m2_jk_graph_stats(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def set_jk_block_tasks(self, X, Y, bt):
        """ket chunks per work item of class combination (X | Y); 0: default"""
        self._chk(self.L.m2_set_jk_block_tasks(self.h, int(X), int(Y), int(bt)))

    def jk_launch(self):
        """(host_totals, {(X, Y): (chunks per item, grid bound)}) of the
        combinations with bra pairs"""
        ht = C.c_int()
        bt = np.zeros(36, np.int32)
        gb = np.zeros(36, np.int32)
        self._chk(self.L.m2_get_jk_launch(self.h, C.byref(ht), bt.ctypes.data_as(_ip), gb.ctypes.data_as(_ip)))
        return bool(ht.value), {(c // 6, c % 6): (int(bt[c]), int(gb[c])) for c in range(36) if bt[c] > 0}

    def jk_autotune(self, dD_ptr, precision=1e-12, mixed=False, repeats=3):
        """times the chunks-per-item candidates on the device density at
        dD_ptr, keeps the fastest per class combination; returns the mean
        build time (ms) before and after"""
        b, a = C.c_double(), C.c_double()
        self._chk(self.L.m2_jk_autotune(self.h, C.c_void_p(dD_ptr), precision, M2_MIXED if mixed else M2_FP64,
                                        int(repeats), C.byref(b), C.byref(a)))
        return b.value, a.value

    def set_prim_screening_fp32(self, relative):
        """primitive screening of the FP32 quartets of a mixed build, relative
        to precision (default 0.1)"""
        self._chk(self.L.m2_set_prim_screening_fp32(self.h, float(relative)))

    def set_prim_screening(self, relative):
        """primitive quartets below relative * precision (bound q_ab q_cd
        max|D|) are skipped; 0 computes all of them (default 1e-6)"""
        self._chk(self.L.m2_set_prim_screening(self.h, float(relative)))

    def df_setup(self):
        self._chk(self.L.m2_df_setup(self.h))

    def jk_df(self, Cocc, fp32=True):
        """J, K over Cocc Cocc^T (Cocc: n x nocc)."""
        Cf = np.asfortranarray(Cocc, float)
        n, nocc = Cf.shape
        J, K = np.zeros((n, n)), np.zeros((n, n))
        self._chk(self.L.m2_jk_df(self.h, Cf.ctypes.data_as(_dp), nocc, int(fp32), _p(J), _p(K)))
        return J, K

    def jk_df_device(self, dC_ptr, nocc, dJ_ptr, dK_ptr, fp32=True):
        self._chk(self.L.m2_jk_df_device(self.h, C.c_void_p(dC_ptr), nocc, int(fp32), C.c_void_p(dJ_ptr),
                                         C.c_void_p(dK_ptr)))

    def df_timing(self):
        out = np.zeros(4)
        self._chk(self.L.m2_df_timing(self.h, _p(out)))
        return out

    def eigh(self, A):
        A = np.asfortranarray(A, float)
        n = A.shape[0]
        w, V = np.zeros(n), np.zeros((n, n), order="F")
        self._chk(self.L.m2_eigh(self.h, n, A.ctypes.data_as(_dp), _p(w), V.ctypes.data_as(_dp)))
        return w, V

    def grad(self, mol, D, W, precision=1e-12, mixed=False):
        """dE/dR (natom x 3) for density D and energy-weighted density W.
        mixed: derivative quartets below the FP32 threshold
        (set_mixed_threshold) in FP32 (m2_grad_ex with M2_MIXED)."""
        Z = np.asarray(mol.Z, float)
        X = np.ascontiguousarray(mol.xyz, float)
        D = np.ascontiguousarray(D, float)
        W = np.ascontiguousarray(W, float)
        g = np.zeros(mol.natom * 3)
        self._chk(self.L.m2_grad_ex(self.h, mol.natom, _p(Z), _p(X), _p(D), _p(W), precision,
                                    M2_MIXED if mixed else 0, _p(g)))
        return g.reshape(mol.natom, 3)

    def rhf(self, mol, basis="cc-pvdz", pure=True, aux=None, gradient=False, guess="soad", **kw):
        """Run the host C++ SCF (molecular::SCF::rhf) on `mol`.  guess:
        "hcore" or "soad" (STO-3G superposition of atomic densities, the
        reference's default).  Keyword options as SCFOptions (defaults: the
        reference's RHF defaults)."""
        if guess not in ("hcore", "soad"):
            raise ValueError("guess must be 'hcore' or 'soad'")
        o = SCFOptions(diis=1, maxdiis=8, maxiter=40, enconverge=1e-7, densconverge=1e-4, precision=1e-12,
                       df=0, mixed=0, mixed_threshold=1e-4, fp64_final=2, guess=1 if guess == "soad" else 0,
                       diis_start=2)
        for k, v in kw.items():
            setattr(o, k, v)
        sh, shell_args = _scf_shells(mol, basis, aux, guess)
        Z = np.asarray(mol.Z, float)
        X = np.ascontiguousarray(mol.xyz, float)
        nbf = sum((2 * s.l + 1) if pure else (s.l + 1) * (s.l + 2) // 2 for s in sh)
        D = np.zeros((nbf, nbf))
        Cm = np.zeros((nbf, nbf))
        eps = np.zeros(nbf)
        log = np.zeros((max(1, o.maxiter), 5))
        g = np.zeros(mol.natom * 3)
        res = SCFResult()
        st = self.L.m2_rhf(self.h, mol.natom, _p(Z), _p(X), mol.charge, mol.multiplicity, int(pure),
                           *shell_args[:-1], C.byref(o), C.byref(res), _p(D), _p(Cm), _p(eps), _p(log),
                           _p(g) if gradient else None)
        if st != 0:
            raise M2Error(st, self.L.m2_rhf_last_error().decode())
        self.nbf = nbf
        self.nsh = len(sh)
        return {"energy": res.energy, "one_e": res.one_e, "two_e": res.two_e, "enuc": res.enuc,
                "iterations": res.iterations, "converged": bool(res.converged), "n_full": res.n_full,
                "n_incremental": res.n_incremental, "n_fp32": res.n_fp32, "jk_seconds": res.jk_seconds,
                "total_seconds": res.total_seconds, "D": D, "C": Cm.T.copy(), "eps": eps,
                "log": log[:res.iterations], "grad": g.reshape(mol.natom, 3) if gradient else None,
                "grad_seconds": res.grad_seconds}

    def rhf_batch(self, mols, basis="cc-pvdz", pure=True, nworkers=8, guess="soad", **kw):
        """Closed-shell SCF::rhf of every molecule in `mols` through
        m2_rhf_batch (guess "hcore") or m2_rhf_batch_min (guess "soad", with
        each molecule's STO-3G shells): nworkers host threads, each with its
        own engine context and HIP stream on this engine's device, take
        molecules largest basis first.  Returns one result dict per molecule
        (energy, iterations, converged, n_full, n_incremental, n_fp32,
        jk_seconds, total_seconds) and the wall seconds of the whole batch."""
        import time
        if guess not in ("hcore", "soad"):
            raise ValueError("guess must be 'hcore' or 'soad'")
        o = SCFOptions(diis=1, maxdiis=8, maxiter=40, enconverge=1e-7, densconverge=1e-4, precision=1e-12,
                       df=0, mixed=0, mixed_threshold=1e-4, fp64_final=2, guess=1 if guess == "soad" else 0,
                       diis_start=2)
        for k, v in kw.items():
            setattr(o, k, v)
        nmol = len(mols)
        natom = np.array([m.natom for m in mols], np.int32)
        charge = np.array([m.charge for m in mols], np.int32)
        Z = np.concatenate([np.asarray(m.Z, float) for m in mols])
        X = np.concatenate([np.ascontiguousarray(m.xyz, float).reshape(-1) for m in mols])
        arrs = [shell_arrays(m.shells(basis)) for m in mols]
        nsh = np.array([len(a[0]) for a in arrs], np.int32)
        l, npr, ex, cf, cen, at = (np.ascontiguousarray(np.concatenate([a[k] for a in arrs])) for k in range(6))
        res = (SCFResult * nmol)()
        ip = lambda a: a.ctypes.data_as(_ip)
        t = time.perf_counter()
        if guess == "soad":
            marr = [shell_arrays(m.shells("sto-3g")) for m in mols]
            nshm = np.array([len(a[0]) for a in marr], np.int32)
            lm, npm, exm, cfm, cenm, atm = (np.ascontiguousarray(np.concatenate([a[k] for a in marr]))
                                            for k in range(6))
            st = self.L.m2_rhf_batch_min(self.L.m2_ctx_device(self.h), int(nworkers), nmol, ip(natom), ip(charge),
                                         _p(Z), _p(X), int(pure), ip(nsh), ip(l), ip(npr), _p(ex), _p(cf), _p(cen),
                                         ip(at), ip(nshm), ip(lm), ip(npm), _p(exm), _p(cfm), _p(cenm), ip(atm),
                                         C.byref(o), res)
        else:
            st = self.L.m2_rhf_batch(self.L.m2_ctx_device(self.h), int(nworkers), nmol, ip(natom), ip(charge),
                                     _p(Z), _p(X), int(pure), ip(nsh), ip(l), ip(npr), _p(ex), _p(cf), _p(cen),
                                     ip(at), C.byref(o), res)
        wall = time.perf_counter() - t
        if st != 0:
            raise M2Error(st, self.L.m2_rhf_last_error().decode())
        out = [{"energy": r.energy, "iterations": r.iterations, "converged": bool(r.converged),
                "n_full": r.n_full, "n_incremental": r.n_incremental, "n_fp32": r.n_fp32,
                "jk_seconds": r.jk_seconds, "total_seconds": r.total_seconds} for r in res]
        return out, wall

    def rhf_hessian_cphf(self, mol, basis="cc-pvdz", pure=True, guess="soad", step=1e-4, cphf_tol=1e-4,
                         cphf_maxiter=30, **kw):
        """Fock::compute_hessian through m2_rhf_hessian_cphf: SCF, then the
        analytic RHF Hessian (3N x 3N, Eh/bohr^2) with the CPHF response
        (the reference's fixed-point solver: cphf_tol 1e-4, 30 steps).
        Returns (H, result dict with energy, converged, cphf_unconverged,
        cphf_steps)."""
        if guess not in ("hcore", "soad"):
            raise ValueError("guess must be 'hcore' or 'soad'")
        o = SCFOptions(diis=1, maxdiis=8, maxiter=40, enconverge=1e-7, densconverge=1e-4, precision=1e-12,
                       df=0, mixed=0, mixed_threshold=1e-4, fp64_final=2, guess=1 if guess == "soad" else 0,
                       diis_start=2)
        for k, v in kw.items():
            setattr(o, k, v)
        _, args = _scf_shells(mol, basis, None, guess)
        n = 3 * mol.natom
        H = np.zeros((n, n))
        res = SCFResult()
        stats = np.zeros(2, np.int32)
        st = self.L.m2_rhf_hessian_cphf(self.h, mol.natom, _p(np.asarray(mol.Z, float)),
                                        _p(np.ascontiguousarray(mol.xyz, float)), mol.charge, mol.multiplicity,
                                        int(pure), *args[0:7], *args[14:21], C.byref(o), float(step),
                                        float(cphf_tol), int(cphf_maxiter), _p(H), C.byref(res),
                                        stats.ctypes.data_as(_ip))
        if st != 0:
            raise M2Error(st, self.L.m2_rhf_last_error().decode())
        return H, {"energy": res.energy, "converged": bool(res.converged), "iterations": res.iterations,
                   "cphf_unconverged": int(stats[0]), "cphf_steps": int(stats[1]),
                   "total_seconds": res.total_seconds}

    def rhf_hessian(self, mol, basis="cc-pvdz", pure=True, aux=None, guess="soad", step=0.01, nworkers=8, **kw):
        """Fock::compute_hessian_numeric through m2_rhf_hessian: the RHF
        nuclear Hessian (3N x 3N, Eh/bohr^2) from central differences (step
        in bohr) of the analytic gradient over the 6N displaced geometries,
        which run as concurrent SCFs on nworkers engine contexts of this
        engine's device.  Options as for rhf().  Returns the Hessian and one
        result dict per displaced run (order: atom, axis, +step, -step)."""
        if guess not in ("hcore", "soad"):
            raise ValueError("guess must be 'hcore' or 'soad'")
        o = SCFOptions(diis=1, maxdiis=8, maxiter=40, enconverge=1e-7, densconverge=1e-4, precision=1e-12,
                       df=0, mixed=0, mixed_threshold=1e-4, fp64_final=2, guess=1 if guess == "soad" else 0,
                       diis_start=2)
        for k, v in kw.items():
            setattr(o, k, v)
        _, shell_args = _scf_shells(mol, basis, aux, guess)
        n = 3 * mol.natom
        H = np.zeros((n, n))
        res = (SCFResult * (2 * n))()
        st = self.L.m2_rhf_hessian(self.L.m2_ctx_device(self.h), int(nworkers), mol.natom,
                                   _p(np.asarray(mol.Z, float)), _p(np.ascontiguousarray(mol.xyz, float)),
                                   mol.charge, mol.multiplicity, int(pure), *shell_args[:-1], C.byref(o),
                                   float(step), _p(H), res)
        if st != 0:
            raise M2Error(st, self.L.m2_rhf_last_error().decode())
        runs = [{"energy": r.energy, "iterations": r.iterations, "converged": bool(r.converged),
                 "total_seconds": r.total_seconds, "grad_seconds": r.grad_seconds} for r in res]
        return H, runs

    def run_deck(self, deck_text, basislib, pure=False, maxres=64):
        """Run the hf / rhf / uhf commands of a reference input deck on this
        engine's device (m2_deck_run).  pure=False: Cartesian d functions, as
        the reference's recorded outputs.  One dict per deck command: command,
        molecule, method ("rhf", "uhf" or None when not an SCF command),
        energy, iterations, converged, seconds."""
        res = (DeckResult * maxres)()
        n = C.c_int(0)
        st = self.L.m2_deck_run(deck_text.encode(), basislib.encode(), self.L.m2_ctx_device(self.h), int(pure), res,
                                maxres, C.byref(n))
        if st != 0:
            raise M2Error(st, self.L.m2_rhf_last_error().decode())
        names = {0: "rhf", 1: "uhf", -1: None}
        return [{"command": r.command, "molecule": r.molecule, "method": names[r.method], "energy": r.energy,
                 "iterations": r.iterations, "converged": bool(r.converged), "seconds": r.seconds}
                for r in res[:min(n.value, maxres)]]

    def jk_uhf(self, Da, Db, multiplier=1.0, precision=1e-12, mixed=False):
        """UnrestrictedFock::formJKdirect: (Ja, Ka, Jb, Kb), screened with
        the block norms of Da + Db."""
        Da = np.ascontiguousarray(Da, float)
        Db = np.ascontiguousarray(Db, float)
        out = [np.zeros_like(Da) for _ in range(4)]
        self._chk(self.L.m2_jk_direct_uhf(self.h, _p(Da), _p(Db), multiplier, precision, M2_MIXED if mixed else 0,
                                          *[_p(x) for x in out]))
        return tuple(out)

    def uhf(self, mol, basis="cc-pvdz", pure=True, aux=None, guess="soad", **kw):
        """Run the host C++ unrestricted SCF (molecular::SCF::uhf)."""
        if guess not in ("hcore", "soad"):
            raise ValueError("guess must be 'hcore' or 'soad'")
        o = SCFOptions(diis=1, maxdiis=8, maxiter=40, enconverge=1e-7, densconverge=1e-4, precision=1e-12,
                       df=0, mixed=0, mixed_threshold=1e-4, fp64_final=2, guess=1 if guess == "soad" else 0,
                       diis_start=2)
        for k, v in kw.items():
            setattr(o, k, v)
        sh = mol.shells(basis)
        l, npr, ex, cf, cen, at = shell_arrays(sh)
        z1 = np.zeros(1, np.int32)
        f1 = np.zeros(1)
        la, npa, exa, cfa, cena, ata = shell_arrays(mol.shells(aux)) if aux else (z1, z1, f1, f1, f1, z1)
        lm, npm, exm, cfm, cenm, atm = shell_arrays(mol.shells("sto-3g")) if guess == "soad" else \
            (z1, z1, f1, f1, f1, z1)
        n = sum((2 * s.l + 1) if pure else (s.l + 1) * (s.l + 2) // 2 for s in sh)
        Da, Db = np.zeros((n, n)), np.zeros((n, n))
        ea, eb = np.zeros(n), np.zeros(n)
        log = np.zeros((max(1, o.maxiter), 5))
        res = SCFResult()
        ip = lambda a: a.ctypes.data_as(_ip)
        st = self.L.m2_uhf(self.h, mol.natom, _p(np.asarray(mol.Z, float)), _p(np.ascontiguousarray(mol.xyz, float)),
                           mol.charge, mol.multiplicity, int(pure),
                           len(sh), ip(l), ip(npr), _p(ex), _p(cf), _p(cen), ip(at),
                           len(la) if aux else 0, ip(la), ip(npa), _p(exa), _p(cfa), _p(cena), ip(ata),
                           len(lm) if guess == "soad" else 0, ip(lm), ip(npm), _p(exm), _p(cfm), _p(cenm), ip(atm),
                           C.byref(o), C.byref(res), _p(Da), _p(Db), _p(ea), _p(eb), _p(log))
        if st != 0:
            raise M2Error(st, self.L.m2_rhf_last_error().decode())
        self.nbf = n
        self.nsh = len(sh)
        nrow = max(0, res.iterations - 1)
        return {"energy": res.energy, "iterations": res.iterations, "converged": bool(res.converged),
                "Da": Da, "Db": Db, "eps_a": ea, "eps_b": eb, "log": log[:nrow], "jk_seconds": res.jk_seconds,
                "total_seconds": res.total_seconds}

    def soad_fock(self, mol, basis="cc-pvdz", pure=True):
        """Two-electron part G of the SOAD guess Fock matrix in `basis`
        (Fock::compute_soad_guess): sum_ls D(ls) [2 (uv|ls) - (ul|vs)] over the
        STO-3G atomic occupations."""
        sh = mol.shells(basis)
        l, npr, ex, cf, cen, at = shell_arrays(sh)
        lm, npm, exm, cfm, cenm, atm = shell_arrays(mol.shells("sto-3g"))
        n = sum((2 * s.l + 1) if pure else (s.l + 1) * (s.l + 2) // 2 for s in sh)
        G = np.zeros((n, n))
        ip = lambda a: a.ctypes.data_as(_ip)
        Z = np.asarray(mol.Z, float)
        X = np.ascontiguousarray(mol.xyz, float)
        st = self.L.m2_soad_fock(self.L.m2_ctx_device(self.h), mol.natom, _p(Z), _p(X), int(pure),
                                 len(sh), ip(l), ip(npr), _p(ex), _p(cf), _p(cen), ip(at),
                                 len(lm), ip(lm), ip(npm), _p(exm), _p(cfm), _p(cenm), ip(atm), _p(G))
        if st != 0:
            raise M2Error(st, self.L.m2_rhf_last_error().decode())
        return G

    def last_stats(self):
        """(FP64 quartets, FP32 quartets, work items) of the last J/K build;
        the quartet counts are -1 (not counted)"""
        a, b, c = C.c_longlong(), C.c_longlong(), C.c_longlong()
        self._chk(self.L.m2_last_stats(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value
