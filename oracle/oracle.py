"""Python side of the CPU oracle.  TEST INFRASTRUCTURE ONLY.

Loads ``oracle/lib/libm2oracle.so`` (built from ``oracle/m2_oracle.c`` by
``oracle/Makefile``) and restates in numpy the reference's SCF driver:

  * SCF::rhf loop, calcE, calcErr, convergence test (src/scf.cpp:29-185);
  * DIISEngine::compute / solve (src/diis.cpp:15-116), including its quirks
    (the damped diagonal is overwritten by the normalised one; DIIS only
    updated from iteration 3 on);
  * Fock::makeFock incremental (Delta-D) logic, addDiis, average, transform,
    diagonalise, makeDens (src/fockbuild.cpp:176-222, src/fock.cpp:114-244);
  * the density-fitted J/K of compute_2body_fock_df (src/fockbuild.cpp:585-660);
  * the fragment-local density-fitted J/K of compute_2body_fock_df_local with
    build_blocked_eris and build_domains (src/fockbuild.cpp:662-1047).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "lib", "libm2oracle.so")
_lib = None

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIBPATH) or \
            os.path.getmtime(LIBPATH) < os.path.getmtime(os.path.join(HERE, "m2_oracle.c")):
        subprocess.check_call(["make", "-s", "-C", HERE] + (["-B"] if force else []))
    return LIBPATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            build()
        L = C.CDLL(LIBPATH)
        L.or_basis_create.restype = C.c_void_p
        L.or_basis_create.argtypes = [C.c_int, _ip, _ip, _dp, _dp, _dp, _ip, C.c_int]
        L.or_basis_free.argtypes = [C.c_void_p]
        L.or_basis_nbf.argtypes = [C.c_void_p]
        L.or_basis_nbf.restype = C.c_int
        L.or_basis_coefs.argtypes = [C.c_void_p, C.c_int, _dp]
        L.or_boys.argtypes = [C.c_int, C.c_double, _dp]
        L.or_sph_coef.argtypes = [C.c_int, C.c_int, C.c_int]
        L.or_sph_coef.restype = C.c_double
        L.or_shell_quartet.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, _dp]
        L.or_schwarz.argtypes = [C.c_void_p, _dp]
        L.or_shellblock_norm.argtypes = [C.c_void_p, _dp, _dp]
        L.or_jk_direct.argtypes = [C.c_void_p, _dp, _dp, C.c_double, C.c_double, _dp, _dp, C.c_int]
        L.or_jk_direct_range_s.argtypes = [C.c_void_p, _dp, _dp, _dp, C.c_double, C.c_double, _dp, _dp, C.c_int,
                                           C.c_int, C.c_int, C.POINTER(C.c_long)]
        L.or_jk_direct_range_s.restype = C.c_long
        L.or_jk_direct.restype = C.c_long
        L.or_jk_direct_range.argtypes = [C.c_void_p, _dp, _dp, C.c_double, C.c_double, _dp, _dp, C.c_int,
                                         C.c_int, C.c_int, C.POINTER(C.c_long)]
        L.or_jk_direct_range.restype = C.c_long
        L.or_eri_full.argtypes = [C.c_void_p, _dp]
        L.or_jk_incore.argtypes = [C.c_int, _dp, _dp, C.c_double, _dp, _dp]
        L.or_one_body.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp, _dp, _dp]
        L.or_eri3_packed.argtypes = [C.c_void_p, C.c_void_p, _dp]
        L.or_eri2.argtypes = [C.c_void_p, _dp]
        L.or_grad_2e.argtypes = [C.c_void_p, _dp, _dp, C.c_double, C.c_int, _dp, C.c_int]
        L.or_grad_1e.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp, _dp, _dp]
        L.or_grad_nuc.argtypes = [C.c_int, _dp, _dp, _dp]
        L.or_num_threads.restype = C.c_int
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_dp)


def boys(mmax: int, T: float) -> np.ndarray:
    F = np.zeros(mmax + 1)
    lib().or_boys(mmax, float(T), _p(F))
    return F


class Basis:
    """A shell list (molecular2_amd.molecules.Shell) normalised by the oracle."""

    def __init__(self, shells, pure: bool = True):
        self.shells = shells
        self.pure = pure
        nsh = len(shells)
        l = np.array([s.l for s in shells], np.int32)
        npr = np.array([len(s.exps) for s in shells], np.int32)
        exps = np.concatenate([np.asarray(s.exps, float) for s in shells])
        coefs = np.concatenate([np.asarray(s.coefs, float) for s in shells])
        cen = np.array([s.center for s in shells], float).reshape(-1)
        at = np.array([s.atom for s in shells], np.int32)
        self._keep = (l, npr, exps, coefs, cen, at)
        self.h = lib().or_basis_create(nsh, l.ctypes.data_as(_ip), npr.ctypes.data_as(_ip), _p(exps),
                                       _p(coefs), _p(cen), at.ctypes.data_as(_ip), int(pure))
        self.nbf = lib().or_basis_nbf(self.h)
        self.nsh = nsh
        sizes = [(2 * s.l + 1) if pure else (s.l + 1) * (s.l + 2) // 2 for s in shells]
        self.sizes = sizes
        self.offsets = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(int)

    def __del__(self):
        try:
            lib().or_basis_free(self.h)
        except Exception:
            pass

    def norm_coefs(self, s: int) -> np.ndarray:
        out = np.zeros(len(self.shells[s].exps))
        lib().or_basis_coefs(self.h, s, _p(out))
        return out

    def quartet(self, s1, s2, s3, s4) -> np.ndarray:
        sz = [self.sizes[s] for s in (s1, s2, s3, s4)]
        out = np.zeros(int(np.prod(sz)))
        lib().or_shell_quartet(self.h, s1, s2, s3, s4, _p(out))
        return out.reshape(sz)

    def schwarz(self) -> np.ndarray:
        Q = np.zeros((self.nsh, self.nsh))
        lib().or_schwarz(self.h, _p(Q))
        return Q

    def shellblock_norm(self, A) -> np.ndarray:
        A = np.ascontiguousarray(A, float)
        out = np.zeros((self.nsh, self.nsh))
        lib().or_shellblock_norm(self.h, _p(A), _p(out))
        return out

    def jk_direct(self, D, Q=None, precision=1e-12, multiplier=1.0, nthreads=0):
        D = np.ascontiguousarray(D, float)
        J = np.zeros_like(D)
        K = np.zeros_like(D)
        Qp = _p(np.ascontiguousarray(Q, float)) if Q is not None else None
        if Q is not None:
            Qa = np.ascontiguousarray(Q, float)
            Qp = _p(Qa)
        n = lib().or_jk_direct(self.h, _p(D), Qp, precision, multiplier, _p(J), _p(K), nthreads)
        self.last_quartets = n
        return J, K

    def jk_direct_uhf(self, Da, Db, Q=None, precision=1e-12, multiplier=1.0, nthreads=0):
        """UnrestrictedFock::formJKdirect (src/fockbuild.cpp:267-364): (Ja, Ka,
        Jb, Kb), every quartet screened with the block norms of Da + Db."""
        Da = np.ascontiguousarray(Da, float)
        Db = np.ascontiguousarray(Db, float)
        Dt = np.ascontiguousarray(Da + Db)
        Qp = _p(np.ascontiguousarray(Q, float)) if Q is not None else None
        if Q is not None:
            Qa = np.ascontiguousarray(Q, float)
            Qp = _p(Qa)
        out = []
        for D in (Da, Db):
            J, K = np.zeros_like(D), np.zeros_like(D)
            lib().or_jk_direct_range_s(self.h, _p(D), _p(Dt), Qp, precision, multiplier, _p(J), _p(K), nthreads,
                                       0, self.nsh, None)
            out += [J, K]
        return tuple(out)

    def jk_direct_range(self, D, Q, s1lo, s1hi, precision=1e-12, nthreads=0):
        """Bounded sample of one build: quartets with s1 in [s1lo, s1hi).
        Returns (quartets, primitive quartets)."""
        D = np.ascontiguousarray(D, float)
        Qa = np.ascontiguousarray(Q, float)
        J, K = np.zeros_like(D), np.zeros_like(D)
        cnt = (C.c_long * 2)(0, 0)
        lib().or_jk_direct_range(self.h, _p(D), _p(Qa), precision, 1.0, _p(J), _p(K), nthreads, s1lo, s1hi, cnt)
        return int(cnt[0]), int(cnt[1])

    def eri_full(self) -> np.ndarray:
        n = self.nbf
        out = np.zeros(n ** 4)
        lib().or_eri_full(self.h, _p(out))
        return out.reshape(n, n, n, n)

    def one_body(self, Z, xyz):
        n = self.nbf
        S, T, V = np.zeros((n, n)), np.zeros((n, n)), np.zeros((n, n))
        Zf = np.asarray(Z, float)
        X = np.ascontiguousarray(xyz, float)
        lib().or_one_body(self.h, len(Zf), _p(Zf), _p(X), _p(S), _p(T), _p(V))
        return S, T, V

    def eri3_packed(self, aux: "Basis") -> np.ndarray:
        n = self.nbf
        out = np.zeros((n * (n + 1) // 2, aux.nbf))
        lib().or_eri3_packed(self.h, aux.h, _p(out))
        return out

    def eri2(self) -> np.ndarray:
        out = np.zeros((self.nbf, self.nbf))
        lib().or_eri2(self.h, _p(out))
        return out

    def grad(self, mol, D, W, Q=None, precision=1e-12, nthreads=0) -> np.ndarray:
        """compute_forces (src/rhfgrad.cpp:15-89): returns dE/dR (natom x 3)."""
        g = np.zeros(mol.natom * 3)
        Zf = np.asarray(mol.Z, float)
        X = np.ascontiguousarray(mol.xyz, float)
        D = np.ascontiguousarray(D, float)
        W = np.ascontiguousarray(W, float)
        lib().or_grad_1e(self.h, mol.natom, _p(Zf), _p(X), _p(D), _p(W), _p(g))
        Qa = np.ascontiguousarray(Q if Q is not None else self.schwarz(), float)
        lib().or_grad_2e(self.h, _p(D), _p(Qa), precision, mol.natom, _p(g), nthreads)
        lib().or_grad_nuc(mol.natom, _p(Zf), _p(X), _p(g))
        return g.reshape(mol.natom, 3)


def jk_incore(eri, P, multiplier=1.0):
    """Fock::formJK (src/fockbuild.cpp:36-51) over the full ERI tensor."""
    n = P.shape[0]
    e = np.ascontiguousarray(eri, float).reshape(-1)
    P = np.ascontiguousarray(P, float)
    J, K = np.zeros((n, n)), np.zeros((n, n))
    lib().or_jk_incore(n, _p(e), _p(P), multiplier, _p(J), _p(K))
    return J, K


# ----------------------------------------------------------------------------
# DIIS (src/diis.cpp:15-116)
# ----------------------------------------------------------------------------

class DIIS:
    def __init__(self, maxdiis=8, use=True, damp=0.02):
        self.max = maxdiis if maxdiis > 0 else 8
        self.use = use
        self.damp = damp
        self.errs = []
        self.lastB = np.zeros((1, 1))

    def _solve(self, B, start):
        size = B.shape[0] - start
        norm = 1.0
        scale = 1.0 + self.damp
        if B[start, start] > 1e-10:
            norm = 1.0 / B[start, start]
        nB = np.zeros((size, size))
        for i in range(size - 1):
            nB[i, i] = B[i + start, i + start] * scale
            for j in range(i, size - 1):   # j == i overwrites the damped diagonal
                nB[i, j] = norm * B[i + start, j + start]
                nB[j, i] = nB[i, j]
            nB[i, size - 1] = nB[size - 1, i] = -1.0
        nB[size - 1, size - 1] = 0.0
        rhs = np.zeros(size)
        rhs[size - 1] = -1.0
        s = np.linalg.svd(nB, compute_uv=False)
        rank = int(np.sum(s > s[0] * size * np.finfo(float).eps))
        try:
            w = np.linalg.solve(nB, rhs)
        except np.linalg.LinAlgError:
            w = np.linalg.lstsq(nB, rhs, rcond=None)[0]
        return np.concatenate([w, [rank]])

    def compute(self, errors):
        weights = np.zeros(0)
        it = len(self.errs) + 1
        if self.use:
            if it > self.max:
                self.errs.pop(0)
            new = np.concatenate([np.asarray(e).ravel() for e in errors])
            self.errs.append(new)
            if it == 1:
                self.lastB[0, 0] = new @ new
            if it > 1:
                lim = it if it < self.max else self.max
                B = np.zeros((lim + 1, lim + 1))
                start = 1 if it > lim else 0
                for i in range(lim - 1):
                    for j in range(i, lim - 1):
                        B[i, j] = self.lastB[i + start, j + start]
                        B[j, i] = B[i, j]
                    B[lim, i] = -1.0
                    B[i, lim] = -1.0
                for i in range(lim):
                    B[lim - 1, i] = self.errs[i] @ self.errs[lim - 1]
                    B[i, lim - 1] = B[lim - 1, i]
                B[lim, lim - 1] = -1.0
                B[lim - 1, lim] = -1.0
                self.lastB = B
                weights = self._solve(B, 0)
                rank = weights[-1]
                start = 1
                while rank < (lim - start + 1) and start < lim:
                    weights = self._solve(B, start)
                    start += 1
                    rank = weights[-1]
                weights = weights[:-2]
        return weights


# ----------------------------------------------------------------------------
# SCF::rhf restated (src/scf.cpp:90-185) with the Fock incremental logic
# ----------------------------------------------------------------------------

def _eigh_sorted(A):
    w, U = np.linalg.eigh(A)
    return w, U


def sto3g_occupations(Z: int):
    """Average STO-3G AO occupations of the neutral ground-state atom (NIST
    configurations, electrons smeared evenly over each subshell), in STO-3G
    shell order (sto3g_ao_occupation_vector, inc/sto3g_atomic_density.hpp:56-110)."""
    if not 1 <= Z <= 53:
        raise ValueError("STO-3G occupations are defined for 1 <= Z <= 53")
    nao = 1 if Z <= 2 else 5 if Z <= 10 else 9 if Z <= 18 else 13 if Z <= 20 else 18 if Z <= 36 else \
        22 if Z <= 38 else 27
    occ = []

    def fill(size, ne):
        n = min(ne, 2 * size)
        occ.extend([n / size] * size)
        return ne - n

    ne = fill(1, Z)                                        # 1s
    if Z > 2:
        ne = fill(3, fill(1, ne))                          # 2s 2p
    if Z > 10:
        ne = fill(3, fill(1, ne))                          # 3s 3p
    if 18 < Z <= 36:
        n4s = 1 if Z in (19, 24, 29) else 2
        ne -= n4s
        fill(1, n4s)                                       # 4s
        n4p = min(6, Z - 30 if Z > 30 else 0)
        ne -= n4p
        fill(3, n4p)                                       # 4p
        fill(5, ne)                                        # 3d
    if Z > 36:
        ne = fill(3, fill(1, ne))                          # 4s 4p
        n5s = 1 if Z in (37, 41, 42, 44, 45, 47) else 2
        ne -= n5s
        fill(1, n5s)                                       # 5s
        n5p = min(6, Z - 48 if Z > 48 else 0)
        ne -= n5p
        fill(3, n5p)                                       # 5p
        ne = fill(5, ne)                                   # 3d
        fill(5, ne)                                        # 4d
    # K and Ca: the 3d entries fall outside the 13 STO-3G AOs and are dropped
    return np.array(occ[:nao])


def soad_two_electron(obs_shells, min_shells, Z, pure=True):
    """G(uv) = sum_ls D(ls) [2 (uv|ls) - (ul|vs)] with D = diag(occ) / 2 over
    the minimal basis: compute_2body_fock_general (src/fockbuild.cpp:477-583)
    with the shell-diagonal SOAD density of Fock::compute_soad_guess
    (src/fock.cpp:174-225).  Evaluated as J/K over the orbital basis followed
    by the minimal one, with the density non-zero on the minimal block only,
    and no screening (the reference runs at machine-epsilon precision)."""
    B = Basis(list(obs_shells) + list(min_shells), pure)
    n = sum((2 * s.l + 1) if pure else (s.l + 1) * (s.l + 2) // 2 for s in obs_shells)
    dmin = np.concatenate([0.5 * sto3g_occupations(int(z)) for z in Z])
    if n + len(dmin) != B.nbf:
        raise ValueError("minimal basis does not match the STO-3G occupations")
    D = np.zeros((B.nbf, B.nbf))
    D[np.arange(n, B.nbf), np.arange(n, B.nbf)] = dmin
    J, K = B.jk_direct(D, precision=0.0)
    return 2.0 * J[:n, :n] - K[:n, :n]


class RHF:
    """Closed-shell SCF restatement.  `jk(P, multiplier)` must return (J, K)
    for the direct/in-core paths (jkints = J - K/2) or, with df=True,
    `jk(Cocc)` returns (J, K) over C C^T (jkints = 2J - K)."""

    def __init__(self, S, H, enuc, nel, jk, df=False, maxdiis=8, diis=True, maxiter=40,
                 enconverge=1e-7, densconverge=1e-4, guess="hcore", soad_G=None, diis_start=2):
        if nel % 2:
            raise ValueError("RHF: Molecule has an odd number of electrons.")
        if guess not in ("hcore", "soad") or (guess == "soad" and soad_G is None):
            raise ValueError("guess: 'hcore', or 'soad' with soad_G (soad_two_electron)")
        self.guess = guess
        self.soad_G = soad_G
        # DIIS extrapolation from loop iteration diis_start + 1 on: 2 in the
        # current source (src/scf.cpp:148,154); the build that wrote
        # test/h2otest.out_1 and test/benzene.out_1 extrapolated from
        # iteration 1 (diis_start = 0)
        self.diis_start = diis_start
        self.S, self.H, self.enuc, self.nel = S, H, enuc, nel
        self.nocc = nel // 2
        self.n = S.shape[0]
        self.jk = jk
        self.df = df
        self.diis_on = diis
        self.MAX = maxdiis
        self.maxiter = maxiter
        self.econv = enconverge
        self.dconv = densconverge
        lam, U = _eigh_sorted(S)
        self.X = U @ np.diag(1.0 / np.sqrt(lam)) @ U.T
        # Fock incremental state (src/fock.cpp:57-60)
        self.reset_incremental = False
        self.started_incremental = False
        self.next_reset = 0.0
        self.rms_error = 1.0
        self.incremental_threshold = 1e-5
        self.last_reset = 0
        self.iter = 0
        self.focks = []
        self.dens = None
        self.dens_diff = None
        self.log = []
        self.jk_calls = []

    # Fock methods ---------------------------------------------------------
    def diagonalise(self):
        eps, U = _eigh_sorted(self.fockm)
        self.CP = self.X @ U
        self.eps = eps

    def makeDens(self):
        if self.dens is not None:
            self.dens_diff = -self.dens
        Co = self.CP[:, :self.nocc]
        self.dens = 2.0 * Co @ Co.T
        if self.dens_diff is None:
            self.dens_diff = self.dens.copy()
        else:
            self.dens_diff = self.dens_diff + self.dens

    def makeJK(self, P, multiplier=1.0):
        if self.df:
            J, K = self.jk(P)
            self.jints, self.kints = J, K
            self.jkints = multiplier * (2.0 * J - K)
        else:
            J, K = self.jk(P, multiplier)
            self.jints, self.kints = J, K
            self.jkints = J - 0.5 * K

    def makeFock(self):
        if not self.df:
            P = None
            self.focka = self.fockinc.copy()
            if (not self.started_incremental) and self.rms_error < self.incremental_threshold:
                self.started_incremental = True
                self.reset_incremental = False
                self.next_reset = self.rms_error / 1e1
                self.last_reset = self.iter - 1
            full = self.reset_incremental or not self.started_incremental
            if full:
                self.focka = self.H.copy()
                self.dens_diff = self.dens.copy()
            if self.reset_incremental and self.started_incremental:
                self.reset_incremental = False
                self.last_reset = self.iter
                self.next_reset = self.rms_error / 1e1
            self.jk_calls.append("full" if full else "incremental")
            self.makeJK(self.dens_diff, 1.0)
        else:
            self.focka = self.H.copy()
            self.makeJK(self.CP[:, :self.nocc], 1.0)
        self.focka = self.focka + self.jkints
        if not self.df:
            self.fockinc = self.focka.copy()
        if self.diis_on:
            if self.iter >= self.MAX:
                self.focks.pop(0)
            self.focks.append(self.focka.copy())
            self.iter += 1

    def average(self, w):
        if self.diis_on and self.iter > 2:
            off = len(self.focks) - len(w)
            F = np.zeros_like(self.focka)
            for i in range(off, len(self.focks)):
                F = F + w[i - off] * self.focks[i]
            self.focka = F

    def transform(self, first=False):
        # Fock::transform (src/fock.cpp:156-172)
        if first and self.guess == "hcore":
            self.fockm = self.X.T @ self.H @ self.X
            self.fockinc = self.H.copy()
        elif first:
            # Fock::compute_soad_guess (src/fock.cpp:174-225)
            self.focka = self.H + self.soad_G
            self.fockinc = self.focka.copy()
            self.fockm = self.X.T @ (self.focka @ self.X)
        else:
            self.fockm = self.X.T @ (self.focka @ self.X)

    # SCF methods ----------------------------------------------------------
    def calcE(self):
        self.last_energy = getattr(self, "energy", 0.0)
        one = 0.5 * np.trace(self.dens @ self.H)
        two = 0.5 * np.trace(self.dens @ self.focka) - one
        one *= 2.0
        self.one_E, self.two_E = one, two
        self.energy = one + two + self.enuc

    def calcErr(self):
        F, D, S, X = self.focka, self.dens, self.S, self.X
        e = F @ (D @ S) - S @ (D @ F)
        e = X.T @ e @ X
        self.error = np.linalg.norm(e)
        return e.ravel(order="F")

    def run(self):
        diis = DIIS(self.MAX, self.diis_on)
        self.energy = 0.0
        self.transform(True)
        self.diagonalise()
        self.makeDens()
        old = self.dens.copy()
        self.makeFock()
        errs = [self.calcErr()]
        w = diis.compute(errs)
        errs = []
        self.calcE()
        self.log.append((0, self.energy, 0.0, 0.0))
        self.average(w)
        self.transform(False)
        it = 1
        converged = False
        while not converged and it < self.maxiter:
            self.diagonalise()
            self.makeDens()
            dd = np.linalg.norm(self.dens - old)
            old = self.dens.copy()
            self.makeFock()
            errs.append(self.calcErr())
            self.rms_error = np.linalg.norm(errs[0]) / (old.shape[0] * old.shape[1])
            if self.rms_error < self.next_reset or it - self.last_reset > 8:
                self.reset_incremental = True
            if it > self.diis_start:
                w = diis.compute(errs)
            errs = []
            self.calcE()
            delta = abs(self.energy - self.last_energy)
            self.log.append((it, self.energy, delta, dd))
            if it > self.diis_start:
                self.average(w)
            self.transform(False)
            converged = dd < self.dconv and delta < self.econv
            it += 1
        self.diagonalise()
        self.converged = converged
        self.iterations = it
        return self.energy


# ----------------------------------------------------------------------------
# DF-JK restated (src/fockbuild.cpp:585-660, src/eris.cpp:89-156,304-356)
# ----------------------------------------------------------------------------

class UHF:
    """Unrestricted SCF restatement: SCF::uhf_internal (src/scf.cpp:193-272)
    with UnrestrictedFock (src/fock.cpp:278-361, src/fockbuild.cpp:412-445).
    `jk(Da, Db)` returns (Ja, Ka, Jb, Kb) (F_s = H + Ja + Jb - K_s).  DIIS runs
    from the first iteration on the concatenated alpha and beta error
    vectors; the energy is taken with the DIIS-averaged Fock matrices."""

    def __init__(self, S, H, enuc, nalpha, nbeta, jk, maxdiis=8, diis=True, maxiter=40,
                 enconverge=1e-7, densconverge=1e-4, guess="hcore", soad_G=None):
        if guess not in ("hcore", "soad") or (guess == "soad" and soad_G is None):
            raise ValueError("guess: 'hcore', or 'soad' with soad_G (soad_two_electron)")
        self.S, self.H, self.enuc = S, H, enuc
        self.na, self.nb = nalpha, nbeta
        self.jk = jk
        self.diis_on = diis
        self.MAX = maxdiis
        self.maxiter = maxiter
        self.econv, self.dconv = enconverge, densconverge
        self.guess, self.soad_G = guess, soad_G
        lam, U = _eigh_sorted(S)
        self.X = U @ np.diag(1.0 / np.sqrt(lam)) @ U.T
        self.iter = 0
        self.afocks, self.bfocks = [], []
        self.log = []

    def _calcE(self, D, F):
        one = 0.5 * np.trace(D @ self.H)
        two = 0.5 * np.trace(D @ F) - one
        return 2.0 * one + two

    def _calcErr(self, F, D):
        e = self.X.T @ (F @ (D @ self.S) - self.S @ (D @ F)) @ self.X
        return e.ravel(order="F")

    def _diagonalise(self):
        ea, Ua = _eigh_sorted(self.Fa_mo)
        eb, Ub = _eigh_sorted(self.Fb_mo)
        self.Ca, self.Cb = self.X @ Ua, self.X @ Ub
        self.eps_a, self.eps_b = ea, eb

    def run(self):
        diis = DIIS(self.MAX, self.diis_on)
        F0 = self.H if self.guess == "hcore" else self.H + self.soad_G
        self.Fa_mo = self.Fb_mo = self.X.T @ (F0 @ self.X)
        self._diagonalise()
        n = self.S.shape[0]
        old_a, old_b = np.zeros((n, n)), np.zeros((n, n))
        self.energy = 0.0
        it, converged = 1, False
        while not converged and it < self.maxiter:
            Ca, Cb = self.Ca[:, :self.na], self.Cb[:, :self.nb]
            self.Da, self.Db = Ca @ Ca.T, Cb @ Cb.T
            Ja, Ka, Jb, Kb = self.jk(self.Da, self.Db)
            self.Fa = self.H + Ja + Jb - Ka
            self.Fb = self.H + Ja + Jb - Kb
            if self.diis_on:
                if self.iter >= self.MAX:
                    self.afocks.pop(0)
                    self.bfocks.pop(0)
                self.afocks.append(self.Fa.copy())
                self.bfocks.append(self.Fb.copy())
                self.iter += 1
            w = diis.compute([self._calcErr(self.Fa, self.Da), self._calcErr(self.Fb, self.Db)])
            if self.diis_on and self.iter > 2:
                off = len(self.afocks) - len(w)
                self.Fa = sum(w[i - off] * self.afocks[i] for i in range(off, len(self.afocks)))
                self.Fb = sum(w[i - off] * self.bfocks[i] for i in range(off, len(self.bfocks)))
            ea, eb = self._calcE(self.Da, self.Fa), self._calcE(self.Db, self.Fb)
            self.Fa_mo = self.X.T @ (self.Fa @ self.X)
            self.Fb_mo = self.X.T @ (self.Fb @ self.X)
            self._diagonalise()
            last = self.energy
            self.energy = ea + eb + self.enuc
            delta = abs(self.energy - last)
            dist = np.linalg.norm(self.Da + self.Db - old_a - old_b)
            old_a, old_b = self.Da, self.Db
            self.log.append((it, self.energy, delta, dist))
            converged = delta < self.econv and dist < self.dconv
            it += 1
        self._diagonalise()
        self.converged = converged
        self.iterations = it
        return self.energy


class DFJK:
    def __init__(self, obs: Basis, aux: Basis):
        n = obs.nbf
        xyK = obs.eri3_packed(aux)               # (n(n+1)/2) x naux
        V = aux.eri2()
        L = np.linalg.cholesky(V)                # V = L L^T
        Linv = np.linalg.inv(L)
        self.B = xyK @ Linv.T                    # xyK(xy,K) = row . (L^-T)[:,K]
        self.n = n
        self.naux = aux.nbf
        iu = np.tril_indices(n)
        self._idx = np.zeros((n, n), int)
        self._idx[iu] = iu[0] * (iu[0] + 1) // 2 + iu[1]
        self._idx = np.maximum(self._idx, self._idx.T)

    def full_B(self):
        return self.B[self._idx]                 # n x n x naux

    def __call__(self, Cocc):
        Bf = self.full_B()
        xiK = np.einsum("xyK,yi->xiK", Bf, Cocc)
        K = np.einsum("xiK,yiK->xy", xiK, xiK)
        gam = np.einsum("xiK,xi->K", xiK, Cocc)
        J = Bf @ gam
        return J, K


class LocalDFJK:
    """Fragment-local density-fitted J/K: Fock::build_blocked_eris
    (src/fockbuild.cpp:985-1047), Fock::build_domains (:662-792) and
    Fock::compute_2body_fock_df_local (:794-984), restated over the full
    raw tensor (xy|K) with the screened fragment blocks zeroed.

    frags: list of dicts with start, nbfs, nshells (orbital functions and
    shells, contiguous), auxstart, naux, ndfshells (auxiliary), occ, com (3,),
    radius; mo_thresh, fit_thresh, r_thresh are the first fragment's (the
    reference reads finfo[0]); thrint: the block screening threshold.
    The first call (Cocc, sigmainv, Pt) builds the blocks and the domains,
    as the reference does when its xyK is still empty."""

    def __init__(self, obs, aux, frags, Q, S, mo_thresh=1e-6, fit_thresh=0.05, r_thresh=15.0, thrint=1e-12):
        n = obs.nbf
        xyK = obs.eri3_packed(aux)
        iu = np.tril_indices(n)
        idx = np.zeros((n, n), int)
        idx[iu] = iu[0] * (iu[0] + 1) // 2 + iu[1]
        idx = np.maximum(idx, idx.T)
        self.M = xyK[idx]                         # n x n x naux, raw (xy|K)
        self.V = aux.eri2()
        self.Q = np.asarray(Q, float)             # shell-pair Schwarz (prescreen)
        self.S = np.asarray(S, float)
        self.frags = frags
        self.mo_thresh, self.fit_thresh, self.r_thresh, self.thrint = mo_thresh, fit_thresh, r_thresh, thrint
        self.n, self.naux = n, aux.nbf
        self.ready = False

    def _blocks(self, Pt):
        """build_blocked_eris: which (f1 >= f2, fk) blocks are kept."""
        F = self.frags
        nf = len(F)
        keep = np.zeros((nf, nf, nf), bool)
        sh0 = np.cumsum([0] + [f["nshells"] for f in F])
        for f1 in range(nf):
            a = F[f1]
            for f2 in range(f1 + 1):
                b = F[f2]
                infnorm = np.max(np.abs(self.Q[sh0[f1]:sh0[f1 + 1], sh0[f2]:sh0[f2 + 1]]))
                ptnorm = np.max(np.abs(Pt[a["start"]:a["start"] + a["nbfs"], b["start"]:b["start"] + b["nbfs"]]))
                for fk in range(nf):
                    k = F[fk]
                    Vb = self.V[k["auxstart"]:k["auxstart"] + k["naux"], k["auxstart"]:k["auxstart"] + k["naux"]]
                    infjpq = np.sqrt(np.max(np.abs(Vb)))
                    rxyz = 0.5 * (np.linalg.norm(np.subtract(a["com"], k["com"]))
                                  + np.linalg.norm(np.subtract(b["com"], k["com"])))
                    est = 1.0 if rxyz < 1.0 else 2.0 * infnorm * ptnorm * infjpq / rxyz
                    if est > self.thrint:
                        keep[f1, f2, fk] = keep[f2, f1, fk] = True
        return keep

    def _domains(self, C):
        """build_domains on C = Cocc sigmainv^(1/2) (one LMO per column,
        ordered by fragment)."""
        F = self.frags
        nf = len(F)
        lmo, ao, fit = [], [], []
        i = 0
        for B in range(nf):
            for _ in range(F[B]["occ"]):
                cen, notc, off = [], [], 0
                for A in range(nf):
                    if A == B:
                        add = True
                    else:
                        # the reference sums over occ(A) rows from the
                        # fragment's first function (src/fockbuild.cpp:675-677)
                        add = float(np.sum(C[off:off + F[A]["occ"], i] ** 2)) > self.mo_thresh
                    (cen if add else notc).append(A)
                    off += F[A]["nbfs"]
                aoc = list(cen)
                for A in cen:
                    newnc = []
                    for Cc in notc:
                        sep = np.linalg.norm(np.subtract(F[Cc]["com"], F[A]["com"]))
                        if F[Cc]["radius"] > sep:
                            aoc.append(Cc)
                        else:
                            newnc.append(Cc)
                    notc = newnc
                lmo.append(cen)
                ao.append(aoc)
                i += 1
        w, U = np.linalg.eigh(self.S)
        SC = (U * np.sqrt(w)) @ U.T @ C
        i = 0
        for f1 in range(nf):
            for _ in range(F[f1]["occ"]):
                d = []
                for f2 in range(nf):
                    sep = np.linalg.norm(np.subtract(F[f2]["com"], F[f1]["com"]))
                    s0 = F[f2]["start"]
                    INi = float(np.sum(SC[s0:s0 + F[f2]["nbfs"], i] ** 2))
                    if INi > self.fit_thresh or sep < self.r_thresh:
                        d.append(f2)
                fit.append(d)
                i += 1
        return lmo, ao, fit

    def __call__(self, Cocc, sigmainv, Pt):
        F = self.frags
        n, naux = self.n, self.naux
        w, U = np.linalg.eigh(sigmainv)
        C = np.asarray(Cocc, float) @ ((U * np.sqrt(w)) @ U.T)
        bf_frag = np.concatenate([[k] * f["nbfs"] for k, f in enumerate(F)])
        aux_frag = np.concatenate([[k] * f["naux"] for k, f in enumerate(F)])
        if not self.ready:
            self.keep = self._blocks(Pt)
            self.lmo, self.ao, self.fit = self._domains(C)
            self.Mk = self.M * self.keep[bf_frag[:, None, None], bf_frag[None, :, None], aux_frag[None, None, :]]
            self.ready = True
        Mk = self.Mk
        K = np.zeros((n, n))
        def idx(frs, first, size):   # the functions of fragments frs, in domain order
            return np.concatenate([np.arange(F[a][first], F[a][first] + F[a][size]) for a in frs]
                                  + [np.zeros(0, int)])

        for i in range(C.shape[1]):
            aoi = idx(self.ao[i], "start", "nbfs")
            moi = idx(self.lmo[i], "start", "nbfs")
            fiti = idx(self.fit[i], "auxstart", "naux")
            if fiti.size == 0:      # an empty fit domain adds no exchange
                continue
            uA = np.einsum("uvA,v->uA", Mk[np.ix_(aoi, moi, fiti)], C[moi, i])
            L = np.linalg.cholesky(self.V[np.ix_(fiti, fiti)])
            G = np.linalg.inv(L).T
            Y = uA @ G
            K[np.ix_(aoi, aoi)] += Y @ Y.T
        Pk = np.zeros(naux)
        for f1 in range(len(F)):
            for f2 in range(f1 + 1):
                a, b = F[f1], F[f2]
                blk = Mk[a["start"]:a["start"] + a["nbfs"], b["start"]:b["start"] + b["nbfs"], :]
                Pb = Pt[a["start"]:a["start"] + a["nbfs"], b["start"]:b["start"] + b["nbfs"]]
                Pk += (0.5 if f1 == f2 else 1.0) * np.einsum("xyK,xy->K", blk, Pb)
        Pk = np.linalg.solve(self.V, Pk)
        J = 4.0 * np.einsum("xyK,K->xy", Mk, Pk)
        return J, K


def hessian_numeric(gradient_at, xyz, h=0.01):
    """Fock::compute_hessian_numeric (src/rhfgrad.cpp:362-416) over any
    gradient: gradient_at(xyz) -> (natom, 3) dE/dR at positions xyz (natom, 3,
    bohr).  Row 3i+x is [g(R + h e_ix) - g(R - h e_ix)] / 2h flattened over
    (atom, axis) (:375-405); entries of magnitude <= 1e-12 become zero (:405)
    and the lower triangle is then copied from the upper one (:413-415)."""
    xyz = np.asarray(xyz, float)
    n = xyz.size
    H = np.zeros((n, n))
    for r in range(n):
        p = xyz.copy().reshape(-1)
        p[r] += h
        gp = np.asarray(gradient_at(p.reshape(xyz.shape)), float).reshape(-1)
        p[r] -= 2 * h
        gm = np.asarray(gradient_at(p.reshape(xyz.shape)), float).reshape(-1)
        row = (gp - gm) / (2 * h)
        H[r] = np.where(np.abs(row) > 1e-12, row, 0.0)
    low = np.tril_indices(n, -1)
    H[low] = H.T[low]
    return H
