"""Independent McMurchie-Davidson ERIs in numpy/mpmath (test helper).

Hermite expansion coefficients E^{ij}_t and Hermite Coulomb integrals
R^n_{tuv} (Helgaker, Joergensen, Olsen, ch. 9), contracted over primitives
with Cartesian normalisation of x^l exp(-a r^2) and unit-norm contraction
(the libint2 convention).  Shares no code with the C oracle or the HIP
engine; used on tiny shells only.
"""
import math
from functools import lru_cache

import mpmath
import numpy as np


def boys_mp(n, T):
    if T < 1e-15:
        return 1.0 / (2 * n + 1)
    return float(mpmath.mpf(0.5) * mpmath.power(T, -(n + 0.5)) * mpmath.gammainc(n + 0.5, 0, T))


def hermite_E(i, j, t, Qx, a, b):
    p = a + b
    q = a * b / p
    if t < 0 or t > i + j:
        return 0.0
    if i == j == t == 0:
        return math.exp(-q * Qx * Qx)
    if j == 0:
        return (1 / (2 * p)) * hermite_E(i - 1, j, t - 1, Qx, a, b) - (q * Qx / a) * hermite_E(i - 1, j, t, Qx, a, b) \
            + (t + 1) * hermite_E(i - 1, j, t + 1, Qx, a, b)
    return (1 / (2 * p)) * hermite_E(i, j - 1, t - 1, Qx, a, b) + (q * Qx / b) * hermite_E(i, j - 1, t, Qx, a, b) \
        + (t + 1) * hermite_E(i, j - 1, t + 1, Qx, a, b)


def hermite_R(t, u, v, n, p, PCx, PCy, PCz, RPC):
    T = p * RPC * RPC
    if t == u == v == 0:
        return (-2 * p) ** n * boys_mp(n, T)
    if t == u == 0:
        val = 0.0
        if v > 1:
            val += (v - 1) * hermite_R(t, u, v - 2, n + 1, p, PCx, PCy, PCz, RPC)
        return val + PCz * hermite_R(t, u, v - 1, n + 1, p, PCx, PCy, PCz, RPC)
    if t == 0:
        val = 0.0
        if u > 1:
            val += (u - 1) * hermite_R(t, u - 2, v, n + 1, p, PCx, PCy, PCz, RPC)
        return val + PCy * hermite_R(t, u - 1, v, n + 1, p, PCx, PCy, PCz, RPC)
    val = 0.0
    if t > 1:
        val += (t - 1) * hermite_R(t - 2, u, v, n + 1, p, PCx, PCy, PCz, RPC)
    return val + PCx * hermite_R(t - 1, u, v, n + 1, p, PCx, PCy, PCz, RPC)


def prim_eri(a, la, A, b, lb, B, c, lc, C, d, ld, D):
    """primitive (ab|cd) of unnormalised Cartesian Gaussians x^l..."""
    p, q = a + b, c + d
    al = p * q / (p + q)
    P = (a * np.asarray(A) + b * np.asarray(B)) / p
    Q = (c * np.asarray(C) + d * np.asarray(D)) / q
    RPQ = np.linalg.norm(P - Q)
    val = 0.0
    for t in range(la[0] + lb[0] + 1):
        for u in range(la[1] + lb[1] + 1):
            for v in range(la[2] + lb[2] + 1):
                e1 = hermite_E(la[0], lb[0], t, A[0] - B[0], a, b) * hermite_E(la[1], lb[1], u, A[1] - B[1], a, b) * \
                    hermite_E(la[2], lb[2], v, A[2] - B[2], a, b)
                if e1 == 0:
                    continue
                for tau in range(lc[0] + ld[0] + 1):
                    for nu in range(lc[1] + ld[1] + 1):
                        for phi in range(lc[2] + ld[2] + 1):
                            e2 = hermite_E(lc[0], ld[0], tau, C[0] - D[0], c, d) * \
                                hermite_E(lc[1], ld[1], nu, C[1] - D[1], c, d) * \
                                hermite_E(lc[2], ld[2], phi, C[2] - D[2], c, d)
                            if e2 == 0:
                                continue
                            val += e1 * e2 * (-1) ** (tau + nu + phi) * hermite_R(
                                t + tau, u + nu, v + phi, 0, al, P[0] - Q[0], P[1] - Q[1], P[2] - Q[2], RPQ)
    return val * 2 * math.pi ** 2.5 / (p * q * math.sqrt(p + q))


def comps(l):
    return [(x, y, l - x - y) for x in range(l, -1, -1) for y in range(l - x, -1, -1)]


def dfact(n):
    r = 1
    while n > 1:
        r *= n
        n -= 2
    return r


def norm_coefs(l, exps, coefs):
    """x^l-component primitive normalisation, then unit-norm contraction."""
    c = np.array([cf * (2 * a / math.pi) ** 0.75 * (4 * a) ** (l / 2) / math.sqrt(dfact(2 * l - 1))
                  for a, cf in zip(exps, coefs)])
    s = 0.0
    for i, ai in enumerate(exps):
        for j, aj in enumerate(exps):
            g = ai + aj
            s += c[i] * c[j] * dfact(2 * l - 1) / (2 * g) ** l * (math.pi / g) ** 1.5
    return c / math.sqrt(s)


def shell_quartet(sh4):
    """sh4: four (l, exps, raw coefs, center) -> Cartesian contracted block"""
    cs = [norm_coefs(l, e, c) for (l, e, c, _) in sh4]
    blocks = [comps(s[0]) for s in sh4]
    out = np.zeros([len(b) for b in blocks])
    for i, ca in enumerate(blocks[0]):
        for j, cb in enumerate(blocks[1]):
            for k, cc in enumerate(blocks[2]):
                for m, cd in enumerate(blocks[3]):
                    v = 0.0
                    for pa, a in enumerate(sh4[0][1]):
                        for pb, b in enumerate(sh4[1][1]):
                            for pc, c in enumerate(sh4[2][1]):
                                for pd, d in enumerate(sh4[3][1]):
                                    v += cs[0][pa] * cs[1][pb] * cs[2][pc] * cs[3][pd] * prim_eri(
                                        a, ca, sh4[0][3], b, cb, sh4[1][3], c, cc, sh4[2][3], d, cd, sh4[3][3])
                    out[i, j, k, m] = v
    return out
