"""Write tests/golden/h2o20_jk.npz: the oracle's formJKdirect restatement
(src/fockbuild.cpp:64-164) on the (H2O)20 cc-pVDZ cluster (seed 20, 480
spherical basis functions), at the density the GPU parity test rebuilds.

The density is the core-Hamiltonian guess 2 C_occ C_occ^T (100 lowest
orbitals of S^-1/2 H S^-1/2 with the oracle's one-electron integrals),
symmetrised and rounded to FP32 so that it is stored exactly: the fixture
holds it as float32 (upper triangle), and J and K (upper triangles, FP64) at
precision 1e-12 with the oracle's Schwarz bounds, plus the number of
quartets the oracle computed.  Takes a few minutes on 8 cores.

Usage:  python tests/golden/make_h2o20_jk.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from molecular2_amd import molecules as M  # noqa: E402
import oracle  # noqa: E402


def core_density(ob, mol):
    S, T, V = ob.one_body(mol.Z, mol.xyz)
    w, U = np.linalg.eigh(S)
    X = U @ np.diag(w ** -0.5) @ U.T
    e, Cp = np.linalg.eigh(X @ (T + V) @ X)
    Co = (X @ Cp)[:, : mol.nel // 2]
    D = 2.0 * Co @ Co.T
    D = 0.5 * (D + D.T)
    return D.astype(np.float32)


def main():
    mol = M.water_cluster(20, seed=20)
    ob = oracle.Basis(mol.shells("cc-pvdz"), pure=True)
    n = ob.nbf
    D32 = core_density(ob, mol)
    D = D32.astype(np.float64)
    Q = ob.schwarz()
    t = time.time()
    J, K = ob.jk_direct(D, Q, 1e-12, 1.0)
    print("oracle (H2O)20 J/K: %d quartets in %.1f s" % (ob.last_quartets, time.time() - t))
    iu = np.triu_indices(n)
    np.savez_compressed(os.path.join(HERE, "h2o20_jk.npz"), nbf=n, precision=1e-12,
                        quartets=ob.last_quartets, D32=D32[iu], J=J[iu], K=K[iu])


if __name__ == "__main__":
    main()
