"""Oracle RHF energies of the small molecules of the seeded random H/C/N/O
batch (molecules.random_batch(256, seed=256), the first six with at most 9
atoms; cc-pVDZ, spherical d) for the batched-SCF parity test.  Core-
Hamiltonian guess, enconverge 1e-10, densconverge 1e-7, maxiter 100 -- the
options the GPU batch test runs with.  Each SCF is the oracle's restatement
of SCF::rhf (src/scf.cpp:29-185) over its formJKdirect restatement at
precision 1e-12.

    python tests/golden/make_batch_oracle.py
"""
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import numpy as np  # noqa: E402

import oracle as O  # noqa: E402
from molecular2_amd import molecules as M  # noqa: E402

KW = dict(enconverge=1e-10, densconverge=1e-7, maxiter=100)


def small_batch():
    return [(i, m) for i, m in enumerate(M.random_batch(256, seed=256)) if m.natom <= 9][:6]


def oracle_rhf(mol):
    ob = O.Basis(mol.shells("cc-pvdz"), pure=True)
    S, T, V = ob.one_body(mol.Z, mol.xyz)
    Q = ob.schwarz()
    nthreads = min(16, os.cpu_count() or 1)
    scf = O.RHF(S, T + V, mol.enuc(), mol.nel, lambda P, mult: ob.jk_direct(P, Q, 1e-12, mult, nthreads),
                guess="hcore", **KW)
    e = scf.run()
    return e, bool(scf.converged), int(scf.iterations), S.shape[0]


def main(out):
    rows = []
    for idx, mol in small_batch():
        t = time.time()
        e, conv, it, n = oracle_rhf(mol)
        print("molecule %d: %d atoms, %d bf, E = %.10f, converged %s in %d iterations (%.1f s)" %
              (idx, mol.natom, n, e, conv, it, time.time() - t), flush=True)
        rows.append({"index": idx, "natom": mol.natom, "nbf": n, "energy": e, "converged": conv, "iterations": it})
    with open(out, "w") as fh:
        json.dump({"batch": "random_batch(256, seed=256), first six with <= 9 atoms", "options": KW,
                   "guess": "hcore", "molecules": rows}, fh, indent=1)


if __name__ == "__main__":
    main(os.path.join(os.path.dirname(os.path.abspath(__file__)), "batch_oracle.json"))
