"""Device RHF gradient (m2_grad / Fock::compute_forces) against the oracle's
compute_forces restatement (itself pinned by finite differences)."""
import numpy as np
import pytest

from molecular2_amd import molecules as M

pytestmark = pytest.mark.gpu


def _oracle_grad(oracle, mol, pure, D, C, eps):
    b = oracle.Basis(mol.shells("cc-pvdz"), pure=pure)
    nocc = mol.nel // 2
    Co = C[:, :nocc]
    W = Co @ np.diag(eps[:nocc]) @ Co.T
    return b.grad(mol, D, W, precision=1e-12), W


@pytest.mark.parametrize("pure", [True, False])
def test_water_gradient(oracle, engine, pure):
    mol = M.Molecule([8, 1, 1], np.array([[0.0, 0.05, -0.1], [1.45, 1.05, 0.2], [-1.5, 1.2, -0.1]]))
    print("case: water gradient pure=%s" % pure, flush=True)
    r = engine.rhf(mol, pure=pure, gradient=True, enconverge=1e-11, densconverge=1e-9, guess="hcore")
    assert r["converged"]
    go, W = _oracle_grad(oracle, mol, pure, r["D"], r["C"], r["eps"])
    g = r["grad"]
    assert np.max(np.abs(g - go)) < 1e-9, np.max(np.abs(g - go))
    assert np.max(np.abs(g.sum(axis=0))) < 1e-8
    # the same through the bare C ABI entry point
    g2 = engine.grad(mol, r["D"], W)
    assert np.max(np.abs(g2 - go)) < 1e-9


def test_benzene_gradient(oracle, engine):
    mol = M.benzene()
    xyz = mol.xyz.copy()
    xyz[0, 0] += 0.05
    xyz[7, 2] += 0.1
    mol = M.Molecule(list(mol.Z), xyz)
    print("case: distorted benzene gradient", flush=True)
    r = engine.rhf(mol, pure=True, gradient=True, enconverge=1e-10, densconverge=1e-7, guess="hcore")
    go, _ = _oracle_grad(oracle, mol, True, r["D"], r["C"], r["eps"])
    assert np.max(np.abs(r["grad"] - go)) < 1e-8, np.max(np.abs(r["grad"] - go))


def test_gradient_finite_difference_on_device(engine):
    """central FD (h = 1e-4 bohr) of the device SCF energy"""
    mol = M.Molecule([8, 1, 1], np.array([[0.0, 0.05, -0.1], [1.45, 1.05, 0.2], [-1.5, 1.2, -0.1]]))
    print("case: device FD gradient", flush=True)
    kw = dict(enconverge=1e-12, densconverge=1e-10, maxiter=80, precision=1e-14)
    g = engine.rhf(mol, gradient=True, **kw, guess="hcore")["grad"]
    h = 1e-4
    for atom, axis in ((0, 1), (2, 0)):
        ep = engine.rhf(mol.displaced(atom, axis, h), **kw, guess="hcore")["energy"]
        em = engine.rhf(mol.displaced(atom, axis, -h), **kw, guess="hcore")["energy"]
        assert abs(g[atom, axis] - (ep - em) / (2 * h)) < 2e-7


def test_mixed_gradient_against_fp64(engine):
    """M2_MIXED: derivative quartets below the FP32 threshold
    (1e-4) in FP32.  Same density: within 1e-7 Eh/bohr of the FP64 gradient;
    through a mixed-precision SCF: within the 1e-6 Eh/bohr target."""
    mol = M.water_cluster(4, seed=4)
    print("case: (H2O)4 mixed-precision gradient", flush=True)
    kw = dict(enconverge=1e-10, densconverge=1e-8, maxiter=80)
    r = engine.rhf(mol, gradient=True, **kw, guess="hcore")
    nocc = mol.nel // 2
    Co = r["C"][:, :nocc]
    W = Co @ np.diag(r["eps"][:nocc]) @ Co.T
    engine.set_mixed_threshold(1e-4)
    g64 = engine.grad(mol, r["D"], W)
    g32 = engine.grad(mol, r["D"], W, mixed=True)
    assert np.max(np.abs(g64 - r["grad"])) < 1e-10
    d = np.max(np.abs(g32 - g64))
    print("same density: max |dg| %.2e" % d, flush=True)
    assert 0 < d < 1e-7
    rm = engine.rhf(mol, gradient=True, mixed=1, **kw, guess="hcore")
    dm = np.max(np.abs(rm["grad"] - r["grad"]))
    print("mixed SCF + gradient: max |dg| %.2e" % dm, flush=True)
    assert dm < 1e-6
