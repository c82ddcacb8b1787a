"""Device numerical RHF Hessian (m2_rhf_hessian / Fock::compute_hessian_numeric,
src/rhfgrad.cpp:362-416) against the oracle's restatement over the oracle's
gradient, and the translational / rotational identities on a larger case."""
import numpy as np
import pytest

from molecular2_amd import molecules as M

pytestmark = pytest.mark.gpu

WATER = np.array([[0.0, 0.05, -0.1], [1.45, 1.05, 0.2], [-1.5, 1.2, -0.1]])
TIGHT = dict(enconverge=1e-11, densconverge=1e-9, maxiter=80, precision=1e-14)


def _oracle_grad(oracle, mol, pure):
    b = oracle.Basis(mol.shells("cc-pvdz"), pure=pure)
    S, T, V = b.one_body(mol.Z, mol.xyz)
    Q = b.schwarz()
    r = oracle.RHF(S, T + V, mol.enuc(), mol.nel, lambda P, m: b.jk_direct(P, Q, 1e-14, m),
                   enconverge=1e-12, densconverge=1e-10, maxiter=80)
    r.run()
    assert r.converged
    nocc = mol.nel // 2
    Co = r.CP[:, :nocc]
    return b.grad(mol, r.dens, Co @ np.diag(r.eps[:nocc]) @ Co.T, precision=1e-14)


@pytest.mark.parametrize("pure", [True, False])
def test_water_hessian_matches_oracle(oracle, engine, pure):
    mol = M.Molecule([8, 1, 1], WATER)
    print("case: water cc-pVDZ numerical Hessian pure=%s" % pure, flush=True)
    H, runs = engine.rhf_hessian(mol, pure=pure, **TIGHT, guess="hcore")
    assert len(runs) == 18 and all(r["converged"] for r in runs)
    Ho = oracle.hessian_numeric(lambda x: _oracle_grad(oracle, M.Molecule([8, 1, 1], x), pure), WATER)
    assert np.max(np.abs(H - Ho)) < 2e-6, np.max(np.abs(H - Ho))
    assert np.array_equal(H, H.T)


def _rotation_residue(H, xyz, g):
    """max over the three rotation generators G of |H (G x) - G g| / max|G x|,
    about the centroid (the translation identity makes the centre free); the
    central difference's truncation leaves about 1e-4 per unit of |G x|"""
    x = xyz - xyz.mean(axis=0)
    out = 0.0
    for axis in range(3):
        G = np.zeros((3, 3))
        i, j = [k for k in range(3) if k != axis]
        G[i, j], G[j, i] = -1.0, 1.0
        v = (x @ G.T).reshape(-1)
        out = max(out, np.max(np.abs(H @ v - (g @ G.T).reshape(-1))) / max(1.0, np.max(np.abs(v))))
    return out


def test_water_dimer_hessian_identities(engine):
    """(H2O)2 cc-pVDZ, SOAD guess, 36 displaced SCFs on 4 contexts: the
    Hessian is translation invariant and rotation covariant with the device
    gradient at the undisplaced geometry; one worker gives the same matrix."""
    mol = M.water_cluster(2, seed=2)
    print("case: (H2O)2 numerical Hessian, soad guess", flush=True)
    H, runs = engine.rhf_hessian(mol, guess="soad", nworkers=4, **TIGHT)
    assert len(runs) == 36 and all(r["converged"] for r in runs)
    g = engine.rhf(mol, guess="soad", gradient=True, **TIGHT)["grad"]
    for a in range(3):
        assert np.max(np.abs(H[:, a::3].sum(axis=1))) < 1e-4
    assert _rotation_residue(H, mol.xyz, g) < 2e-4
    H1, _ = engine.rhf_hessian(mol, guess="soad", nworkers=1, **TIGHT)
    assert np.max(np.abs(H1 - H)) < 1e-7
