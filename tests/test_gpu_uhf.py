"""Unrestricted Fock build and SCF on the device against the oracle."""
import numpy as np
import pytest

from molecular2_amd import molecules as M

pytestmark = pytest.mark.gpu


def _oh():
    return M.Molecule([8, 1], np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 0.97]]) * M.TOBOHR, 0, 2, "OH")


def _o2():
    return M.Molecule([8, 8], np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 1.21]]) * M.TOBOHR, 0, 3, "O2")


@pytest.mark.parametrize("pure", [True, False])
def test_jk_uhf_matches_oracle(oracle, engine, pure):
    mol = M.benzene()
    print("case: UHF J/K benzene pure=%s" % pure, flush=True)
    engine.set_basis(mol.shells("cc-pvdz"), pure=pure)
    ob = oracle.Basis(mol.shells("cc-pvdz"), pure=pure)
    n = ob.nbf
    rng = np.random.default_rng(3)
    Ca, Cb = rng.normal(size=(n, 21)), rng.normal(size=(n, 20))
    Da, Db = Ca @ Ca.T / n, Cb @ Cb.T / n
    got = engine.jk_uhf(Da, Db)
    ref = ob.jk_direct_uhf(Da, Db, ob.schwarz(), 1e-12)
    for g, r in zip(got, ref):
        assert np.max(np.abs(g - r)) < 1e-10


@pytest.mark.parametrize("mol_fn,na,nb", [(_oh, 5, 4), (_o2, 9, 7)])
def test_uhf_energy_matches_oracle(oracle, engine, mol_fn, na, nb):
    mol = mol_fn()
    print("case: C++ UHF %s" % mol.name, flush=True)
    r = engine.uhf(mol, pure=True, enconverge=1e-10, densconverge=1e-8, maxiter=80, guess="hcore")
    ob = oracle.Basis(mol.shells("cc-pvdz"), pure=True)
    S, T, V = ob.one_body(mol.Z, mol.xyz)
    Q = ob.schwarz()
    u = oracle.UHF(S, T + V, mol.enuc(), na, nb, lambda Da, Db: ob.jk_direct_uhf(Da, Db, Q, 1e-12),
                   enconverge=1e-10, densconverge=1e-8, maxiter=80)
    E = u.run()
    assert r["converged"] and u.converged
    assert abs(r["energy"] - E) < 1e-8
    assert abs(np.trace(r["Da"] @ S) - na) < 1e-8 and abs(np.trace(r["Db"] @ S) - nb) < 1e-8
    # the same trajectory
    nrow = min(len(r["log"]), len(u.log))
    for k in range(min(nrow, 6)):
        assert abs(r["log"][k, 0] - u.log[k][1]) < 1e-8, k


def test_closed_shell_uhf_equals_rhf_device(engine):
    mol = M.water()
    print("case: C++ UHF closed shell vs RHF", flush=True)
    ru = engine.uhf(mol, pure=True, enconverge=1e-10, densconverge=1e-8, guess="hcore")
    rr = engine.rhf(mol, pure=True, enconverge=1e-10, densconverge=1e-8, guess="hcore")
    assert ru["converged"] and abs(ru["energy"] - rr["energy"]) < 1e-9


def test_df_uhf_closed_shell_equals_df_rhf(engine):
    mol = M.water()
    print("case: C++ DF-UHF closed shell vs DF-RHF", flush=True)
    ru = engine.uhf(mol, pure=True, aux="cc-pvdz-jkfit", df=1, enconverge=1e-10, densconverge=1e-8, guess="hcore")
    rr = engine.rhf(mol, pure=True, aux="cc-pvdz-jkfit", df=1, enconverge=1e-10, densconverge=1e-8, guess="hcore")
    assert ru["converged"] and abs(ru["energy"] - rr["energy"]) < 1e-9
