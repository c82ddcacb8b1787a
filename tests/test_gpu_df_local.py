"""Fragment-local DF-JK on the device (m2_jk_df_local) against the oracle's
restatement of compute_2body_fock_df_local, build_blocked_eris and
build_domains (src/fockbuild.cpp:662-1047), and against the device's own
global DF-JK where the local fit must equal it.

  * (H2O)3 with the third monomer moved 4 A further out, the reference's
    default thresholds: the LMO, AO and fit domains equal the oracle's and
    J, K agree to 1e-10; a second call with other orbitals reuses the
    domains (as the reference does once its blocks exist) and agrees with
    the oracle's second call;
  * every block and fragment in every domain: J = 2 J_DF and K = K_DF of
    m2_jk_df (FP64) at Cocc sigmainv^(1/2);
  * a reset rebuilds the domains under new thresholds.
"""
import numpy as np
import pytest

from molecular2_amd import molecules as M
from test_oracle_df_local import local_orbitals

pytestmark = pytest.mark.gpu

GROUPS = [[0, 1, 2], [3, 4, 5], [6, 7, 8]]


def _trimer():
    mol = M.water_cluster(3, seed=3)
    xyz = mol.xyz.copy()
    xyz[6:9, 2] += 4.0 * M.TOBOHR
    return M.Molecule(list(mol.Z), xyz, name="(H2O)3 stretched")


@pytest.fixture(scope="module")
def system(oracle, engine):
    mol = _trimer()
    frags = M.fragment_info(mol, GROUPS)
    Cocc, sig, Pt, S = local_orbitals(oracle, mol, GROUPS)
    ob = oracle.Basis(mol.shells("cc-pvdz"), pure=True)
    ab = oracle.Basis(mol.shells("cc-pvdz-jkfit"), pure=True)
    return dict(mol=mol, frags=frags, Cocc=Cocc, sig=sig, Pt=Pt, S=S, ob=ob, ab=ab)


def _set(engine, mol):
    engine.set_basis(mol.shells("cc-pvdz"), pure=True)
    engine.set_basis(mol.shells("cc-pvdz-jkfit"), pure=True, which=1)


def test_local_df_matches_oracle(oracle, engine, system):
    print("case: local DF-JK (H2O)3 default thresholds", flush=True)
    s = system
    _set(engine, s["mol"])
    loc = oracle.LocalDFJK(s["ob"], s["ab"], s["frags"], s["ob"].schwarz(), s["S"])
    Jo, Ko = loc(s["Cocc"], s["sig"], s["Pt"])
    J, K = engine.jk_df_local(s["frags"], s["Cocc"], s["sig"], s["Pt"])
    nocc = s["Cocc"].shape[1]
    for which, ref in enumerate((loc.lmo, loc.ao, loc.fit)):
        assert engine.df_local_domains(which, nocc, 3) == ref, which
    # the test system has truncated domains (otherwise it pins nothing local)
    assert any(len(d) < 3 for d in loc.fit) and any(len(d) < 3 for d in loc.lmo)
    print("  fit domain sizes %s; max|dJ| %.1e max|dK| %.1e" % ([len(d) for d in loc.fit], np.max(np.abs(J - Jo)),
                                                                 np.max(np.abs(K - Ko))), flush=True)
    assert np.max(np.abs(J - Jo)) < 1e-10 and np.max(np.abs(K - Ko)) < 1e-10
    # second call, other orbitals: the domains and blocks of the first call
    rng = np.random.default_rng(3)
    C2 = s["Cocc"] + 1e-2 * rng.normal(size=s["Cocc"].shape) * (np.abs(s["Cocc"]) > 0)
    sig2 = np.linalg.inv(C2.T @ s["S"] @ C2)
    P2 = C2 @ sig2 @ C2.T
    Jo2, Ko2 = loc(C2, sig2, P2)
    J2, K2 = engine.jk_df_local(s["frags"], C2, sig2, P2)
    assert np.max(np.abs(J2 - Jo2)) < 1e-10 and np.max(np.abs(K2 - Ko2)) < 1e-10
    assert np.max(np.abs(K2 - K)) > 1e-6   # the call did use the new orbitals


def test_local_df_all_inclusive_equals_global_df(engine, system):
    print("case: local DF-JK all-inclusive vs global DF", flush=True)
    s = system
    _set(engine, s["mol"])
    engine.df_local_reset()
    J, K = engine.jk_df_local(s["frags"], s["Cocc"], s["sig"], s["Pt"], mo_thresh=-1.0, fit_thresh=-1.0,
                              r_thresh=1e9, thrint=-1.0)
    w, U = np.linalg.eigh(s["sig"])
    Cp = s["Cocc"] @ ((U * np.sqrt(w)) @ U.T)
    Jg, Kg = engine.jk_df(Cp, fp32=False)
    print("  max|J - 2 J_DF| %.1e max|K - K_DF| %.1e" % (np.max(np.abs(J - 2 * Jg)), np.max(np.abs(K - Kg))),
          flush=True)
    assert np.max(np.abs(J - 2.0 * Jg)) < 1e-10 and np.max(np.abs(K - Kg)) < 1e-10


def test_local_df_reset_rebuilds_domains(oracle, engine, system):
    """A reset rebuilds the domains: every fit domain full, then every fit
    domain empty (r_thresh 0, fit_thresh 1e9: no LMO adds exchange, the
    Coulomb part is unchanged), each equal to the oracle's."""
    print("case: local DF-JK reset", flush=True)
    s = system
    _set(engine, s["mol"])
    nocc = s["Cocc"].shape[1]
    engine.df_local_reset()
    J1, _ = engine.jk_df_local(s["frags"], s["Cocc"], s["sig"], s["Pt"], fit_thresh=-1.0)
    assert engine.df_local_domains(2, nocc, 3) == [[0, 1, 2]] * nocc
    engine.df_local_reset()
    J0, K0 = engine.jk_df_local(s["frags"], s["Cocc"], s["sig"], s["Pt"], r_thresh=0.0, fit_thresh=1e9)
    assert engine.df_local_domains(2, nocc, 3) == [[]] * nocc
    assert np.max(np.abs(K0)) == 0.0
    loc = oracle.LocalDFJK(s["ob"], s["ab"], s["frags"], s["ob"].schwarz(), s["S"], r_thresh=0.0, fit_thresh=1e9)
    Jo, Ko = loc(s["Cocc"], s["sig"], s["Pt"])
    assert np.max(np.abs(Ko)) == 0.0 and np.max(np.abs(J0 - Jo)) < 1e-10
    assert np.max(np.abs(J0 - J1)) < 1e-10
