"""Parity at the workload's own size: the (H2O)20 cc-pVDZ cluster (seed 20,
480 spherical basis functions, 240 shells) and the batched random molecules.

  * direct J/K (FP64) against the oracle's formJKdirect restatement
    (src/fockbuild.cpp:64-164) stored in tests/golden/h2o20_jk.npz
    (written by tests/golden/make_h2o20_jk.py), to 1e-10, with the number of
    computed quartets; mixed precision against the same fixture;
  * shell-block norms (src/integrals.cpp:451-469) of a full density, of a
    density difference and of a non-symmetric matrix against the oracle;
  * density-fitted J/K (src/fockbuild.cpp:585-660): FP32 MFMA against FP64 at
    the converged orbitals, as the two-electron energy tr(P (2J - K));
  * the RHF gradient (src/rhfgrad.cpp:15-89): translation and rotation
    invariance, central finite differences of the device SCF energy
    (h = 1e-4 bohr) on two components, mixed precision against FP64.
"""
import os

import numpy as np
import pytest

from molecular2_amd import molecules as M

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _unpack(n, tri):
    A = np.zeros((n, n))
    iu = np.triu_indices(n)
    A[iu] = tri
    A[(iu[1], iu[0])] = tri
    return A


@pytest.fixture(scope="module")
def h2o20():
    return M.water_cluster(20, seed=20)


@pytest.fixture(scope="module")
def fixture_jk():
    d = np.load(os.path.join(GOLDEN, "h2o20_jk.npz"))
    n = int(d["nbf"])
    return dict(D=_unpack(n, d["D32"].astype(np.float64)), J=_unpack(n, d["J"]), K=_unpack(n, d["K"]),
                quartets=int(d["quartets"]), precision=float(d["precision"]))


def test_h2o20_jk_fp64_vs_oracle(engine, h2o20, fixture_jk):
    print("case: (H2O)20 FP64 direct J/K vs oracle fixture", flush=True)
    nbf = engine.set_basis(h2o20.shells("cc-pvdz"), pure=True)
    assert nbf == 480
    D = fixture_jk["D"]
    engine.set_profiling(True, counters=True)
    engine.reset_profile()
    try:
        J, K = engine.jk(D, 1.0, fixture_jk["precision"])
        quartets = sum(v[2] for v in engine.profile().values())
    finally:
        engine.set_profiling(False)
    dj = np.max(np.abs(J - fixture_jk["J"]))
    dk = np.max(np.abs(K - fixture_jk["K"]))
    print("  max|dJ| %.2e max|dK| %.2e quartets %d (oracle %d)" % (dj, dk, quartets, fixture_jk["quartets"]),
          flush=True)
    assert dj < 1e-10 and dk < 1e-10, (dj, dk)
    # The same screening predicate on both sides; the device and oracle
    # Schwarz bounds agree to ~1e-15 relative, so only quartets whose bound
    # lies within that of the threshold may flip (none expected).
    assert abs(quartets - fixture_jk["quartets"]) <= 10, (quartets, fixture_jk["quartets"])


def test_h2o20_jk_mixed_vs_oracle(engine, h2o20, fixture_jk):
    print("case: (H2O)20 mixed J/K vs oracle fixture", flush=True)
    engine.set_basis(h2o20.shells("cc-pvdz"), pure=True)
    D = fixture_jk["D"]
    try:
        for split in (1e-4, 1e-2):
            engine.set_mixed_threshold(split)
            J, K = engine.jk(D, 1.0, fixture_jk["precision"], mixed=True)
            dj = np.max(np.abs(J - fixture_jk["J"]))
            dk = np.max(np.abs(K - fixture_jk["K"]))
            dE = abs(0.5 * np.sum(D * ((J - 0.5 * K) - (fixture_jk["J"] - 0.5 * fixture_jk["K"]))))
            print("  split %.0e: max|dJ| %.2e max|dK| %.2e |dE2| %.2e" % (split, dj, dk, dE), flush=True)
            assert dj < 1e-7 and dk < 1e-7 and dE < 1e-6
    finally:
        engine.set_mixed_threshold(1e-4)


def test_shellblock_norm_vs_oracle(oracle, engine, h2o20, fixture_jk):
    """m2_shellblock_norm against the oracle's compute_shellblock_norm for a
    full density, a density difference and a non-symmetric matrix (rows =
    first shell index), on water and on (H2O)20."""
    for mol, pure in ((M.water(), False), (h2o20, True)):
        print("case: shellblock norms natom=%d" % mol.natom, flush=True)
        sh = mol.shells("cc-pvdz")
        ob = oracle.Basis(sh, pure=pure)
        engine.set_basis(sh, pure=pure)
        n = ob.nbf
        rng = np.random.default_rng(n)
        if mol.natom == 60:
            D = fixture_jk["D"]
        else:
            C = rng.normal(size=(n, 5))
            D = 2 * C @ C.T
        D2 = D + 1e-3 * (lambda X: X + X.T)(rng.normal(size=(n, n)))
        for A in (D, D2 - D, rng.normal(size=(n, n))):
            got = engine.shellblock_norm(A)
            np.testing.assert_array_equal(got, ob.shellblock_norm(A))


def test_jk_unscreened_equals_oracle_without_schwarz(oracle, engine):
    """precision 0 computes every unique quartet, which is what the reference
    does for an empty Schwarz matrix (src/fockbuild.cpp:74,114); the host
    Fock::formJKdirect passes 0 for an empty Schwarz."""
    mol = M.water()
    print("case: water unscreened J/K", flush=True)
    sh = mol.shells("cc-pvdz")
    ob = oracle.Basis(sh, pure=True)
    engine.set_basis(sh, pure=True)
    rng = np.random.default_rng(9)
    C = rng.normal(size=(ob.nbf, 6))
    D = 2 * C @ C.T / ob.nbf
    Jo, Ko = ob.jk_direct(D, None, 0.0, 1.0)
    engine.set_profiling(True, counters=True)
    engine.reset_profile()
    try:
        J, K = engine.jk(D, 1.0, 0.0)
        quartets = sum(v[2] for v in engine.profile().values())
    finally:
        engine.set_profiling(False)
    assert np.max(np.abs(J - Jo)) < 1e-12 and np.max(np.abs(K - Ko)) < 1e-12
    nsh = ob.nsh
    npair = nsh * (nsh + 1) // 2
    assert quartets == ob.last_quartets == npair * (npair + 1) // 2


def test_rejected_basis_leaves_context_empty(engine):
    """A failed m2_basis_set (an f shell in the orbital basis) leaves no
    basis behind: J/K, one-body and Schwarz calls return an error."""
    from molecular2_amd.engine import M2Error
    print("case: rejected basis", flush=True)
    mol = M.water()
    sh = mol.shells("cc-pvdz")
    engine.set_basis(sh, pure=True)
    bad = list(sh)
    bad[3] = M.Shell(3, [1.0], [1.0], bad[3].center, bad[3].atom)
    with pytest.raises(M2Error):
        engine.set_basis(bad, pure=True)
    D = np.eye(24)
    with pytest.raises(M2Error):
        engine.jk(D, 1.0, 1e-12)
    with pytest.raises(M2Error):
        engine.schwarz()
    with pytest.raises(M2Error):
        engine.one_body(mol.Z, mol.xyz)
    engine.set_basis(sh, pure=True)   # usable again afterwards
    J, K = engine.jk(D, 1.0, 1e-12)
    assert np.isfinite(J).all()


def _pair_energy(Cocc, J, K):
    P = Cocc @ Cocc.T
    return float(np.sum(P * (2.0 * J - K)))


@pytest.mark.parametrize("name", ["water", "h2o20"])
def test_df_fp32_energy_error(engine, h2o20, name):
    """FP32 (MFMA GEMMs with FP32 accumulation; gamma, the slice sum of K and
    the Coulomb contraction in FP64) against FP64 density-fitted J/K at the
    converged FP64-DF orbitals, as the two-electron energy tr(P (2J - K)),
    P = C_occ C_occ^T.  The error is at the FP32 level relative to that energy
    (water 4e-7 Eh; (H2O)20 6e-4 Eh of ~5e3 Eh, dominated by the FP32
    accumulation of the exchange GEMM).  A mixed-precision DF SCF therefore
    finishes with FP64 iterations, and its energy is within 1 uEh of the FP64
    DF SCF's."""
    mol = M.water() if name == "water" else h2o20
    print("case: DF FP32 vs FP64 energy, %s" % name, flush=True)
    kw = dict(pure=True, aux="cc-pvdz-jkfit", df=1, enconverge=1e-10, densconverge=1e-7, maxiter=100)
    r = engine.rhf(mol, **kw)
    assert r["converged"]
    Co = r["C"][:, : mol.nel // 2]
    J64, K64 = engine.jk_df(Co, fp32=False)
    J32, K32 = engine.jk_df(Co, fp32=True)
    E64 = _pair_energy(Co, J64, K64)
    dE = abs(_pair_energy(Co, J32, K32) - E64)
    print("  E2 %.6f |dE2| %.2e max|dJ| %.2e max|dK| %.2e" % (E64, dE, np.max(np.abs(J32 - J64)),
                                                              np.max(np.abs(K32 - K64))), flush=True)
    assert dE < 5e-7 * abs(E64)
    if name == "water":
        assert dE < 1e-6
    rm = engine.rhf(mol, mixed=1, **kw)
    assert rm["converged"] and rm["n_fp32"] > 0
    print("  mixed DF SCF: |dE| %.2e after %d FP32 builds" % (abs(rm["energy"] - r["energy"]), rm["n_fp32"]),
          flush=True)
    assert abs(rm["energy"] - r["energy"]) < 1e-6


def test_h2o20_gradient_invariances_fd_and_mixed(engine, h2o20):
    """(H2O)20 RHF gradient: sum_A g_A = 0, sum_A R_A x g_A = 0, two
    four-point finite differences of the SCF energy (h = 2e-3 bohr, SOAD
    start), and the mixed-precision gradient within 1e-6 Eh/bohr of FP64."""
    mol = h2o20
    print("case: (H2O)20 gradient", flush=True)
    # At the default screening (1e-12) each incremental Delta-D build drops
    # ~1e8 quartets below the threshold, and the incremental Fock matrix
    # drifts by ~1e-9 Eh between resets (the reference's own scheme,
    # src/fockbuild.cpp:176-203), which keeps |dE| from settling at 1e-10.
    # Screening at 1e-14 and convergence to 1e-11 leave each SCF energy
    # ~1e-10 Eh of drift of either sign (two-point differences at h = 1e-4
    # have landed +-5e-7 from the same analytic value), so the derivative is
    # the four-point stencil at h = 2e-3: noise <= 18e-10 / (12 h) = 7.5e-8,
    # truncation h^4 E^(5) / 30 ~ 1e-11.
    kw = dict(enconverge=1e-11, densconverge=1e-8, maxiter=100, precision=1e-14)
    r = engine.rhf(mol, gradient=True, **kw)
    assert r["converged"]
    g = r["grad"]
    assert np.max(np.abs(g.sum(axis=0))) < 1e-6
    torque = np.cross(mol.xyz, g).sum(axis=0)
    assert np.max(np.abs(torque)) < 1e-5, torque
    h = 2e-3
    for atom, axis in ((0, 0), (37, 2)):
        e = {}
        for s in (-2, -1, 1, 2):
            r_s = engine.rhf(mol.displaced(atom, axis, s * h), **kw)
            assert r_s["converged"]
            e[s] = r_s["energy"]
        fd = (8 * (e[1] - e[-1]) - (e[2] - e[-2])) / (12 * h)
        print("  g[%d,%d] analytic %.9f fd %.9f" % (atom, axis, g[atom, axis], fd), flush=True)
        assert abs(g[atom, axis] - fd) < 5e-7, (atom, axis, g[atom, axis], fd)
    rm = engine.rhf(mol, gradient=True, mixed=1, **kw)
    assert rm["converged"]
    assert abs(rm["energy"] - r["energy"]) < 1e-6
    dg = np.max(np.abs(rm["grad"] - g))
    print("  mixed: |dE| %.2e max|dg| %.2e" % (abs(rm["energy"] - r["energy"]), dg), flush=True)
    assert dg < 1e-6
