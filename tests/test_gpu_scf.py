"""Host C++ SCF (molecular::SCF::rhf over the HIP engine) against the
reference's recorded energies and the oracle's SCF restatement."""
import numpy as np
import pytest

from molecular2_amd import molecules as M

pytestmark = pytest.mark.gpu


def test_water_6d_energy_pin(engine):
    """test/h2otest.out_1:124: E(RHF/cc-pVDZ, 6 Cartesian d) = -75.990178761701."""
    print("case: C++ SCF water 6d", flush=True)
    r = engine.rhf(M.water_h2otest(), pure=False, enconverge=1e-11, densconverge=1e-9)
    assert r["converged"]
    assert abs(r["energy"] - (-75.990178761701)) < 1e-8
    # D S D = 2 D and tr(D S) = N_el for the converged closed-shell density
    S, T, V = engine.one_body(M.water_h2otest().Z, M.water_h2otest().xyz)
    D = r["D"]
    assert abs(np.trace(D @ S) - 10) < 1e-8
    assert np.max(np.abs(D @ S @ D - 2 * D)) < 1e-7


def test_benzene_6d_energy_pin(engine):
    """test/benzene.out_1:178: E = -230.723053 (printed to 1e-6)."""
    print("case: C++ SCF benzene 6d", flush=True)
    r = engine.rhf(M.benzene_dz_ref(), pure=False, enconverge=1e-9, densconverge=1e-6)
    assert r["converged"]
    assert abs(r["energy"] - (-230.723053)) < 1.5e-6


def test_water_5d_matches_oracle_scf(oracle, engine):
    mol = M.water()
    print("case: C++ SCF water 5d vs oracle", flush=True)
    r = engine.rhf(mol, pure=True, enconverge=1e-10, densconverge=1e-8)
    ob = oracle.Basis(mol.shells("cc-pvdz"), pure=True)
    S, T, V = ob.one_body(mol.Z, mol.xyz)
    Q = ob.schwarz()
    scf = oracle.RHF(S, T + V, mol.enuc(), mol.nel, lambda P, m: ob.jk_direct(P, Q, 1e-12, m),
                     enconverge=1e-10, densconverge=1e-8)
    E = scf.run()
    assert abs(r["energy"] - E) < 1e-8


def test_mixed_precision_scf_benzene(engine):
    """Mixed precision (FP32 quartets below 1e-4, FP32 Delta-D builds, final
    FP64 iterations) reaches the FP64 energy to 1 microhartree."""
    mol = M.benzene()
    print("case: mixed-precision SCF benzene", flush=True)
    r64 = engine.rhf(mol, pure=True, enconverge=1e-9, densconverge=1e-6)
    rmx = engine.rhf(mol, pure=True, enconverge=1e-9, densconverge=1e-6, mixed=1, mixed_threshold=1e-4)
    assert r64["converged"] and rmx["converged"]
    assert rmx["n_fp32"] > 0
    assert abs(rmx["energy"] - r64["energy"]) < 1e-6


def test_odd_electrons_rejected(engine):
    mol = M.Molecule([8, 1], np.array([[0, 0, 0], [0, 0, 1.8]]))
    with pytest.raises(Exception, match="odd number of electrons"):
        engine.rhf(mol, pure=True)
