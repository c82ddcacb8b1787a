"""The analytic RHF Hessian with CPHF response (m2_rhf_hessian_cphf, the
reference's Fock::compute_hessian, src/rhfgrad.cpp:91-359, and its CPHF
solver, src/cphf.cpp:4-34) on water cc-pVDZ, pinned by:

  * the numerical Hessian of central differences of the analytic gradient
    over re-converged SCFs (m2_rhf_hessian, Fock::compute_hessian_numeric),
    which shares nothing with the CPHF path but the integrals;
  * translational invariance: every column sums to zero over the atoms for
    each axis;
  * the derivative-integral step: 1e-4 and 2e-4 bohr agree;
  * the reference's CPHF settings (1e-4, 30 steps) against a tight solve.
"""
import numpy as np
import pytest

from molecular2_amd import molecules as M

pytestmark = pytest.mark.gpu

TIGHT = dict(enconverge=1e-11, densconverge=1e-9, maxiter=100, precision=1e-14)


@pytest.fixture(scope="module")
def water_hessian(engine):
    print("case: water CPHF Hessian (tight)", flush=True)
    H, r = engine.rhf_hessian_cphf(M.water(), cphf_tol=1e-10, cphf_maxiter=200, **TIGHT)
    assert r["converged"] and r["cphf_unconverged"] == 0
    print("  %.2f s, %d CPHF steps" % (r["total_seconds"], r["cphf_steps"]), flush=True)
    return H


def test_cphf_hessian_matches_numerical(engine, water_hessian):
    H = water_hessian
    print("case: water numerical Hessian", flush=True)
    Hn, runs = engine.rhf_hessian(M.water(), step=2e-3, **TIGHT)
    assert all(r["converged"] for r in runs)
    d = np.max(np.abs(H - Hn))
    wa, wn = np.linalg.eigvalsh(H), np.linalg.eigvalsh(Hn)
    print("  max|H_cphf - H_num| %.2e; eigenvalues %s vs %s" % (d, np.round(wa, 5), np.round(wn, 5)), flush=True)
    assert d < 2e-5, d
    assert np.max(np.abs(wa - wn)) < 2e-5


def test_cphf_hessian_translational_invariance(water_hessian):
    H = water_hessian
    assert np.max(np.abs(H - H.T)) == 0.0
    for ax in range(3):
        s = H[ax::3, :].sum(axis=0)
        assert np.max(np.abs(s)) < 1e-6, (ax, s)


def test_cphf_hessian_step_and_reference_solver(engine, water_hessian):
    H = water_hessian
    print("case: water CPHF Hessian step 2e-4 and reference CPHF settings", flush=True)
    H2, _ = engine.rhf_hessian_cphf(M.water(), step=2e-4, cphf_tol=1e-10, cphf_maxiter=200, **TIGHT)
    assert np.max(np.abs(H2 - H)) < 1e-6, np.max(np.abs(H2 - H))
    Hr, r = engine.rhf_hessian_cphf(M.water(), **TIGHT)   # cphf_tol 1e-4, 30 steps
    print("  reference CPHF settings: max|dH| %.2e, %d steps" % (np.max(np.abs(Hr - H)), r["cphf_steps"]), flush=True)
    assert np.max(np.abs(Hr - H)) < 1e-3
