"""Batched SCF (m2_rhf_batch): molecules of the random H/C/N/O batch run by
several host workers, each with its own context and stream, give the same
energies as one-at-a-time runs, and mixed precision stays within 1 uEh of
FP64 for every molecule."""
import numpy as np
import pytest

from molecular2_amd import molecules as M

pytestmark = pytest.mark.gpu

KW = dict(enconverge=1e-10, densconverge=1e-7, maxiter=100)


def _small_batch():
    mols = [m for m in M.random_batch(256, seed=256) if m.natom <= 9][:6]
    assert len(mols) == 6
    return mols


def test_batch_matches_single_runs(engine):
    mols = _small_batch()
    print("case: batch of %d molecules, 4 workers" % len(mols), flush=True)
    out, wall = engine.rhf_batch(mols, nworkers=4, **KW)
    assert wall > 0 and len(out) == len(mols)
    for m, r in zip(mols, out):
        s = engine.rhf(m, **KW)
        assert r["converged"] == s["converged"]
        if s["converged"]:
            assert abs(r["energy"] - s["energy"]) < 1e-9, (m.natom, r["energy"], s["energy"])


def test_batch_mixed_within_one_microhartree(engine):
    mols = _small_batch()
    print("case: batch mixed vs fp64", flush=True)
    r64, _ = engine.rhf_batch(mols, nworkers=4, **KW)
    rmx, _ = engine.rhf_batch(mols, nworkers=4, mixed=1, **KW)
    nconv = 0
    for a, b in zip(r64, rmx):
        if a["converged"] and b["converged"]:
            nconv += 1
            assert abs(a["energy"] - b["energy"]) < 1e-6
    assert nconv >= 4


def test_batch_rejects_soad_guess(engine):
    from molecular2_amd.engine import M2Error
    with pytest.raises(M2Error, match="guess"):
        engine.rhf_batch(_small_batch()[:1], guess=1)
