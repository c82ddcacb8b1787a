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
    out, wall = engine.rhf_batch(mols, nworkers=4, **KW, guess="hcore")
    assert wall > 0 and len(out) == len(mols)
    for m, r in zip(mols, out):
        s = engine.rhf(m, **KW, guess="hcore")
        assert r["converged"] == s["converged"]
        if s["converged"]:
            assert abs(r["energy"] - s["energy"]) < 1e-9, (m.natom, r["energy"], s["energy"])


def test_batch_mixed_within_one_microhartree(engine):
    mols = _small_batch()
    print("case: batch mixed vs fp64", flush=True)
    r64, _ = engine.rhf_batch(mols, nworkers=4, **KW, guess="hcore")
    rmx, _ = engine.rhf_batch(mols, nworkers=4, mixed=1, **KW, guess="hcore")
    nconv = 0
    for a, b in zip(r64, rmx):
        if a["converged"] and b["converged"]:
            nconv += 1
            assert abs(a["energy"] - b["energy"]) < 1e-6
    assert nconv >= 4


def test_batch_soad_guess_matches_single_runs(engine):
    """m2_rhf_batch_min with the STO-3G shells: the SOAD-started batch gives
    the single SOAD-started runs' energies (the iteration counts of the slowly
    converging molecule vary with FP64 atomics noise, so they are not
    compared)."""
    mols = _small_batch()[:4]
    # one of them needs ~100 iterations from this guess: room enough that no
    # run ends at the limit, where FP64 atomics noise decides convergence
    kw = dict(KW, maxiter=160)
    print("case: SOAD batch of %d molecules, 4 workers" % len(mols), flush=True)
    out, _ = engine.rhf_batch(mols, nworkers=4, guess="soad", **kw)
    for m, r in zip(mols, out):
        s = engine.rhf(m, guess="soad", **kw)
        assert r["converged"] and s["converged"], (m.natom, r["iterations"], s["iterations"])
        assert abs(r["energy"] - s["energy"]) < 5e-9, (m.natom, r["energy"], s["energy"])


def test_batch_without_minimal_shells_rejects_soad(engine):
    """m2_rhf_batch has no minimal shells: guess = 1 is refused before any run"""
    import ctypes as C
    from molecular2_amd.engine import SCFOptions, SCFResult
    L = engine.L
    o = SCFOptions(diis=1, maxdiis=8, maxiter=40, enconverge=1e-7, densconverge=1e-4, precision=1e-12, df=0,
                   mixed=0, mixed_threshold=1e-4, fp64_final=2, guess=1, diis_start=2)
    res = (SCFResult * 1)()
    z = np.zeros(8, np.int32)
    d = np.zeros(8)
    ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    st = L.m2_rhf_batch(0, 1, 1, ip(z), ip(z), dp(d), dp(d), 1, ip(z), ip(z), ip(z), dp(d), dp(d), dp(d), ip(z),
                        C.byref(o), res)
    assert st != 0 and b"minimal shells" in L.m2_rhf_last_error()


def test_batch_matches_oracle_energies(engine):
    """The batch against the oracle's SCF::rhf restatement over its
    formJKdirect restatement (tests/golden/batch_oracle.json, written by
    tests/golden/make_batch_oracle.py with the same molecules, guess and
    convergence options): every molecule both converge agrees to 1e-7 Eh."""
    import json
    import os
    with open(os.path.join(os.path.dirname(__file__), "golden", "batch_oracle.json")) as fh:
        ref = json.load(fh)
    batch = M.random_batch(256, seed=256)
    mols = [batch[r["index"]] for r in ref["molecules"]]
    assert [m.natom for m in mols] == [m.natom for m in _small_batch()]
    print("case: batch of %d molecules vs oracle energies" % len(mols), flush=True)
    out, _ = engine.rhf_batch(mols, nworkers=4, guess=ref["guess"], **ref["options"])
    both = 0
    for r, o in zip(out, ref["molecules"]):
        if r["converged"] and o["converged"]:
            both += 1
            assert abs(r["energy"] - o["energy"]) < 1e-7, (o["index"], r["energy"], o["energy"])
    assert both >= 4
