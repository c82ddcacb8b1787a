"""CPU-side checks of the native library and its host logic (no GPU calls)."""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "molecular2_amd", "csrc"))


def _declared_symbols():
    syms = set()
    for name in os.listdir(os.path.join(ROOT, "include")):
        if name.endswith(".h"):
            text = open(os.path.join(ROOT, "include", name)).read()
            syms |= set(re.findall(r"\b(m2_[a-z0-9_]+)\s*\(", text))
    return syms


def test_library_exports_every_declared_symbol():
    from molecular2_amd import engine
    if not os.path.exists(engine.LIBPATH):
        pytest.skip("libm2.so not built")
    lib = ctypes.CDLL(engine.LIBPATH)          # loads without touching the GPU
    syms = _declared_symbols()
    assert len(syms) >= 20
    for s in sorted(syms):
        assert hasattr(lib, s), s
    assert hasattr(lib, "m2_rhf")


def test_autograd_operator_registers():
    """libm2_torch.so loads without a GPU and registers m2::rhf_energy with
    the schema the Python wrapper calls."""
    from molecular2_amd import autograd
    if not os.path.exists(autograd.LIBPATH):
        pytest.skip("libm2_torch.so not built")
    import torch
    autograd.load()
    schema = str(torch.ops.m2.rhf_energy.default._schema)
    assert schema.startswith("m2::rhf_energy(Tensor R, Tensor Z, Tensor l, Tensor nprim")
    assert schema.endswith("-> Tensor")


def test_hessian_rejects_empty_molecule():
    """m2_rhf_hessian checks its arguments before any engine context exists"""
    from molecular2_amd import engine
    if not os.path.exists(engine.LIBPATH):
        pytest.skip("libm2.so not built")
    L = engine.lib()
    o = engine.SCFOptions()
    H = np.zeros(1)
    nul = ctypes.POINTER(ctypes.c_int)()
    dnul = ctypes.POINTER(ctypes.c_double)()
    st = L.m2_rhf_hessian(0, 1, 0, dnul, dnul, 0, 1, 1, 0, nul, nul, dnul, dnul, dnul, nul,
                          0, nul, nul, dnul, dnul, dnul, nul, 0, nul, nul, dnul, dnul, dnul, nul,
                          ctypes.byref(o), 0.01, H.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), None)
    assert st != 0
    assert b"null or empty" in L.m2_rhf_last_error()


def test_engine_fails_loudly_without_library(monkeypatch):
    from molecular2_amd import engine
    monkeypatch.setattr(engine, "LIBPATH", "/nonexistent/libm2.so")
    monkeypatch.setattr(engine, "_lib", None)
    with pytest.raises(RuntimeError, match="not built"):
        engine.lib()


def test_solid_harmonics_orthonormal_generator():
    import gen_eri
    for l in range(5):
        rows = np.array(gen_eri.solid_harmonics(l))
        cs = gen_eri.comps(l)
        G = np.zeros((len(cs), len(cs)))
        for i, a in enumerate(cs):
            for j, b in enumerate(cs):
                if all((a[k] + b[k]) % 2 == 0 for k in range(3)):
                    G[i, j] = np.prod([gen_eri.dfact(a[k] + b[k] - 1) for k in range(3)]) / gen_eri.dfact(2 * l - 1)
        np.testing.assert_allclose(rows @ G @ rows.T, np.eye(2 * l + 1), atol=1e-13)


def test_generator_programs_are_level_consistent():
    """Every instruction of a heavy-class program reads only slots written at
    an earlier level (or the base / accumulator slots), and no level writes a
    slot it also reads."""
    import gen_eri
    for cls in gen_eri.HEAVY:
        for pure in (0, 1):
            P = gen_eri.gen_program(*cls, pure)
            written = set(range(P["base"], P["base"] + P["L"] + 1))
            for lvl in P["vrr_levels"]:
                dsts = {d for d, _ in lvl}
                for d, t in lvl:
                    for src, sid, c in t:
                        assert src in written and src not in dsts
                written |= dsts
            live = {a for a, _ in P["acc"]}
            for lvl in P["post_levels"]:
                dsts = {d for d, _ in lvl}
                for d, t in lvl:
                    for src, sid, c in t:
                        assert src not in dsts
                live |= dsts
            assert len(P["out"]) == np.prod(P["sizes"])
            assert P["nslots"] * 8 <= 64 * 1024


def test_flop_model_committed_matches_generator():
    import json
    import gen_eri
    with open(os.path.join(ROOT, "molecular2_amd", "flop_model.json")) as fh:
        model = json.load(fh)["classes"]
    for key, v in model.items():
        X, Y = map(int, key.split(","))
        la, lb = gen_eri.PAIR_CLASSES[X]
        lc, ld = gen_eri.PAIR_CLASSES[Y]
        f = gen_eri.flop_model(la, lb, lc, ld)
        assert f["f_prim"] == v["f_prim"] and f["f_cont_pure"] == v["f_cont_pure"]


def test_water_cluster_generator():
    from molecular2_amd import molecules as M
    a = M.water_cluster(20, seed=20)
    b = M.water_cluster(20, seed=20)
    np.testing.assert_array_equal(a.xyz, b.xyz)
    assert a.natom == 60 and a.nel == 200
    shells = a.shells("cc-pvdz")
    assert len(shells) == 240 and M.nbf_of(shells, True) == 480
    assert M.nbf_of(a.shells("cc-pvdz-jkfit"), True) == 2320
    # intermolecular O..H, H..H >= 1 A
    xyz = a.xyz / M.TOBOHR
    for i in range(60):
        for j in range(i + 1, 60):
            if i // 3 != j // 3 and not (a.Z[i] == 8 and a.Z[j] == 8):
                assert np.linalg.norm(xyz[i] - xyz[j]) >= 1.0


def test_random_molecules_are_closed_shell():
    from molecular2_amd import molecules as M
    for m in M.random_batch(20, seed=3):
        assert m.nel % 2 == 0 and 3 <= m.natom <= 30
    c = M.random_chain(12, seed=12)
    assert c.natom == 12 and c.nel % 2 == 0


def test_missing_basis_raises_basislib():
    from molecular2_amd import molecules as M
    mol = M.Molecule([8], np.zeros((1, 3)))
    with pytest.raises(KeyError, match="BASISLIB"):
        mol.shells("nonexistent-basis")


def _decode_words(words, npass, nw=7):
    """[pass][word][lane] int table -> per pass, per lane, the word list."""
    out = []
    for p in range(npass):
        out.append([[words[(p * nw + k) * 64 + j] for k in range(nw)] for j in range(64)])
    return out


def test_heavy_pass_words_encode_the_programs():
    """The pass-ordered VRR / post words of the heavy J/K classes decode back
    to exactly the levelled instructions of gen_program (destination, terms,
    scalar ids, coefficients; one level per run of passes)."""
    import tempfile
    import gen_eri as G
    with tempfile.TemporaryDirectory() as d:
        G.write_programs(d)
        text = open(os.path.join(d, "programs.inc")).read()

    def arr(name):
        m = re.search(r"constexpr (?:int|double|unsigned long long) %s\[\d+\] = \{(.*?)\};" % name, text, re.S)
        return [float(x.replace("ull", "")) if "." in x or "e" in x else int(x.replace("ull", ""))
                for x in m.group(1).replace("\n", " ").split(",") if x.strip()]
    vdesc, vwords, vend = arr("VRRW_DESC"), arr("VRRW_WORDS"), arr("VRRW_LEVEND")
    pdesc, pwords, pcoef, pend = arr("POSTW_DESC"), arr("POSTW_WORDS"), arr("POSTW_COEF"), arr("POSTW_END")
    for k in range(len(vdesc) // 6):
        la, lb, lc, ld, npass, off = vdesc[6 * k:6 * k + 6]
        P = G.gen_program(la, lb, lc, ld, 1)
        passes = _decode_words(vwords[off:], npass)
        got, cur = [], []
        for p, lanes in enumerate(passes):
            for w in lanes:
                if w[0] < 0:
                    continue
                dst, nt = w[0] & 0xFFFF, w[0] >> 16
                terms = []
                for t in w[1:1 + nt]:
                    c = (t >> 18) & 63
                    terms.append((t & 8191, (t >> 13) & 31, float(c - 64 if c >= 32 else c)))
                cur.append((dst, terms))
            if (vend[k] >> p) & 1:
                got.append(cur)
                cur = []
        assert got == [[(dst, [tuple(x) for x in t]) for dst, t in lvl] for lvl in P["vrr_levels"]]
    for k in range(len(pdesc) // 10):
        la, lb, lc, ld, pure, npass, off, coff, ncoef, eoff = pdesc[10 * k:10 * k + 10]
        P = G.gen_program(la, lb, lc, ld, pure)
        passes = _decode_words(pwords[off:], npass)
        got, cur = [], []
        for p, lanes in enumerate(passes):
            for w in lanes:
                if w[0] < 0:
                    continue
                terms = []
                for t in w[1:1 + (w[0] >> 16)]:
                    sid = t >> 13
                    if sid >= G.NSCALAR:
                        terms.append((t & 8191, G.SID["one"], pcoef[coff + sid - G.NSCALAR]))
                    else:
                        terms.append((t & 8191, sid, 1.0))
                cur.append((w[0] & 0xFFFF, terms))
            if pend[eoff + p]:
                got.append(cur)
                cur = []
        assert got == [[(dst, [tuple(x) for x in t]) for dst, t in lvl] for lvl in P["post_levels"]]


def _contraction_terms(code):
    """(g component, Gamma index, sign, coefficient) of every term of the
    g[k] += lines"""
    out = []
    for line in code.split("\n"):
        m = re.match(r"\s+g\[(\d)\] \+= R\(0\) (.*);", line)
        if not m:
            continue
        for sg, coef, gi in re.findall(r"([+-]) (?:(R\([-0-9.e]+\)) \* )?Gam\[(\d+)\]", m.group(2)):
            out.append((int(m.group(1)), int(gi), sg, coef))
    return sorted(out)


def test_derivative_class_split_passes_partition_the_contraction():
    """The passes of a class split over the c components of Gamma together
    contract exactly the terms of the unsplit centre pass."""
    import gen_eri as G
    for c in sorted(G.DERIV_CSPLIT):
        for m in (1, 2, 4):
            whole = _contraction_terms(G.gen_deriv_class(*c, mask=m))
            parts = []
            for k in range(1, len(G.comps(c[2])) + 1):
                parts += _contraction_terms(G.gen_deriv_class(*c, mask=m | (k << 3)))
            assert whole and any(t[2] == "-" for t in whole) and sorted(parts) == whole, (c, m)
