"""Host-side reading of the reference's own input files (no GPU): input
decks (ProgramController::parse, src/ProgramController.cpp:137-236;
Molecule::init / parseGeomLine / parseBasis, src/molecule.cpp:20-240) and the
XML basis library (BasisReader, src/basisreader.cpp:27-132), through the
C ABI m2_deck_molecule.  Fixtures: decks copied from the reference's test/
directory and a library in its XML layout (tests/golden/make_basislib.py)."""
import os

import numpy as np
import pytest

from molecular2_amd import molecules as M

HERE = os.path.dirname(os.path.abspath(__file__))
DECKS = os.path.join(HERE, "golden", "decks")
LIB = os.path.join(HERE, "golden", "basislib")


def _deck(name):
    with open(os.path.join(DECKS, name)) as fh:
        return fh.read()


def _load(text, imol=0):
    from molecular2_amd import engine
    if not os.path.exists(engine.LIBPATH):
        pytest.skip("libm2.so not built")
    return engine.deck_molecule(text, imol, LIB)


def _same_shells(a, b):
    assert len(a) == len(b)
    for x, y in zip(a, b):
        assert x.l == y.l and x.atom == y.atom
        assert list(x.exps) == list(y.exps) and list(x.coefs) == list(y.coefs)
        np.testing.assert_array_equal(np.asarray(x.center), np.asarray(y.center))


def test_benzene_deck_angstrom_geometry_and_library_basis():
    mol, orb, aux, mini = _load(_deck("benzene-dz.mol"))
    ref = M.benzene_dz_ref()
    assert list(mol.Z) == list(ref.Z) and mol.charge == 0 and mol.multiplicity == 1
    np.testing.assert_array_equal(mol.xyz, ref.xyz)          # angstrom -> bohr, 1.889726124565
    _same_shells(orb, ref.shells("cc-pvdz"))                # library XML == packaged cc-pVDZ
    _same_shells(mini, ref.shells("sto-3g"))
    assert aux == []                                        # the deck names no jkbasis


def test_h2otest_deck_bohr_geometry_comments_and_sto3g():
    mol, orb, aux, mini = _load(_deck("h2otest.mol"))
    ref = M.water_h2otest()
    assert list(mol.Z) == [8, 1, 1]
    np.testing.assert_array_equal(mol.xyz, ref.xyz)          # no "angstrom": bohr as written
    _same_shells(orb, ref.shells("sto-3g"))


def test_methane_deck_and_jkbasis_construct():
    mol, orb, aux, _ = _load(_deck("methane.mol"))
    assert list(mol.Z) == [6, 1, 1, 1, 1]
    assert abs(mol.xyz[1, 0] - 1.18377) < 1e-15
    # a jkbasis construct and a per-element basis line, in the reference's syntax
    deck = ("molecule = {\n charge, 1\n multiplicity, 2\n geometry = {\n  N, 0.0, 0.0, 0.0\n  H, 0.0, 0.0, 1.9\n }\n"
            " basis = {\n  default, cc-pvdz\n  h, sto-3g\n }\n jkbasis = {\n  default, cc-pvdz\n }\n}\nhf()\n")
    mol, orb, aux, _ = _load(deck)
    assert mol.charge == 1 and mol.multiplicity == 2 and list(mol.Z) == [7, 1]
    nh = M.Molecule([7, 1], np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 1.9]]))
    assert [s.l for s in orb] == [s.l for s in nh.shells("cc-pvdz") if s.atom == 0] + \
        [s.l for s in nh.shells("sto-3g") if s.atom == 1]
    _same_shells(aux, nh.shells("cc-pvdz-jkfit"))


def test_second_molecule_construct_and_errors():
    from molecular2_amd.engine import M2Error
    two = _deck("h2otest.mol") + "\n" + _deck("benzene-dz.mol")
    mol, _, _, _ = _load(two, 1)
    assert mol.natom == 12
    with pytest.raises(M2Error, match="no molecule construct"):
        _load(two, 2)
    with pytest.raises(M2Error, match="BASISLIB: Could not read basis avqz for atom o"):
        _load(_deck("oxygen.mol"))                        # avqz is not in the fixture library
    with pytest.raises(M2Error, match="INPUT: unknown element"):
        _load("molecule = {\n geometry = {\n  Qq, 0, 0, 0\n }\n basis = {\n  default, sto-3g\n }\n}\n")
    with pytest.raises(M2Error, match="Missing coordinate"):
        _load("molecule = {\n geometry = {\n  H, 0, 0\n }\n basis = {\n  default, sto-3g\n }\n}\n")


SETS = os.path.join(HERE, "golden", "basissets")


def test_molpro_basis_file_matches_library_data():
    """The Molpro text basissets/CC-PVDZfirst.basis read by the library
    (m2_basis_file_shells) gives, for H, C, N and O, the shells of the
    committed cc-pVDZ data (converted offline from the same file): one
    segmented shell per "c, i.j" contraction line over the exponents i..j,
    in file order, raw coefficients."""
    from molecular2_amd import engine
    if not os.path.exists(engine.LIBPATH):
        pytest.skip("libm2.so not built")
    mol = M.from_angstrom(["O", "H", "C", "N", "H"], [[0, 0, 0], [0.96, 0, 0], [0, 1.5, 0], [0, 0, 1.6],
                                                      [1.0, 1.0, 1.0]])
    got = engine.basis_file_shells(os.path.join(SETS, "CC-PVDZfirst.basis"), mol)
    _same_shells(got, mol.shells("cc-pvdz"))
    with pytest.raises(engine.M2Error, match="BASISLIB"):
        engine.basis_file_shells(os.path.join(SETS, "missing.basis"), mol)


def test_deck_naming_a_molpro_file():
    """A deck whose basis names a ".basis" file reads it from the basissets
    directory beside the library and gets the same shells as the library's
    cc-pvdz entry."""
    text = _deck("benzene-dz.mol")
    assert "cc-pvdz" in text
    mol, orb, _, _ = _load(text)
    mol2, orb2, _, _ = _load(text.replace("cc-pvdz", "CC-PVDZfirst.basis"))
    np.testing.assert_array_equal(mol.xyz, mol2.xyz)
    _same_shells(orb2, orb)
