"""Write the input-file fixtures of the deck and basis-library tests.

  decks/*.mol   -- input decks of the reference's own tests (test/*.mol),
                   copied unchanged: h2otest, benzene-dz, methane, oxygen,
                   h2odimer, cdthe.
  basislib/     -- a basis library in the reference's XML layout
                   (basislibrary/: basis.list, jkfit.list,
                   bases/build/<kind>N.xml; read by src/basisreader.cpp:27-132)
                   for H, C, N, O: the two .list files copied; basis11.xml
                   (STO-3G) and jkfit4.xml (cc-pVDZ-JKFIT) cut to those
                   elements; basis12.xml ("test", He) copied; basis13.xml (cc-pVDZ) written from the Molpro
                   text basissets/CC-PVDZfirst.basis, one <Shell> per
                   contraction line, because the reference's shipped
                   basis13.xml holds only Na-Ar.
  basissets/    -- the Molpro text basissets/CC-PVDZfirst.basis, copied
                   unchanged (read by decks that name the file directly).

Usage:  python tests/golden/make_basislib.py /root/reference
"""
import os
import shutil
import sys
import xml.etree.ElementTree as ET

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "molecular2_amd", "data"))
import convert_basis  # noqa: E402

KEEP = ["h", "he", "c", "n", "o"]
LNAME = {0: "s", 1: "p", 2: "d", 3: "f"}


def write_xml(path, root_name, elements):
    lines = ['<root name="%s">' % root_name]
    for sym in KEEP:
        shells = elements.get(sym.capitalize())
        if not shells:
            continue
        lines.append('  <%s nshells="%d">' % (sym, len(shells)))
        for sh in shells:
            lines.append('    <Shell lval="%s" nexp="%d">' % (LNAME[sh["l"]], len(sh["exps"])))
            for x, c in zip(sh["exps"], sh["coefs"]):
                lines.append('      <xc c=" %r" x=" %r"/>' % (c, x))
            lines.append("    </Shell>")
        lines.append("  </%s>" % sym)
    lines.append("</root>")
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def cut_xml(src, dst):
    root = ET.parse(src).getroot()
    for el in list(root):
        if el.tag not in KEEP:
            root.remove(el)
    ET.ElementTree(root).write(dst)


def main(ref):
    for deck in ("h2otest", "benzene-dz", "methane", "oxygen", "h2odimer", "cdthe"):
        shutil.copyfile(os.path.join(ref, "test", deck + ".mol"), os.path.join(HERE, "decks", deck + ".mol"))
    lib = os.path.join(HERE, "basislib")
    for lst in ("basis.list", "jkfit.list"):
        shutil.copyfile(os.path.join(ref, "basislibrary", lst), os.path.join(lib, lst))
    build = os.path.join(lib, "bases", "build")
    os.makedirs(build, exist_ok=True)
    cut_xml(os.path.join(ref, "basislibrary/bases/build/basis11.xml"), os.path.join(build, "basis11.xml"))
    cut_xml(os.path.join(ref, "basislibrary/bases/build/jkfit4.xml"), os.path.join(build, "jkfit4.xml"))
    # the "test" basis of test/cdthe.mol (He only), copied unchanged
    shutil.copyfile(os.path.join(ref, "basislibrary/bases/build/basis12.xml"), os.path.join(build, "basis12.xml"))
    dz = convert_basis.parse_molpro(os.path.join(ref, "basissets/CC-PVDZfirst.basis"), {s.capitalize() for s in KEEP})
    write_xml(os.path.join(build, "basis13.xml"), "vdz", dz)
    # the Molpro text itself, beside the library as in the reference's tree
    # (basissets/ next to basislibrary/), for decks that name the file
    sets = os.path.join(HERE, "basissets")
    os.makedirs(sets, exist_ok=True)
    shutil.copyfile(os.path.join(ref, "basissets", "CC-PVDZfirst.basis"), os.path.join(sets, "CC-PVDZfirst.basis"))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
