"""Convert the reference's basis-set sources into the small JSON files this
package ships (run once at development time; the JSON outputs are committed).

Sources (public numeric basis data inside the reference tree):
  * cc-pVDZ, H-Ne: Molpro text ``basissets/CC-PVDZfirst.basis``.  Each
    ``c, i.j, coefs...`` line is one segmented shell over primitives i..j of
    the preceding exponent line, which is how the reference splits general
    contractions into separate shells (src/basisreader.cpp:62-86 reads the
    XML built from this text with one <Shell> per contraction line).
  * cc-pVDZ-JKFIT: ``basislibrary/bases/build/jkfit4.xml`` (jkfit.list maps
    cc-pvdz -> 4).
  * STO-3G: ``basislibrary/bases/build/basis11.xml`` (basis.list: sto-3g -> 11).
  * "test" basis (He, 5 uncontracted s): ``basis12.xml`` (used by the
    reference's test/cdthe.mol).

Output format: {"name": ..., "elements": {"H": [{"l": 0, "exps": [...],
"coefs": [...]}, ...]}} with shells in file order and raw (unnormalised)
contraction coefficients, exactly as the reference hands them to
libint2::Shell (src/basis.cpp:33-66).

Usage:  python convert_basis.py /root/reference
"""
import json
import os
import re
import sys
import xml.etree.ElementTree as ET

LVAL = {"s": 0, "p": 1, "d": 2, "f": 3, "g": 4, "h": 5, "i": 6}
SYMBOLS = ["H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne"]


def parse_molpro(path, keep):
    out = {}
    cur_l = None
    cur_el = None
    cur_exps = None
    with open(path) as fh:
        for raw in fh:
            line = raw.strip()
            if not line or line.startswith("!") or line.startswith("basis") or line.startswith("}"):
                continue
            parts = [p.strip() for p in line.split(",")]
            if parts[0] in LVAL:
                cur_l = LVAL[parts[0]]
                cur_el = parts[1].capitalize()
                cur_exps = [float(x) for x in parts[2:] if x]
            elif parts[0] == "c":
                if cur_el not in keep:
                    continue
                lo, hi = (int(v) for v in parts[1].split("."))
                coefs = [float(x) for x in parts[2:] if x]
                exps = cur_exps[lo - 1:hi]
                assert len(exps) == len(coefs), (cur_el, line)
                out.setdefault(cur_el, []).append({"l": cur_l, "exps": exps, "coefs": coefs})
    return out


def parse_xml(path, keep):
    root = ET.parse(path).getroot()
    out = {}
    for el in root:
        sym = el.tag.capitalize()
        if sym not in keep:
            continue
        shells = []
        for sh in el.findall("Shell"):
            l = LVAL[sh.attrib["lval"].strip()]
            exps, coefs = [], []
            for xc in sh.findall("xc"):
                exps.append(float(xc.attrib["x"]))
                coefs.append(float(xc.attrib["c"]))
            shells.append({"l": l, "exps": exps, "coefs": coefs})
        out[sym] = shells
    return out


def main(ref):
    here = os.path.dirname(os.path.abspath(__file__))
    keep = {"H", "C", "N", "O"}
    jobs = [
        ("cc-pvdz", parse_molpro(os.path.join(ref, "basissets/CC-PVDZfirst.basis"), keep)),
        ("cc-pvdz-jkfit", parse_xml(os.path.join(ref, "basislibrary/bases/build/jkfit4.xml"), keep)),
        ("sto-3g", parse_xml(os.path.join(ref, "basislibrary/bases/build/basis11.xml"), keep)),
        ("test", parse_xml(os.path.join(ref, "basislibrary/bases/build/basis12.xml"), {"He"})),
    ]
    for name, data in jobs:
        with open(os.path.join(here, name + ".json"), "w") as fh:
            json.dump({"name": name, "elements": data}, fh, indent=1)
        print(name, {k: len(v) for k, v in data.items()})


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
