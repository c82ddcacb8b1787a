"""Extract the 20 water monomers of the reference's test/h2oclusters/h2o_20.mol
(geometry in Angstrom) and the monomer RHF energies its recorded run printed
(test/h2oclusters/h2o_20.ref, "Monomer k energy", to 1e-6 Hartree; density
fitted, cc-pVDZ with the cc-pVDZ-JKFIT auxiliary basis) into a JSON fixture.

    python tests/golden/make_h2o20_monomers.py /root/reference
"""
import json
import os
import re
import sys

Z = {"H": 1, "O": 8}


def main(ref):
    d = os.path.join(ref, "test", "h2oclusters")
    atoms = []
    with open(os.path.join(d, "h2o_20.mol")) as fh:
        on = False
        for line in fh:
            if "geometry" in line:
                on = True
                continue
            if on:
                if "}" in line:
                    break
                p = [x.strip() for x in line.split(",")]
                if len(p) == 4 and p[0].upper() in Z:
                    atoms.append([Z[p[0].upper()], float(p[1]), float(p[2]), float(p[3])])
    energies = {}
    with open(os.path.join(d, "h2o_20.ref")) as fh:
        for line in fh:
            m = re.match(r"Monomer (\d+) energy = (-?\d+\.\d+) Hartree", line.strip())
            if m:
                energies[int(m.group(1))] = float(m.group(2))
    assert len(atoms) == 60 and len(energies) == 20
    mono = [{"Z": [a[0] for a in atoms[3 * k:3 * k + 3]],
             "xyz_angstrom": [a[1:] for a in atoms[3 * k:3 * k + 3]],
             "energy": energies[k + 1]} for k in range(20)]
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "h2o20_monomers.json")
    with open(out, "w") as fh:
        json.dump({"source": "reference test/h2oclusters/h2o_20.mol + h2o_20.ref",
                   "method": "RHF, density fitted (cc-pVDZ / cc-pVDZ-JKFIT, spherical)", "monomers": mono},
                  fh, indent=1)
    print(len(mono), "monomers ->", out)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
