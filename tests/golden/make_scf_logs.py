"""Extract the RHF iteration tables (iteration, energy, |delta E|, delta D)
of the reference's recorded runs into a JSON fixture:

  test/h2otest.out_1  water, cc-pVDZ with 6 Cartesian d functions
                      (geometry of water_h2otest(), bohr), in-core
  test/benzene.out_1  benzene, cc-pVDZ 6d (test/benzene-dz.mol), direct

    python tests/golden/make_scf_logs.py /root/reference
"""
import json
import os
import sys


def table(path):
    rows, on = [], False
    with open(path) as fh:
        for line in fh:
            if "Iteration" in line and "Energy" in line:
                on = True
                continue
            if on:
                p = line.split()
                if len(p) == 5 and p[0].isdigit():
                    rows.append([int(p[0]), float(p[1]), float(p[2]), float(p[3])])
                elif rows and not line.strip():
                    break
    return rows


def main(ref):
    out = {"h2otest": {"source": "reference test/h2otest.out_1", "basis": "cc-pVDZ, 6d",
                       "rows": table(os.path.join(ref, "test", "h2otest.out_1"))},
           "benzene": {"source": "reference test/benzene.out_1", "basis": "cc-pVDZ, 6d",
                       "rows": table(os.path.join(ref, "test", "benzene.out_1"))}}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "scf_logs.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1)
    print({k: len(v["rows"]) for k, v in out.items()}, "->", path)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
