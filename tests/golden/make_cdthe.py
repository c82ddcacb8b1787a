"""Convert the reference's test/cdthe.ints (printERI output for He with the
5-function uncontracted 'test' s basis, test/cdthe.mol) into a JSON fixture:
rows (c1, c2, c3, c4, value) where value = multiplier * (c4 c3|c2 c1) with
the multiplier of IntegralEngine::printERI (src/integrals.cpp:203-233).

    python tests/golden/make_cdthe.py /root/reference
"""
import json
import os
import sys


def main(ref):
    rows = []
    with open(os.path.join(ref, "test", "cdthe.ints")) as fh:
        for line in fh:
            p = line.split()
            if len(p) == 5 and all(x.isdigit() for x in p[:4]):
                rows.append([int(p[0]), int(p[1]), int(p[2]), int(p[3]), float(p[4])])
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cdthe_ints.json")
    with open(out, "w") as fh:
        json.dump({"source": "reference test/cdthe.ints", "basis": "test (He: s 0.25, 0.5, 1, 2, 4)",
                   "rows": rows}, fh, indent=0)
    print(len(rows), "rows ->", out)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
