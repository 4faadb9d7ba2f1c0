"""Writes w7x_modes.txt: the W7-X standard (EIM) boundary modes, 288 rows of
(m, n, Rmn, Zmn) in the VMEC form R = sum Rmn cos(m theta - n nfp phi),
Z = sum Zmn sin(m theta - n nfp phi), nfp = 5, read from the data table of the
reference's test/w7x_grf.m (function w7x_modes).  Run from the repository root:

    python tests/golden/make_w7x_modes.py /root/reference
"""
import os
import sys


def main(ref):
    src = open(os.path.join(ref, "test", "w7x_grf.m")).read().splitlines()
    start = next(i for i, l in enumerate(src) if l.startswith("function D4 = w7x_modes"))
    rows, on = [], False
    for line in src[start:]:
        s = line.strip()
        if s.startswith("D = ["):
            on = True
            continue
        if not on:
            continue
        end = s.endswith("];")
        s = s.rstrip("];").rstrip(";").strip()
        if s:
            m, n, r, z = s.split()
            rows.append((int(m), int(n), r, z))
        if end:
            break
    assert len(rows) == 288, len(rows)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "w7x_modes.txt")
    with open(out, "w") as f:
        f.write("# W7-X standard (EIM) boundary modes, nfp = 5: m n Rmn Zmn (test/w7x_grf.m, w7x_modes)\n")
        for m, n, r, z in rows:
            f.write("%d %d %s %s\n" % (m, n, r, z))
    print("wrote", out, len(rows), "modes")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
