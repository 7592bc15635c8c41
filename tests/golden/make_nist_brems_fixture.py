"""Writes nist_brems_water.json (or, with the argument "nrc",
nrc_brems_water.json): the grid of data/nist_brems.data (data/nrc_brems.data)
-- electron kinetic energies T in MeV, photon energy fractions k/T -- and
the scaled bremsstrahlung cross sections (beta^2/Z^2) k dsigma/dk (mb) of
hydrogen and oxygen on it, as init_nist_brems reads them
(egsnrc.mortran:6933-6964).  Run from the repository root with the
reference tree at /root/reference; the JSON is the committed fixture."""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def main(kind="nist"):
    DATA = "/root/reference/HEN_HOUSE/data/%s_brems.data" % kind
    OUT = os.path.join(HERE, "%s_brems_water.json" % kind)
    lines = open(DATA).read().splitlines()
    pos = 1
    nmix, kmix = (int(v) for v in lines[pos].split())
    pos += 1

    def reals(n):
        nonlocal pos
        out = []
        while len(out) < n:
            out += [float(v) for v in lines[pos].split()]
            pos += 1
        return out[:n]

    energy = reals(nmix)
    x = reals(kmix)
    pos += 1   # the "BREMX.DAT" line
    cs = {}
    for iz in range(1, 101):
        block = reals(nmix * kmix)   # T fastest, then k/T
        if iz in (1, 8):
            cs[str(iz)] = [block[k * nmix:(k + 1) * nmix] for k in range(kmix)]
    json.dump({"source": "HEN_HOUSE/data/%s_brems.data" % kind, "T": energy, "x": x,
               "cs_by_Z": cs, "layout": "cs_by_Z[Z][k][n] at x[k], T[n]"},
              open(OUT, "w"))
    print("wrote", OUT, nmix, kmix)


if __name__ == "__main__":
    main(*sys.argv[1:])
