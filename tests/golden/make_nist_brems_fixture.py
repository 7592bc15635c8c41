"""Writes nist_brems_water.json: the grid of data/nist_brems.data (electron
kinetic energies T in MeV, photon energy fractions k/T) and the scaled
bremsstrahlung cross sections (beta^2/Z^2) k dsigma/dk (mb) of hydrogen and
oxygen on it, as init_nist_brems reads them (egsnrc.mortran:6944-6964).
Run from the repository root with the reference tree at /root/reference;
the JSON is the committed fixture."""
import json
import os

DATA = "/root/reference/HEN_HOUSE/data/nist_brems.data"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "nist_brems_water.json")


def main():
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
    json.dump({"source": "HEN_HOUSE/data/nist_brems.data", "T": energy, "x": x,
               "cs_by_Z": cs, "layout": "cs_by_Z[Z][k][n] at x[k], T[n]"},
              open(OUT, "w"))
    print("wrote", OUT, nmix, kmix)


if __name__ == "__main__":
    main()
