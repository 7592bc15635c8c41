"""Writes eii_ik_ta_pb.json: from data/eii_ik.data (the EII cross sections
eii_init reads, egsnrc.mortran:7547-7614) the maximum energy of the grid,
the number of bins and, for tantalum and lead, each shell's cross-section
maximum (barn/atom) and its 250 values relative to it (the cross section at
the upper edge of each log-spaced bin from the binding energy to the
maximum energy).  Run from the repository root with the reference tree at
/root/reference; the JSON is the committed fixture."""
import json
import os

DATA = "/root/reference/HEN_HOUSE/data/eii_ik.data"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "eii_ik_ta_pb.json")


def main():
    toks_lines = open(DATA).read().splitlines()
    nskip = int(toks_lines[0].split()[0])
    pos = 1 + nskip
    emax, nbin = toks_lines[pos].split()[:2]
    emax, nbin = float(emax), int(nbin)
    pos += 1

    def reals(n):
        nonlocal pos
        out = []
        while len(out) < n:
            out += [float(v) for v in toks_lines[pos].split()]
            pos += 1
        return out[:n]

    shells = {}
    for _ in range(100):
        z, nsh = (int(v) for v in toks_lines[pos].split()[:2])
        pos += 1
        for ish in range(1, nsh + 1):
            fmax = reals(1)[0]
            aux = reals(nbin)
            if z in (73, 82):
                shells["%d.%d" % (z, ish)] = {"fmax": fmax, "rel": aux}
    json.dump({"source": "HEN_HOUSE/data/eii_ik.data", "emax": emax, "nbin": nbin,
               "shells": shells}, open(OUT, "w"))
    print("wrote", OUT, emax, nbin, sorted(shells))


if __name__ == "__main__":
    main()
