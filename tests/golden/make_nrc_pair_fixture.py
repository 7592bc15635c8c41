"""Writes nrc_pair_water.json: the header of data/pair_nrc1.data (photon
energies k = E / m c^2 from emin to emax in ne steps of log(k - 2), ne, the
number of positron energy-fraction points nb) and, for hydrogen and oxygen
at six of the energies, the tabulated densities of the positron's share of
the kinetic energy, as init_nrc_pair reads them (egsnrc.mortran:7280-7333:
direct-access records of nb 4-byte words, record 1 the header, element Z's
energy ie in record (Z-1) ne + 1 + ie).  Run from the repository root with
the reference tree at /root/reference; the JSON is the committed fixture."""
import json
import os
import struct

DATA = "/root/reference/HEN_HOUSE/data/pair_nrc1.data"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "nrc_pair_water.json")
ENERGIES = (1, 5, 20, 40, 60, 84)   # 1-based


def main():
    raw = open(DATA, "rb").read()
    emin, emax, ne, nb = struct.unpack("<ffii", raw[:16])
    assert raw[16] == 0   # little-endian data
    rec = 4 * nb
    f = {}
    for z in (1, 8):
        f[str(z)] = {}
        for ie in ENERGIES:
            irec = (z - 1) * ne + 1 + (ie - 1)   # 0-based record
            f[str(z)][str(ie)] = list(struct.unpack("<%df" % nb, raw[irec * rec:(irec + 1) * rec]))
    json.dump({"source": "HEN_HOUSE/data/pair_nrc1.data", "emin": emin, "emax": emax,
               "ne": ne, "nb": nb, "f_by_Z": f}, open(OUT, "w"))
    print("wrote", OUT, emin, emax, ne, nb)


if __name__ == "__main__":
    main()
