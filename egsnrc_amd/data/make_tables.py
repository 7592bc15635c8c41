"""Regenerate the committed cross-section table blobs from the EGSnrc data.

Runs where an EGSnrc HEN_HOUSE tree is available (default
/root/reference/HEN_HOUSE); the blobs it writes are what the engine loads.
"""
import argparse
import ctypes as C
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from egsnrc_amd import _lib  # noqa: E402

BLOBS = {
    # name: (pegs4 file, media[, brems_xsec[, photon data set[, EII data set]]])
    "521icru.egstab": ("521icru.pegs4dat",
                       ["H2O521ICRU", "LUNG521ICRU", "ICRPBONE521ICRU",
                        "AIR521ICRU", "AL521ICRU"]),
    "700icru.egstab": ("700icru.pegs4dat",
                       ["H2O700ICRU", "ICRPBONE700ICRU", "AIR700ICRU", "LUNG700ICRU"]),
    # high-Z media of the tutorial data set (AE 0.7, AP 0.01 MeV): tantalum
    # of tutor2, lead for the atomic relaxation cascade, and water
    "tutor.egstab": ("tutor_data.pegs4dat", ["TA", "PB", "H2O"]),
    # the NIST bremsstrahlung cross sections (ibr_nist = 1) and the NRC pair
    # tables for water and bone
    "521icru_nist.egstab": ("521icru.pegs4dat", ["H2O521ICRU", "ICRPBONE521ICRU"], 1),
    # the NRC bremsstrahlung cross sections (ibr_nist = 2) for water and bone
    "521icru_nrcbrems.egstab": ("521icru.pegs4dat", ["H2O521ICRU", "ICRPBONE521ICRU"], 2),
    # water from the EPDL photon data set ("Photon cross sections = epdl")
    "521icru_epdl.egstab": ("521icru.pegs4dat", ["H2O521ICRU"], 0, "epdl"),
    # the tutorial media with electron impact ionisation (the ik data set)
    "tutor_eii.egstab": ("tutor_data.pegs4dat", ["TA", "PB", "H2O"], 0, "xcom", "ik"),
}


def build(hen_house, name, pegs, media, out_dir=HERE, estepe=0.25, ximax=0.5, brems_xsec=0,
          pair_xsec=None, photon_xsec="xcom", eii=""):
    L = _lib.lib()
    L.egsgpu_tables_build.argtypes = [C.c_char_p, C.c_char_p, C.c_int32,
                                      C.POINTER(C.c_char_p), C.c_void_p,
                                      C.c_char_p]
    arr = (C.c_char_p * len(media))(*[m.encode() for m in media])
    if pair_xsec is None:
        pair_xsec = 1 if brems_xsec == 1 else 0   # the NIST blob carries the NRC pair tables too
    tp = _lib.default_transport(estepe=estepe, ximax=ximax, brems_xsec=brems_xsec,
                                pair_xsec=pair_xsec, photon_xsec=photon_xsec.encode(),
                                eii_xsec=eii.encode())
    out = os.path.join(out_dir, name)
    _lib.check(L.egsgpu_tables_build(
        hen_house.encode(), os.path.join(hen_house, "pegs4", "data", pegs).encode(),
        len(media), arr, C.cast(C.pointer(tp), C.c_void_p), out.encode()))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--hen-house", default="/root/reference/HEN_HOUSE")
    ap.add_argument("--out-dir", default=HERE)
    a = ap.parse_args()
    for name, spec in BLOBS.items():
        pegs, media = spec[0], spec[1]
        nist = spec[2] if len(spec) > 2 else 0
        px = spec[3] if len(spec) > 3 else "xcom"
        eii = spec[4] if len(spec) > 4 else ""
        print("wrote", build(a.hen_house, name, pegs, media, a.out_dir, brems_xsec=nist,
                             photon_xsec=px, eii=eii))


if __name__ == "__main__":
    main()
