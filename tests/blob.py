"""Reader of the cross-section table blob (format in csrc/egs_tables.h)."""
import struct

import numpy as np


def read_blob(path):
    """Return {name: numpy array} for every record of a table blob."""
    out = {}
    with open(path, "rb") as f:
        data = f.read()
    if data[:8] != b"EGSGPUT1":
        raise ValueError("%s is not a table blob" % path)
    ver, nrec = struct.unpack_from("<II", data, 8)
    off = 16
    for _ in range(nrec):
        name = data[off:off + 48].split(b"\0")[0].decode()
        typ, _pad, count = struct.unpack_from("<IIQ", data, off + 48)
        off += 64
        if typ == 1:
            arr = np.frombuffer(data, dtype="<f8", count=count, offset=off)
            nbytes = 8 * count
        else:
            arr = np.frombuffer(data, dtype="<i4", count=count, offset=off)
            nbytes = 4 * count
        out[name] = arr.copy()
        off += nbytes + (-nbytes % 8)
    return out


def medium_names(blob):
    n = int(blob["nmed"][0])
    return ["".join(chr(c) for c in blob["m%d.name" % m]) for m in range(n)]


def pwlf(blob, m, field, x, grid="g", suffix=""):
    """Evaluate a piecewise-linear table ($SET INTERVAL/$EVALUATE,
    egsnrc.macros:1226-1248) at log-energy x (numpy, FP64); suffix picks
    a variant of the table ("_kn": the Klein-Nishina photon rows)."""
    if grid == "g":
        a, b = blob["m%d.ge1" % m][0], blob["m%d.ge0" % m][0]
    else:
        a, b = blob["m%d.eke1" % m][0], blob["m%d.eke0" % m][0]
    L = (a * x + b).astype(np.int64)
    c0 = blob["m%d.%s0%s" % (m, field, suffix)]
    c1 = blob["m%d.%s1%s" % (m, field, suffix)]
    return c1[L - 1] * x + c0[L - 1]
