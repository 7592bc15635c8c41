"""Reference file formats around the voxel dose path.

* ``.egsphant`` voxel phantoms (DOSXYZnrc / egs++ EGS_XYZGeometry,
  egs++/geometry/egs_nd_geometry/egs_nd_geometry.cpp:596-710): number of
  media, media names, ESTEPE values (ignored), nx ny nz, the plane
  positions, one character per voxel giving the medium (z slices of y rows
  of x characters, '1'..'9' then 'A'..), and the mass density of every
  voxel (x fastest).  Plain text or gzip-compressed.
* CT ramps (the same file, :339-357): ``name rho_min rho_max rho_default``
  per line; a voxel of density rho gets the first medium with
  rho_min <= rho < rho_max and the relative density rho / rho_default
  (:740-764).
* ``.3ddose`` dose files (EGS_DoseScoring::outputDoseFile,
  egs++/ausgab_objects/egs_dose_scoring/egs_dose_scoring.cpp:416-479):
  nx ny nz, the planes, the doses and their relative uncertainties.
* A ``.3ddose`` comparator: per-voxel z-scores of two dose files with
  chi^2/dof and the fraction beyond 2 sigma.
"""
import gzip

import numpy as np

_MEDIUM_CHARS = "123456789ABCDEFGHIJKLMNOPQRSTUVWXYZ"


def _open_text(path):
    with open(path, "rb") as f:
        magic = f.read(2)
    if magic == b"\x1f\x8b":
        return gzip.open(path, "rt")
    return open(path, "r")


def read_egsphant(path):
    """Returns dict(media, xb, yb, zb, medium (0-based, x fastest),
    density (g/cm^3, x fastest))."""
    with _open_text(path) as f:
        nmed = int(f.readline().split()[0])
        media = [f.readline().strip() for _ in range(nmed)]
        toks = []
        need = nmed + 3
        while len(toks) < need:
            toks += f.readline().split()
        nx, ny, nz = (int(t) for t in toks[nmed:nmed + 3])
        rest = toks[nmed + 3:]

        def read_floats(n, have):
            vals = list(have)
            while len(vals) < n:
                line = f.readline()
                if not line:
                    raise ValueError("%s: file ends early" % path)
                vals += line.split()
            return np.array(vals[:n], dtype=np.float64), vals[n:]

        xb, rest = read_floats(nx + 1, rest)
        yb, rest = read_floats(ny + 1, rest)
        zb, rest = read_floats(nz + 1, rest)
        if rest:
            raise ValueError("%s: unexpected data after the planes" % path)
        med = np.empty((nz, ny, nx), dtype=np.int32)
        for iz in range(nz):
            iy = 0
            while iy < ny:
                line = f.readline()
                if not line:
                    raise ValueError("%s: medium matrix ends early" % path)
                line = line.strip()
                if not line:
                    continue
                if len(line) < nx:
                    raise ValueError("%s: short medium row" % path)
                med[iz, iy] = [_MEDIUM_CHARS.index(c) for c in line[:nx]]
                iy += 1
        vals = f.read().split()
        if len(vals) < nx * ny * nz:
            raise ValueError("%s: density matrix ends early" % path)
        rho = np.array(vals[:nx * ny * nz], dtype=np.float64)
    if med.max() >= nmed:
        raise ValueError("%s: medium index beyond the %d media" % (path, nmed))
    return dict(media=media, xb=xb, yb=yb, zb=zb, medium=med.reshape(-1),
                density=rho, shape=(nx, ny, nz))


def write_egsphant(path, media, xb, yb, zb, medium, density, estepe=0.25):
    nx, ny, nz = len(xb) - 1, len(yb) - 1, len(zb) - 1
    med = np.asarray(medium).reshape(nz, ny, nx)
    rho = np.asarray(density, dtype=np.float64).reshape(nz, ny, nx)
    with open(path, "w") as f:
        f.write("%d\n" % len(media))
        for m in media:
            f.write(m + "\n")
        f.write("\t".join("%g" % estepe for _ in media) + "\n")
        f.write("%d\t%d\t%d\n" % (nx, ny, nz))
        for b in (xb, yb, zb):
            f.write("\t".join("%.10g" % v for v in b) + "\n")
        for iz in range(nz):
            for iy in range(ny):
                f.write("".join(_MEDIUM_CHARS[i] for i in med[iz, iy]) + "\n")
            f.write("\n")
        for iz in range(nz):
            for iy in range(ny):
                f.write("\t".join("%.6g" % v for v in rho[iz, iy]) + "\n")
            f.write("\n")


def read_ramp(path):
    """CT ramp: list of (name, rho_min, rho_max, rho_default)."""
    out = []
    with open(path) as f:
        for line in f:
            t = line.split()
            if len(t) < 4:
                break
            out.append((t[0], float(t[1]), float(t[2]), float(t[3])))
    if not out:
        raise ValueError("%s: no media in the CT ramp" % path)
    return out


def apply_ramp(density, ramp, epsilon=2.0 ** -50):
    """Media (index into ramp) and relative densities from mass densities.
    Voxels outside every range get -1 (vacuum in the reference)."""
    rho = np.asarray(density, dtype=np.float64)
    med = np.full(rho.shape, -1, dtype=np.int32)
    rel = np.ones(rho.shape)
    for i, (_, lo, hi, rdef) in enumerate(ramp):
        sel = (med < 0) & (rho >= lo) & (rho < hi)
        med[sel] = i
        if rdef > 0:
            r = rho[sel] / rdef
            rel[sel] = np.where(np.abs(r - 1) > epsilon, r, 1.0)
    return med, rel


RM = 0.5109989461


def read_egsphsp(path):
    """An EGSnrc MODE0/MODE2 phase-space file (the records EGS_PhspSource
    reads, egs++/sources/egs_phsp_source/egs_phsp_source.cpp:79-431).
    Returns a dict of header values and per-particle arrays: latch, e
    (kinetic, MeV), x, y, u, v, wt (sign of w), q and new_history."""
    raw = open(path, "rb").read()
    mode = raw[:5]
    if mode not in (b"MODE0", b"MODE2"):
        raise ValueError("%s is not a MODE0 or MODE2 phase-space file" % path)
    recl = 32 if mode == b"MODE2" else 28
    for order in ("<", ">"):
        n, nphot = np.frombuffer(raw[5:13], dtype=order + "i4")
        emax, emin, pinc = np.frombuffer(raw[13:25], dtype=order + "f4")
        if n > 0 and 0 <= nphot <= n and 0 <= emin <= emax and pinc >= 0:
            break
    else:
        raise ValueError("%s: meaningless phase-space header" % path)
    n = int(n)
    fields = [("latch", order + "i4"), ("e", order + "f4"), ("x", order + "f4"),
              ("y", order + "f4"), ("u", order + "f4"), ("v", order + "f4"),
              ("wt", order + "f4")]
    if recl == 32:
        fields.append(("zlast", order + "f4"))
    rec = np.frombuffer(raw[recl:recl * (n + 1)], dtype=np.dtype(fields))
    latch = rec["latch"].astype(np.int64) & 0xFFFFFFFF
    q = np.where(latch & 1073741824, -1, np.where(latch & 536870912, 1, 0)).astype(np.int8)
    e = rec["e"].astype(np.float64)
    new_history = e < 0
    e = np.abs(e) - np.where(q != 0, RM, 0.0)
    return {"mode2": recl == 32, "n": n, "nphot": int(nphot), "emax": float(emax),
            "emin": float(emin), "pinc": float(pinc), "latch": latch, "e": e,
            "x": rec["x"].copy(), "y": rec["y"].copy(), "u": rec["u"].copy(),
            "v": rec["v"].copy(), "wt": rec["wt"].copy(), "q": q, "new_history": new_history}


def write_egsphsp(path, e, x, y, u, v, wt, q, pinc, new_history=None, latch=None):
    """Write a MODE0 file: e kinetic (MeV), wt with the sign of w, q the
    charge; new_history marks the first particle of each primary history
    (negative energy), latch bits 29/30 are set from q."""
    n = len(e)
    q = np.asarray(q, dtype=np.int64)
    lat = np.zeros(n, np.int64) if latch is None else np.asarray(latch, np.int64)
    lat = lat | np.where(q < 0, 1073741824, 0) | np.where(q > 0, 536870912, 0)
    etot = np.asarray(e, np.float64) + np.where(q != 0, RM, 0.0)
    if new_history is not None:
        etot = np.where(new_history, -etot, etot)
    ekin = np.asarray(e, np.float64)
    nphot = int(np.sum(q == 0))
    rec = np.zeros(n, dtype=[("latch", "<i4"), ("e", "<f4"), ("x", "<f4"), ("y", "<f4"),
                             ("u", "<f4"), ("v", "<f4"), ("wt", "<f4")])
    rec["latch"] = lat.astype(np.uint32).view(np.int32) if lat.dtype != np.int32 else lat
    rec["e"] = etot
    rec["x"], rec["y"], rec["u"], rec["v"], rec["wt"] = x, y, u, v, wt
    head = b"MODE0" + np.array([n, nphot], "<i4").tobytes() + np.array(
        [ekin.max(), ekin.min(), pinc], "<f4").tobytes()
    with open(path, "wb") as f:
        f.write(head + b"\0" * (28 - len(head)))
        f.write(rec.tobytes())


def read_3ddose(path):
    """Returns (xb, yb, zb, dose, rel_err), x fastest."""
    with _open_text(path) as f:
        toks = f.read().split()
    nx, ny, nz = (int(t) for t in toks[:3])
    v = np.array(toks[3:], dtype=np.float64)
    k = 0
    xb = v[k:k + nx + 1]; k += nx + 1
    yb = v[k:k + ny + 1]; k += ny + 1
    zb = v[k:k + nz + 1]; k += nz + 1
    n = nx * ny * nz
    if len(v) < k + 2 * n:
        raise ValueError("%s: truncated .3ddose" % path)
    return xb, yb, zb, v[k:k + n], v[k + n:k + 2 * n]


def compare_3ddose(a, b, frac=0.1):
    """Compare two .3ddose results (paths or read_3ddose tuples) voxel by
    voxel over voxels above frac * the larger maximum dose:
    z = (Da - Db) / sqrt(sa^2 + sb^2) with absolute uncertainties
    s = D * rel_err.  Returns dict(voxels, chi2_per_dof,
    frac_beyond_2sigma, max_abs_z, mean_rel_diff)."""
    if isinstance(a, str):
        a = read_3ddose(a)
    if isinstance(b, str):
        b = read_3ddose(b)
    for k in range(3):
        if len(a[k]) != len(b[k]) or not np.allclose(a[k], b[k], rtol=1e-5, atol=1e-6):
            raise ValueError("the two dose files have different voxel grids")
    da, ea = a[3], a[4] * a[3]
    db, eb = b[3], b[4] * b[3]
    dmax = max(da.max(), db.max())
    m = (np.maximum(da, db) > frac * dmax) & (ea * ea + eb * eb > 0)
    z = (da[m] - db[m]) / np.sqrt(ea[m] ** 2 + eb[m] ** 2)
    return {"voxels": int(m.sum()),
            "chi2_per_dof": float(np.mean(z * z)) if z.size else 0.0,
            "frac_beyond_2sigma": float(np.mean(np.abs(z) > 2)) if z.size else 0.0,
            "max_abs_z": float(np.max(np.abs(z))) if z.size else 0.0,
            "mean_rel_diff": float(np.mean((da[m] - db[m]) / db[m])) if z.size else 0.0}


def main(argv=None):
    """python -m egsnrc_amd.phantom_io compare a.3ddose b.3ddose"""
    import argparse
    import json
    ap = argparse.ArgumentParser(prog="egsnrc_amd.phantom_io")
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("compare", help="compare two .3ddose files")
    c.add_argument("a")
    c.add_argument("b")
    c.add_argument("--frac", type=float, default=0.1)
    args = ap.parse_args(argv)
    print(json.dumps(compare_3ddose(args.a, args.b, args.frac)))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
