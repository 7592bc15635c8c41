"""Write tests/golden/photon_xs_521icru.json: photon interaction data of
the media of egsnrc_amd/data/521icru.egstab computed here, independently of
the table builder (egsnrc_amd/csrc/egs_tables.cpp), from the EGSnrc data
files, so that a test can hold the table blob to them.

Run in a checkout that has the EGSnrc tree (default /root/reference):

    python tests/golden/make_photon_xs.py [HEN_HOUSE]

Definitions (the state egs_init_user_photon leaves behind,
src/egsnrc.mortran:7973-8284, with IRAYLR = 1 and IBCMP = 3):
  * per element, XCOM photo-absorption, Rayleigh, pair and triplet cross
    sections (data/<prefix>_*.data, barn/atom; prefix xcom unless given),
    interpolated linearly in
    (ln E, ln sigma), the pair (triplet) data after dividing by
    (1 - 2m/E)^3 ((1 - 4m/E)^3) (egsi_get_data, :8776-8838);
  * Compton: the Klein-Nishina total cross section per electron (Heitler's
    closed form) times Z, times the bound-Compton ratio of
    data/compton_sigma.data for the mixture (interpolated linearly in
    (ln E, ln ratio)), as init_compton corrects it for IBCMP > 1;
  * a medium's atom fractions PZ and density from the PEGS4 file
    (pegs4/data/521icru.pegs4dat);
  * gmfp = 1/(n sigma) without Rayleigh (cm), gbr1 = (pair + triplet)/sigma,
    gbr2 = gbr1 + Compton/sigma, cohe = sigma/(sigma + Rayleigh), with
    n = rho N_A / sum(PZ A).
Energies: the XCOM photo-absorption grid points of all elements of the
medium inside the table range, at least 2 % away from any absorption edge.
"""
import json
import math
import os
import sys

HH = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/HEN_HOUSE"
# the photon data set ("Photon cross sections": xcom, epdl or si); the
# xcom fixture covers every medium of the 521icru blob, the others water
PREFIX = sys.argv[2] if len(sys.argv) > 2 else "xcom"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                   "photon_xs_521icru.json" if PREFIX == "xcom"
                   else "photon_xs_521icru_%s.json" % PREFIX)
MEDIA = (["H2O521ICRU", "LUNG521ICRU", "ICRPBONE521ICRU", "AIR521ICRU", "AL521ICRU"]
         if PREFIX == "xcom" else ["H2O521ICRU"])
RM = 0.5109989461
R0 = 2.8179403262e-13          # classical electron radius (cm)
NA = 6.02214076e23             # Avogadro's number (1/mol)


def numbers(path):
    with open(path) as f:
        return [float(t) for t in f.read().replace("D", "E").split()]


def element_tables(path):
    """{Z: (lnE list, ln sigma list)} for Z = 1..100."""
    v = numbers(path)
    out, k = {}, 0
    for z in range(1, 101):
        n = int(v[k]); k += 1
        pts = v[k:k + 2 * n]; k += 2 * n
        out[z] = (pts[0::2], pts[1::2])
    return out


def interp_loglog(le, lf, x):
    """Linear interpolation in (ln E, ln sigma); 0 below the table."""
    if x < le[0]:
        return None
    for j in range(len(le) - 1):
        if le[j] <= x < le[j + 1]:
            p = (x - le[j]) / (le[j + 1] - le[j])
            return math.exp(p * lf[j + 1] + (1 - p) * lf[j])
    return math.exp(lf[-1])


def threshold_xs(tab, z, e, eth):
    """Pair/triplet: interpolate ln(sigma / (1 - eth/E)^3), zero below eth."""
    if e <= eth:
        return 0.0
    le, lf = tab[z]
    g = [f - 3 * math.log(1 - eth / math.exp(x)) for x, f in zip(le, lf)]
    le = [math.log(eth)] + list(le)
    g = [g[0]] + g
    s = interp_loglog(le, g, math.log(e))
    return (s or 0.0) * (1 - eth / e) ** 3


def kn_per_electron(e):
    """Klein-Nishina total cross section per electron (barn)."""
    k = e / RM
    a = 1 + 2 * k
    s = 2 * math.pi * R0 * R0 * ((1 + k) / k ** 2 * (2 * (1 + k) / a - math.log(a) / k)
                                 + math.log(a) / (2 * k) - (1 + 3 * k) / a ** 2)
    return s * 1e24


def pegs_media(path):
    """{name: (rho, [(Z, A, PZ), ...])} from the PEGS4 file's headers."""
    out = {}
    lines = open(path).read().splitlines()
    for i, l in enumerate(lines):
        if "MEDIUM=" not in l:
            continue
        name = l.split("MEDIUM=")[1].split(",")[0].strip()
        comp = lines[i + 1]
        rho = float(comp.split("RHO=")[1].split(",")[0])
        ne = int(comp.split("NE=")[1].split(",")[0])
        els = []
        for k in range(ne):
            t = lines[i + 2 + k]
            z = float(t.split("Z=")[1].split(",")[0])
            a = float(t.split("A=")[1].split(",")[0])
            pz = float(t.split("PZ=")[1].split(",")[0])
            els.append((int(round(z)), a, pz))
        ae, ap, ue, up = [float(x) for x in lines[i + 2 + ne].split()[1:5]]
        out[name] = (rho, els, ap, up)
    return out


def main():
    data = os.path.join(HH, "data")
    photo = element_tables(os.path.join(data, PREFIX + "_photo.data"))
    rayl = element_tables(os.path.join(data, PREFIX + "_rayleigh.data"))
    pair = element_tables(os.path.join(data, PREFIX + "_pair.data"))
    trip = element_tables(os.path.join(data, PREFIX + "_triplet.data"))
    cs = numbers(os.path.join(data, "compton_sigma.data"))
    cs_emin, cs_emax, cs_n = cs[0], cs[1], int(cs[2])
    cs_el = {z: cs[3 + (z - 1) * cs_n: 3 + z * cs_n] for z in range(1, 101)}
    media = pegs_media(os.path.join(HH, "pegs4", "data", "521icru.pegs4dat"))
    result = {"source": "make_photon_xs.py", "media": {}}
    for name in MEDIA:
        rho, els, ap, up = media[name]
        sum_a = sum(pz * a for z, a, pz in els)
        sum_z = sum(pz * z for z, a, pz in els)
        n_per_barn = rho * NA / sum_a * 1e-24
        edges = []
        for z, a, pz in els:
            le = photo[z][0]
            edges += [math.exp(le[j]) for j in range(1, len(le)) if le[j] - le[j - 1] < 1e-5]
        grid = sorted(set(round(math.exp(x), 9) for z, a, pz in els for x in photo[z][0]))
        energies = [e for e in grid if ap * 1.02 < e < up / 1.02 and
                    all(abs(e / ed - 1) > 0.02 for ed in edges)]
        # mixture bound-Compton ratio on its own log grid
        dle = math.log(cs_emax / cs_emin) / (cs_n - 1)
        ratio = [math.log(sum(pz * z * cs_el[z][j] for z, a, pz in els) / sum_z)
                 for j in range(cs_n)]
        rows = []
        for e in energies:
            le = math.log(e)
            s_ph = sum(pz * interp_loglog(*photo[z], le) for z, a, pz in els)
            s_ra = sum(pz * interp_loglog(*rayl[z], le) for z, a, pz in els)
            s_pp = sum(pz * (threshold_xs(pair, z, e, 2 * RM) + threshold_xs(trip, z, e, 4 * RM))
                       for z, a, pz in els)
            if e <= cs_emin:
                bc = math.exp(ratio[0])
            elif e < cs_emax:
                aj = 1 + math.log(e / cs_emin) / dle
                j = int(aj)
                aj -= j
                bc = math.exp(ratio[j - 1] * (1 - aj) + ratio[j] * aj)
            else:
                bc = 1.0
            s_co = sum_z * kn_per_electron(e) * bc
            sigma = s_ph + s_pp + s_co
            rows.append(dict(e=e, gmfp=1 / (sigma * n_per_barn), gbr1=s_pp / sigma,
                             gbr2=(s_pp + s_co) / sigma, cohe=sigma / (sigma + s_ra),
                             mu_rho_total=(sigma + s_ra) * n_per_barn / rho,
                             compton_ratio=bc))
        result["media"][name] = dict(rho=rho, rows=rows)
    with open(OUT, "w") as f:
        json.dump(result, f, indent=0)
    print("wrote", OUT, {k: len(v["rows"]) for k, v in result["media"].items()})


if __name__ == "__main__":
    main()
