// egs_tables.cpp -- table builder and blob I/O (host, FP64).
//
// Each builder step cites the reference routine whose result it
// reproduces; the arithmetic is restated, the data files are read in the
// record layout the reference expects.
#include "egs_tables.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>
#include <limits>
#include <sstream>

#include "../../include/egsgpu.h"
#include "egsgpu_common.h"

namespace egsgpu {

const std::vector<double> &Blob::F(const std::string &k) const {
    auto it = f.find(k);
    if (it == f.end()) throw std::runtime_error("table blob has no field " + k);
    return it->second;
}
const std::vector<int32_t> &Blob::I(const std::string &k) const {
    auto it = i.find(k);
    if (it == i.end()) throw std::runtime_error("table blob has no field " + k);
    return it->second;
}

// ---------------------------------------------------------------- blob I/O
int blob_write(const Blob &b, const std::string &path) {
    FILE *fp = fopen(path.c_str(), "wb");
    if (!fp) { set_error("cannot write %s", path.c_str()); return EGSGPU_E_IO; }
    const char magic[8] = {'E', 'G', 'S', 'G', 'P', 'U', 'T', '1'};
    uint32_t ver = 1, nrec = (uint32_t)(b.f.size() + b.i.size());
    fwrite(magic, 1, 8, fp);
    fwrite(&ver, 4, 1, fp);
    fwrite(&nrec, 4, 1, fp);
    auto rec = [&](const std::string &name, uint32_t type, uint64_t count,
                   const void *data, size_t esz) {
        char nm[48] = {0};
        strncpy(nm, name.c_str(), 47);
        uint32_t pad = 0;
        fwrite(nm, 1, 48, fp);
        fwrite(&type, 4, 1, fp);
        fwrite(&pad, 4, 1, fp);
        fwrite(&count, 8, 1, fp);
        fwrite(data, esz, count, fp);
        size_t bytes = esz * count;
        static const char zeros[8] = {0};
        if (bytes % 8) fwrite(zeros, 1, 8 - bytes % 8, fp);
    };
    for (auto &kv : b.f) rec(kv.first, 1, kv.second.size(), kv.second.data(), 8);
    for (auto &kv : b.i) rec(kv.first, 2, kv.second.size(), kv.second.data(), 4);
    bool ok = !ferror(fp);
    fclose(fp);
    if (!ok) { set_error("write error on %s", path.c_str()); return EGSGPU_E_IO; }
    return EGSGPU_OK;
}

double bca_blcmin(const Blob &b, const std::vector<int> &media, double ecutmn) {
    // the blob keeps blcc and xcc as mscati leaves them after this point:
    // blcc times exp(2 gamma_E - 1) and xcc squared
    const double beta2 = (ecutmn - kRM) * (ecutmn + kRM) / (ecutmn * ecutmn);
    double best = 1e30;
    for (int m : media) {
        const double blcc = b.f1(Blob::mk(m, "blcc")) / 1.16699413758864573;
        const double xcc2 = b.f1(Blob::mk(m, "xcc"));
        const double t = blcc * beta2 * ecutmn * ecutmn / xcc2;
        best = std::min(best, std::log(t / std::log(t)));
    }
    return best;
}

int blob_read(const std::string &path, Blob &b) {
    std::ifstream in(path, std::ios::binary);
    if (!in) { set_error("cannot open table blob %s", path.c_str()); return EGSGPU_E_IO; }
    char magic[8];
    uint32_t ver, nrec;
    in.read(magic, 8);
    in.read((char *)&ver, 4);
    in.read((char *)&nrec, 4);
    if (!in || memcmp(magic, "EGSGPUT1", 8) != 0 || ver != 1) {
        set_error("%s is not a version-1 table blob", path.c_str());
        return EGSGPU_E_FORMAT;
    }
    for (uint32_t r = 0; r < nrec; r++) {
        char nm[48];
        uint32_t type, pad;
        uint64_t count;
        in.read(nm, 48);
        in.read((char *)&type, 4);
        in.read((char *)&pad, 4);
        in.read((char *)&count, 8);
        if (!in || count > (1ull << 32)) { set_error("truncated blob %s", path.c_str()); return EGSGPU_E_FORMAT; }
        nm[47] = 0;
        size_t bytes;
        if (type == 1) {
            std::vector<double> v(count);
            in.read((char *)v.data(), 8 * count);
            b.f[nm] = std::move(v);
            bytes = 8 * count;
        } else if (type == 2) {
            std::vector<int32_t> v(count);
            in.read((char *)v.data(), 4 * count);
            b.i[nm] = std::move(v);
            bytes = 4 * count;
        } else {
            set_error("bad record type in %s", path.c_str());
            return EGSGPU_E_FORMAT;
        }
        if (bytes % 8) in.ignore(8 - bytes % 8);
        if (!in) { set_error("truncated blob %s", path.c_str()); return EGSGPU_E_FORMAT; }
    }
    return EGSGPU_OK;
}

// ------------------------------------------------------------ PEGS4 read
struct PegsMedium {
    std::string name;
    double rho = 0;
    int nne = 0;
    std::vector<double> z, wa, pz, rhoz;
    double rlc, ae, ap, ue, up;
    int msge, mge, mseke, meke, mleke, mcmfp, mrange, irayl;
    double delcm, alphi[2], bpar[2], delpos[2];
    double xr0, teff0, blcc, xcc;
    double eke0, eke1;
    std::vector<double> esig0, esig1, psig0, psig1, ededx0, ededx1, pdedx0,
        pdedx1, ebr10, ebr11, pbr10, pbr11, pbr20, pbr21, tmxs0, tmxs1;
};

static double field_after(const std::string &s, const char *key) {
    size_t p = s.find(key);
    if (p == std::string::npos) throw std::runtime_error(std::string("PEGS4 line lacks ") + key);
    return std::stod(s.substr(p + strlen(key)));
}

static std::string trim(const std::string &s) {
    size_t a = s.find_first_not_of(" \t\r"), b = s.find_last_not_of(" \t\r");
    return a == std::string::npos ? "" : s.substr(a, b - a + 1);
}

// HATCH medium read, src/egsnrc.mortran:2011-2176, and DUNIT=1 scaling
// :2185-2200 (lengths in cm; PEGS4 stores them in radiation lengths).
static void read_pegs(const std::string &path, const std::vector<std::string> &want,
                      std::vector<PegsMedium> &out) {
    out.assign(want.size(), PegsMedium());
    std::vector<bool> found(want.size(), false);
    Records r(path);
    if (!r.ok()) throw std::runtime_error("cannot open PEGS4 file " + path);
    std::string s;
    size_t nfound = 0;
    while (nfound < want.size() && r.line(s)) {
        if (s.compare(0, 8, " MEDIUM=") != 0) continue;
        std::string nm = trim(s.substr(8, 24));
        size_t c = nm.find(',');
        if (c != std::string::npos) nm = trim(nm.substr(0, c));
        int im = -1;
        for (size_t k = 0; k < want.size(); k++)
            if (!found[k] && want[k] == nm) { im = (int)k; break; }
        if (im < 0) continue;
        PegsMedium &m = out[im];
        m.name = nm;
        r.line(s);
        m.rho = field_after(s, "RHO=");
        m.nne = (int)field_after(s, "NE=");
        for (int e = 0; e < m.nne; e++) {
            r.line(s);
            m.z.push_back(field_after(s, ",Z="));
            m.wa.push_back(field_after(s, ",A="));
            m.pz.push_back(field_after(s, ",PZ="));
            m.rhoz.push_back(field_after(s, ",RHOZ="));
        }
        auto v = r.reals(5);
        m.rlc = v[0]; m.ae = v[1]; m.ap = v[2]; m.ue = v[3]; m.up = v[4];
        auto iv = r.ints(8);
        m.msge = iv[0]; m.mge = iv[1]; m.mseke = iv[2]; m.meke = iv[3];
        m.mleke = iv[4]; m.mcmfp = iv[5]; m.mrange = iv[6]; m.irayl = iv[7];
        r.reals(36);                      // DL1..DL6: recomputed by fix_brems
        v = r.reals(7);
        m.delcm = v[0];
        for (int k = 0; k < 2; k++) { m.alphi[k] = v[1 + 3 * k]; m.bpar[k] = v[2 + 3 * k]; m.delpos[k] = v[3 + 3 * k]; }
        v = r.reals(4);
        m.xr0 = v[0]; m.teff0 = v[1]; m.blcc = v[2]; m.xcc = v[3];
        v = r.reals(2);
        m.eke0 = v[0]; m.eke1 = v[1];
        int n = m.meke;
        v = r.reals(16 * (size_t)n);
        std::vector<double> *cols[16] = {&m.esig0, &m.esig1, &m.psig0, &m.psig1,
                                         &m.ededx0, &m.ededx1, &m.pdedx0, &m.pdedx1,
                                         &m.ebr10, &m.ebr11, &m.pbr10, &m.pbr11,
                                         &m.pbr20, &m.pbr21, &m.tmxs0, &m.tmxs1};
        for (int k = 0; k < 16; k++) cols[k]->resize(n);
        for (int i = 0; i < n; i++)
            for (int k = 0; k < 16; k++) (*cols[k])[i] = v[16 * i + k];
        r.reals(3);                       // EBINDA, GE0, GE1 (PEGS4 photon grid)
        r.reals(6 * (size_t)m.mge);       // PEGS4 GMFP/GBR1/GBR2: rebuilt from XCOM
        if (m.irayl == 1) {
            auto ngr = r.ints(1);
            r.reals(2);
            r.reals(2 * (size_t)ngr[0]);
            r.reals(2 * (size_t)m.mge);
        }
        // DUNIT = 1 cm scaling
        double dfact = m.rlc, dfacti = 1.0 / dfact;
        for (int i = 0; i < n; i++) {
            m.esig0[i] *= dfacti; m.esig1[i] *= dfacti;
            m.psig0[i] *= dfacti; m.psig1[i] *= dfacti;
            m.ededx0[i] *= dfacti; m.ededx1[i] *= dfacti;
            m.pdedx0[i] *= dfacti; m.pdedx1[i] *= dfacti;
            m.tmxs0[i] *= dfact; m.tmxs1[i] *= dfact;
        }
        m.teff0 *= dfact;
        m.blcc *= dfacti;
        m.xcc *= std::sqrt(dfacti);
        found[im] = true;
        nfound++;
    }
    for (size_t k = 0; k < want.size(); k++)
        if (!found[k]) throw std::runtime_error("medium " + want[k] + " not in " + path);
}

// ------------------------------------------------------ photon (XCOM)
// egs_KN_sigma0, src/egsnrc.mortran:8729-8744 (total Klein-Nishina, barn)
static double kn_sigma0(double e) {
    const double con = 0.1274783851;
    double ko = e / kRM;
    if (ko < 0.01) return 8. * con / 3. * (1 - ko * (2 - ko * (5.2 - 13.3 * ko))) / kRM;
    double c1 = 1. / (ko * ko), c2 = 1. - 2 * (1 + ko) * c1, c3 = (1 + 2 * ko) * c1;
    double eps2 = 1, eps1 = 1. / (1 + 2 * ko);
    return (c1 * (1. / eps1 - 1. / eps2) + c2 * std::log(eps2 / eps1) +
            eps2 * (c3 + 0.5 * eps2) - eps1 * (c3 + 0.5 * eps1)) / e * con;
}

struct ElementFile {
    // per element Z=1..100: log-energy and log-cross-section points
    std::vector<std::vector<double>> le, lf;
};

static void read_element_file(const std::string &path, ElementFile &ef) {
    Records r(path);
    if (!r.ok()) throw std::runtime_error("cannot open " + path);
    ef.le.assign(101, {});
    ef.lf.assign(101, {});
    for (int iz = 1; iz <= 100; iz++) {
        auto nd = r.ints(1);
        auto v = r.reals(2 * (size_t)nd[0]);
        for (long k = 0; k < nd[0]; k++) {
            ef.le[iz].push_back(v[2 * k]);
            ef.lf[iz].push_back(v[2 * k + 1]);
        }
    }
}

// egsi_get_data, src/egsnrc.mortran:8776-8838 (flag 0: photo/Rayleigh,
// 1: pair with threshold 2m, 2: triplet with threshold 4m)
static std::vector<double> get_data(int flag, const ElementFile &ef, int n,
                                    const std::vector<double> &zs,
                                    const std::vector<double> &pzs, double ge1,
                                    double ge0) {
    std::vector<double> data(n, 0.0);
    for (size_t i = 0; i < zs.size(); i++) {
        int iz = (int)(zs[i] + 0.5);
        std::vector<double> et = ef.le[iz], ft = ef.lf[iz];
        double eth = 0;
        if (flag == 1 || flag == 2) {
            eth = flag == 1 ? 2 * kRM : 4 * kRM;
            for (size_t k = 0; k < et.size(); k++)
                ft[k] = ft[k] - 3 * std::log(1 - eth / std::exp(et[k]));
            et.insert(et.begin(), std::log(eth));
            ft.insert(ft.begin(), ft[0]);
        }
        int nd = (int)et.size();
        for (int k = 1; k <= n; k++) {
            double gle = (k - ge0) / ge1, e = std::exp(gle), sig;
            if (gle < et[0] || gle >= et[nd - 1]) {
                if (flag == 0)
                    throw std::runtime_error("photon energy outside the cross-section data range");
                sig = gle < et[0] ? 0 : std::exp(ft[nd - 1]);
            } else {
                int kk = 0;
                for (kk = 0; kk < nd - 1; kk++)
                    if (gle >= et[kk] && gle < et[kk + 1]) break;
                double p = (gle - et[kk]) / (et[kk + 1] - et[kk]);
                sig = std::exp(p * ft[kk + 1] + (1 - p) * ft[kk]);
            }
            if ((flag == 1 || flag == 2) && e > eth) sig = sig * std::pow(1 - eth / e, 3);
            data[k - 1] += pzs[i] * sig;
        }
    }
    return data;
}

// zero(), src/egs_utilities.mortran:2165: smallest positive double
static double tiny_positive() { return std::numeric_limits<double>::denorm_min(); }

// prepare_rayleigh_data, src/egsnrc.mortran:8598-8727
static void prepare_rayleigh(int ndat, std::vector<double> &x, std::vector<double> &f,
                             int ne, double emin, double emax,
                             std::vector<double> &pe, int ncbin,
                             std::vector<double> &fcum, std::vector<int32_t> &iarr,
                             std::vector<double> &barr, std::vector<double> &carr) {
    // arrays are used with Fortran 1-based indices below (element 0 unused)
    std::vector<double> X(ndat + 1), F(ndat + 1), FC(ndat + 1), B(ndat + 2, 0.0), Cc(ndat + 1), PE(ne + 1);
    std::vector<int32_t> IA(ncbin + 1);
    for (int i = 1; i <= ndat; i++) { X[i] = x[i - 1]; F[i] = f[i - 1]; if (F[i] == 0.0) F[i] = tiny_positive(); }
    double sum0 = 0;
    FC[1] = 0;
    for (int i = 1; i <= ndat - 1; i++) {
        double b = std::log(F[i + 1] / F[i]) / std::log(X[i + 1] / X[i]);
        B[i] = b;
        double x1 = X[i], x2 = X[i + 1];
        double px1 = std::pow(x1, 2 * b), px2 = std::pow(x2, 2 * b);
        sum0 += F[i] * F[i] * (x2 * x2 * px2 - x1 * x1 * px1) / ((1 + b) * px1);
        FC[i + 1] = sum0;
    }
    double dle = std::log(emax / emin) / (ne - 1);
    int i = 1, k = 0;
    for (int j = 1; j <= ne; j++) {
        double e = emin * std::exp(dle * (j - 1));
        double xmax = 20.607544 * 2 * e / kRM;
        for (k = i; k <= ndat - 1; k++)
            if (xmax >= X[k] && xmax < X[k + 1]) break;
        i = k;
        double b = B[i];
        double x1 = X[i], x2 = xmax;
        double px1 = std::pow(x1, 2 * b), px2 = std::pow(x2, 2 * b);
        PE[j] = FC[i] + F[i] * F[i] * (x2 * x2 * px2 - x1 * x1 * px1) / ((1 + b) * px1);
    }
    IA[ncbin] = i;
    double anorm = 1.0 / std::sqrt(PE[ne]);
    double anorm1 = 1.005 / PE[ne];
    double anorm2 = 1.0 / PE[ne];
    for (int j = 1; j <= ne; j++) {
        PE[j] *= anorm1;
        if (PE[j] > 1) PE[j] = 1;
    }
    for (int j = 1; j <= ndat; j++) {
        F[j] *= anorm;
        FC[j] *= anorm2;
        Cc[j] = (1 + B[j]) / ((X[j] * F[j]) * (X[j] * F[j]));
    }
    double dw = 1.0 / (ncbin - 1);
    double xold = X[1];
    int ibin = 1;
    double b = B[1];
    double px1 = std::pow(X[1], 2 * b);
    IA[1] = 1;
    for (int ii = 2; ii <= ncbin - 1; ii++) {
        double w = dw;
        for (;;) {
            double x1 = xold, x2 = X[ibin + 1];
            double t = x1 * x1 * std::pow(x1, 2 * b);
            double px2 = std::pow(x2, 2 * b);
            double aux = F[ibin] * F[ibin] * (x2 * x2 * px2 - t) / ((1 + b) * px1);
            if (aux > w) {
                xold = std::exp(std::log(t + w * (1 + b) * px1 / F[ibin] / F[ibin]) / (2 + 2 * b));
                IA[ii] = ibin;
                break;
            }
            w -= aux;
            xold = x2;
            ibin++;
            b = B[ibin];
            px1 = std::pow(xold, 2 * b);
        }
    }
    for (int j = 1; j <= ndat; j++) B[j] = 0.5 / (1 + B[j]);
    x.assign(X.begin() + 1, X.end());
    f.assign(F.begin() + 1, F.end());
    fcum.assign(FC.begin() + 1, FC.end());
    barr.assign(B.begin() + 1, B.begin() + 1 + ndat);
    carr.assign(Cc.begin() + 1, Cc.end());
    iarr.assign(IA.begin() + 1, IA.end());
    pe.assign(PE.begin() + 1, PE.end());
}

// FCOULC / XSIF, src/egsnrc.mortran:2410-2440
static double fcoulc(double Z) {
    const double fine = 137.03604;
    double asq = Z / fine;
    asq = asq * asq;
    return asq * (1.0 / (1.0 + asq) + 0.20206 + asq * (-0.0369 + asq * (0.0083 + asq * (-0.002))));
}
static double xsif(double Z) {
    static const double alrad[4] = {5.31, 4.79, 4.74, 4.71};
    static const double alradp[4] = {6.144, 5.621, 5.805, 5.924};
    if (Z <= 4) {
        int iZ = (int)Z;
        return alradp[iZ - 1] / (alrad[iZ - 1] - fcoulc(Z));
    }
    return std::log(1194.0 * std::pow(Z, -0.666667)) /
           (std::log(184.15 * std::pow(Z, -0.33333)) - fcoulc(Z));
}

// ------------------------------------------------------------- builder
int tables_build(const std::string &hh, const std::string &pegs_file,
                 const std::vector<std::string> &media, const BuildOptions &o,
                 Blob &out) {
    try {
        std::string data = hh + "/data/";
        std::vector<PegsMedium> pm;
        read_pegs(pegs_file, media, pm);
        int nmed = (int)pm.size();
        out = Blob();
        out.i["nmed"] = {nmed};
        out.f["rm"] = {kRM};
        out.f["estepe"] = {o.estepe};
        out.f["ximax"] = {o.ximax};
        out.i["ibcmp"] = {o.bound_compton};
        out.i["iraylr"] = {o.rayleigh};

        // ---- binding energies from the XCOM photo edges
        // (egs_init_user_photon :8041-8062, eadl_relax = true)
        // the photon data set "Photon cross sections" names
        // (egs_init_user_photon :8031-8035): <prefix>_photo.data, ...
        ElementFile photo, rayl, pair, trip;
        read_element_file(data + o.photon_xsec + "_photo.data", photo);
        read_element_file(data + o.photon_xsec + "_rayleigh.data", rayl);
        read_element_file(data + o.photon_xsec + "_pair.data", pair);
        read_element_file(data + o.photon_xsec + "_triplet.data", trip);
        {
            std::vector<int32_t> name;
            for (char c : o.photon_xsec) name.push_back(c);
            out.i["photon_xsec"] = name;
        }
        std::vector<double> bind(101 * kMXSHXSEC, 0.0);   // [iz][k], k = 0..29
        for (int iz = 1; iz <= 100; iz++) {
            const auto &et = photo.le[iz];
            int k = 0;
            for (int j = (int)et.size() - 1; j >= 1; j--) {
                if (et[j] - et[j - 1] < 1e-5) {
                    k++;
                    if (k > kMXSHXSEC) throw std::runtime_error("too many photo edges");
                    bind[iz * kMXSHXSEC + (k - 1)] = std::exp(et[j]);
                }
            }
        }

        // ---- compton_sigma.data (bound Compton total cross-section ratio)
        double bc_emin = 0, bc_emax = 0;
        long bc_ne = 0;
        std::vector<std::vector<double>> bc_el(101);
        {
            Records r(data + "compton_sigma.data");
            if (!r.ok()) throw std::runtime_error("cannot open compton_sigma.data");
            auto h = r.reals(3);
            bc_emin = h[0]; bc_emax = h[1]; bc_ne = (long)h[2];
            for (int iz = 1; iz <= 100; iz++) bc_el[iz] = r.reals(bc_ne);
        }
        // ---- pgs4form.dat atomic form factors (egs_init_rayleigh :8367-8380)
        std::vector<double> xval, aff;
        {
            Records r(hh + "/pegs4/pgs4form.dat");
            if (!r.ok()) throw std::runtime_error("cannot open pgs4form.dat");
            auto v = r.reals(kMXRAYFF + kMXRAYFF * 100);
            xval.assign(v.begin(), v.begin() + kMXRAYFF);
            aff.assign(v.begin() + kMXRAYFF, v.end());   // aff[(Z-1)*100 + i]
        }

        double esige_max = 0, psige_max = 0;
        for (int m = 0; m < nmed; m++) {
            PegsMedium &P = pm[m];
            auto K = [&](const char *fld) { return Blob::mk(m, fld); };
            out.i[K("nne")] = {P.nne};
            out.f[K("z")] = P.z;
            out.f[K("pz")] = P.pz;
            out.f[K("wa")] = P.wa;
            out.f[K("rho")] = {P.rho};
            out.f[K("rlc")] = {P.rlc};
            out.f[K("ae")] = {P.ae};
            out.f[K("ap")] = {P.ap};
            out.f[K("ue")] = {P.ue};
            out.f[K("up")] = {P.up};
            out.f[K("te")] = {P.ae - kRM};
            out.f[K("thmoll")] = {(P.ae - kRM) * 2. + kRM};
            {
                std::vector<int32_t> nm(P.name.begin(), P.name.end());
                out.i[K("name")] = nm;
            }

            // ---- photon cross sections on the XCOM grid (:8063-8284)
            int nge = kMXGE;
            double ge1 = (nge - 1) / std::log(P.up / P.ap);
            double ge0 = 1 - ge1 * std::log(P.ap);
            double sumZ = 0, sumA = 0;
            for (int i = 0; i < P.nne; i++) { sumZ += P.pz[i] * P.z[i]; sumA += P.pz[i] * P.wa[i]; }
            double con2 = P.rho / (sumA * 1.6605655);
            std::vector<int> ord(P.nne);
            for (int i = 0; i < P.nne; i++) ord[i] = i;
            std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return P.z[a] < P.z[b]; });
            std::vector<double> zs(P.nne), pzs(P.nne);
            for (int i = 0; i < P.nne; i++) { zs[i] = P.z[ord[i]]; pzs[i] = P.pz[ord[i]]; }
            auto sig_photo = get_data(0, photo, nge, zs, pzs, ge1, ge0);
            auto sig_rayl = get_data(0, rayl, nge, zs, pzs, ge1, ge0);
            auto sig_pair = get_data(1, pair, nge, zs, pzs, ge1, ge0);
            auto sig_trip = get_data(2, trip, nge, zs, pzs, ge1, ge0);
            std::vector<double> bc_data(bc_ne, 0.0);
            double bc_dle = std::log(bc_emax / bc_emin) / (bc_ne - 1);
            if (o.bound_compton > 1) {
                for (int i = 0; i < P.nne; i++) {
                    int iz = (int)(zs[i] + 0.5);
                    for (long j = 0; j < bc_ne; j++) bc_data[j] += pzs[i] * zs[i] * bc_el[iz][j];
                }
                for (long j = 0; j < bc_ne; j++) bc_data[j] = std::log(bc_data[j] / sumZ);
            }
            // Rayleigh sampling data (egs_init_rayleigh :8284-8450)
            if (o.rayleigh) {
                std::vector<double> xg(kMXRAYFF), ff(kMXRAYFF);
                for (int i = 0; i < kMXRAYFF; i++) {
                    double s = 0;
                    for (int j = 0; j < P.nne; j++) {
                        double a = aff[((int)P.z[j] - 1) * kMXRAYFF + i];
                        s += P.pz[j] * a * a;
                    }
                    ff[i] = std::sqrt(s);
                    xg[i] = xval[i];
                }
                if (xg[0] < 1e-6) xg[0] = 1e-4;
                double emin = std::exp((1 - ge0) / ge1), emax = std::exp((nge - ge0) / ge1);
                std::vector<double> pe, fcum, barr, carr;
                std::vector<int32_t> iarr;
                prepare_rayleigh(kMXRAYFF, xg, ff, nge, emin, emax, pe, kRAYCDF, fcum, iarr, barr, carr);
                std::vector<double> pmax0(nge), pmax1(nge);
                for (int i = 1; i <= nge - 1; i++) {
                    double gle = (i - ge0) / ge1;
                    pmax1[i - 1] = (pe[i] - pe[i - 1]) * ge1;
                    pmax0[i - 1] = pe[i - 1] - pmax1[i - 1] * gle;
                }
                pmax0[nge - 1] = pmax0[nge - 2];
                pmax1[nge - 1] = pmax1[nge - 2];
                out.f[K("ray_x")] = xg;
                out.f[K("ray_fcum")] = fcum;
                out.f[K("ray_b")] = barr;
                out.f[K("ray_c")] = carr;
                out.i[K("ray_i")] = iarr;
                out.f[K("ray_pmax0")] = pmax0;
                out.f[K("ray_pmax1")] = pmax1;
            }
            // the photon rows with the bound-Compton total cross section
            // (compton_sigma.data ratio, IBCMP > 1) or with Klein-Nishina
            // (IBCMP 0 and 1): the build option's set under the plain keys,
            // the Klein-Nishina set also under "<key>_kn" when the option
            // is bound, so that a run may switch bound Compton off
            auto photon_rows = [&](bool bound, const char *suffix) {
                std::vector<double> g0(nge), g1(nge), b10(nge), b11(nge), b20(nge), b21(nge),
                    c0(nge), c1(nge);
                double gmfp_old = 0, gbr1_old = 0, gbr2_old = 0, cohe_old = 0, gmfp = 0, gbr1 = 0,
                       gbr2 = 0, cohe = 0, gle = 0;
                for (int i = 1; i <= nge; i++) {
                    gle = (i - ge0) / ge1;
                    double e = std::exp(gle);
                    double sig_KN = sumZ * kn_sigma0(e);
                    if (bound) {
                        double bcf;
                        if (e <= bc_emin) bcf = std::exp(bc_data[0]);
                        else if (e < bc_emax) {
                            double aj = 1 + std::log(e / bc_emin) / bc_dle;
                            int j = (int)aj;
                            aj -= j;
                            bcf = std::exp(bc_data[j - 1] * (1 - aj) + bc_data[j] * aj);
                        } else bcf = 1;
                        sig_KN *= bcf;
                    }
                    double sig_p = sig_pair[i - 1] + sig_trip[i - 1];
                    double sigma = sig_KN + sig_p + sig_photo[i - 1];
                    gmfp = 1 / (sigma * con2);
                    gbr1 = sig_p / sigma;
                    gbr2 = gbr1 + sig_KN / sigma;
                    cohe = sigma / (sig_rayl[i - 1] + sigma);
                    if (i > 1) {
                        g1[i - 2] = (gmfp - gmfp_old) * ge1; g0[i - 2] = gmfp - g1[i - 2] * gle;
                        b11[i - 2] = (gbr1 - gbr1_old) * ge1; b10[i - 2] = gbr1 - b11[i - 2] * gle;
                        b21[i - 2] = (gbr2 - gbr2_old) * ge1; b20[i - 2] = gbr2 - b21[i - 2] * gle;
                        c1[i - 2] = (cohe - cohe_old) * ge1; c0[i - 2] = cohe - c1[i - 2] * gle;
                    }
                    gmfp_old = gmfp; gbr1_old = gbr1; gbr2_old = gbr2; cohe_old = cohe;
                }
                g1[nge - 1] = g1[nge - 2]; g0[nge - 1] = gmfp - g1[nge - 1] * gle;
                b11[nge - 1] = b11[nge - 2]; b10[nge - 1] = gbr1 - b11[nge - 1] * gle;
                b21[nge - 1] = b21[nge - 2]; b20[nge - 1] = gbr2 - b21[nge - 1] * gle;
                c1[nge - 1] = c1[nge - 2]; c0[nge - 1] = cohe - c1[nge - 1] * gle;
                auto KS = [&](const char *f) { return K(f) + std::string(suffix); };
                out.f[KS("gmfp0")] = g0; out.f[KS("gmfp1")] = g1;
                out.f[KS("gbr10")] = b10; out.f[KS("gbr11")] = b11;
                out.f[KS("gbr20")] = b20; out.f[KS("gbr21")] = b21;
                out.f[KS("cohe0")] = c0; out.f[KS("cohe1")] = c1;
            };
            photon_rows(o.bound_compton > 1, "");
            if (o.bound_compton > 1) photon_rows(false, "_kn");
            out.i[K("mge")] = {nge};
            out.f[K("ge0")] = {ge0};
            out.f[K("ge1")] = {ge1};

            // ---- fix_brems (:2322-2408) and brems/pair angle constants
            //      ($INITIALIZE-BREMS-ANGLE, egsnrc.macros:1981-1993)
            {
                double Zt = 0, Zb = 0, Zf = 0;
                for (int i = 0; i < P.nne; i++) {
                    double Zi = P.z[i], pi = P.pz[i];
                    double fc = fcoulc(Zi), xi = xsif(Zi);
                    double aux = pi * Zi * (Zi + xi);
                    Zt += aux;
                    Zb -= aux * std::log(Zi) / 3;
                    Zf += aux * fc;
                }
                double Zv = (Zb - Zf) / Zt, Zg = Zb / Zt;
                double fmax1 = 2 * (20.863 + 4 * Zg) - 2 * (20.029 + 4 * Zg) / 3;
                double fmax2 = 2 * (20.863 + 4 * Zv) - 2 * (20.029 + 4 * Zv) / 3;
                double dl[6][8];
                double Zs[2] = {Zg, Zv}, fm[2] = {fmax1, fmax2};
                for (int a = 0; a < 2; a++) {
                    double Z = Zs[a], f = fm[a];
                    int l = 2 * a;   // l = 1,2 (Zg) and 3,4 (Zv), 0-based 0,1 / 2,3
                    dl[0][l] = (20.863 + 4 * Z) / f; dl[1][l] = -3.242 / f; dl[2][l] = 0.625 / f;
                    dl[3][l] = (21.12 + 4 * Z) / f; dl[4][l] = -4.184 / f; dl[5][l] = 0.952;
                    dl[0][l + 1] = (20.029 + 4 * Z) / f; dl[1][l + 1] = -1.93 / f; dl[2][l + 1] = -0.086 / f;
                    dl[3][l + 1] = (21.12 + 4 * Z) / f; dl[4][l + 1] = -4.184 / f; dl[5][l + 1] = 0.952;
                    int p = 4 + 2 * a;   // l = 5,6 (Zg) and 7,8 (Zv)
                    dl[0][p] = 3 * (20.863 + 4 * Z) - (20.029 + 4 * Z);
                    dl[1][p] = 3 * (-3.242) - (-1.930);
                    dl[2][p] = 3 * (0.625) - (-0.086);
                    dl[3][p] = 2 * 21.12 + 8 * Z;
                    dl[4][p] = 2 * (-4.184);
                    dl[5][p] = 0.952;
                    dl[0][p + 1] = 3 * (20.863 + 4 * Z) + (20.029 + 4 * Z);
                    dl[1][p + 1] = 3 * (-3.242) + (-1.930);
                    dl[2][p + 1] = 3 * 0.625 + (-0.086);
                    dl[3][p + 1] = 4 * 21.12 + 16 * Z;
                    dl[4][p + 1] = 4 * (-4.184);
                    dl[5][p + 1] = 0.952;
                }
                std::vector<double> dlv(48);
                for (int k = 0; k < 6; k++)
                    for (int l = 0; l < 8; l++) dlv[k * 8 + l] = dl[k][l];
                out.f[K("dl")] = dlv;     // dl[k*8 + l]: DL(k+1) at l+1
                double bpar2 = dl[0][6] / (3 * dl[0][7] + dl[0][6]);
                double bpar1 = 12 * dl[0][7] / (3 * dl[0][7] + dl[0][6]);
                out.f[K("bpar")] = {bpar1, bpar2};
                out.f[K("delcm")] = {P.delcm};
                out.f[K("log_ap")] = {std::log(P.ap)};
                double zb = 0, pzn = 0;
                for (int i = 0; i < P.nne; i++) { zb += P.pz[i] * P.z[i] * (P.z[i] + 1.0); pzn += P.pz[i]; }
                zb = 8.116224E-05 * std::pow(zb / pzn, 1. / 3.);
                out.f[K("zbrang")] = {zb};
                out.f[K("lzbrang")] = {-std::log(zb)};
            }

            // ---- electrons: PEGS4 tables + mscati (:2750-2981)
            {
                int neke = P.meke;
                double blcc = 1.16699413758864573 * P.blcc;
                double xcc = P.xcc * P.xcc;
                auto ev = [&](const std::vector<double> &a0, const std::vector<double> &a1, int L, double x) {
                    return a1[L - 1] * x + a0[L - 1];
                };
                double sigee = 1e-15, sigep = 1e-15, sige_old = -1, sigp_old = -1;
                bool ise = true, isp = true;
                for (int i = 1; i <= neke; i++) {
                    double ei = std::exp(((double)i - P.eke0) / P.eke1);
                    double eil = std::log(ei);
                    double ededx = ev(P.ededx0, P.ededx1, i, eil);
                    double sig = ev(P.esig0, P.esig1, i, eil) / ededx;
                    if (sig > sigee) sigee = sig;
                    if (sig < sige_old) ise = false;
                    sige_old = sig;
                    ededx = ev(P.pdedx0, P.pdedx1, i, eil);
                    sig = ev(P.psig0, P.psig1, i, eil) / ededx;
                    if (sig > sigep) sigep = sig;
                    if (sig < sigp_old) isp = false;
                    sigp_old = sig;
                }
                esige_max = std::max(esige_max, sigee);
                psige_max = std::max(psige_max, sigep);
                std::vector<double> E(neke), r0(neke, 0.0), r1(neke, 0.0);
                double ei = std::exp((1 - P.eke0) / P.eke1);
                E[0] = ei;
                double expeke1 = std::exp(1. / P.eke1) - 1;
                for (int i = 1; i <= neke - 1; i++) {
                    double eip1 = std::exp(((double)(i + 1) - P.eke0) / P.eke1);
                    E[i] = eip1;
                    double eke = 0.5 * (eip1 + ei), elke = std::log(eke);
                    int L = (int)(P.eke1 * elke + P.eke0);
                    double ededx = ev(P.pdedx0, P.pdedx1, L, elke);
                    double aux = P.pdedx1[i - 1] / ededx;
                    r1[i] = r1[i - 1] + (eip1 - ei) / ededx * (1 + aux * (1 + 2 * aux) * std::pow((eip1 - ei) / eke, 2) / 24);
                    ededx = ev(P.ededx0, P.ededx1, L, elke);
                    aux = P.ededx1[i - 1] / ededx;
                    r0[i] = r0[i - 1] + (eip1 - ei) / ededx * (1 + aux * (1 + 2 * aux) * std::pow((eip1 - ei) / eke, 2) / 24);
                    ei = eip1;
                }
                std::vector<double> t0(neke), t1(neke);
                {
                    double eil = (1 - P.eke0) / P.eke1;
                    double e1 = std::exp(eil);
                    double p2 = e1 * (e1 + 2 * kRM), beta2 = p2 / (p2 + kRM * kRM);
                    double chi_a2 = xcc / (4 * p2 * blcc);
                    double dedx0 = ev(P.ededx0, P.ededx1, 1, eil);
                    double estepx = 2 * p2 * beta2 * dedx0 / e1 / xcc / (std::log(1 + 1. / chi_a2) * (1 + chi_a2) - 1);
                    estepx *= o.ximax;
                    if (estepx > o.estepe) estepx = o.estepe;
                    double si = estepx * e1 / dedx0;
                    for (int i = 1; i <= neke - 1; i++) {
                        double elke = ((double)(i + 1) - P.eke0) / P.eke1;
                        double eke = std::exp(elke);
                        int lelke = i + 1;
                        p2 = eke * (eke + 2 * kRM); beta2 = p2 / (p2 + kRM * kRM);
                        chi_a2 = xcc / (4 * p2 * blcc);
                        double ededx = ev(P.ededx0, P.ededx1, lelke, elke);
                        estepx = 2 * p2 * beta2 * ededx / eke / xcc / (std::log(1 + 1. / chi_a2) * (1 + chi_a2) - 1);
                        estepx *= o.ximax;
                        if (estepx > o.estepe) estepx = o.estepe;
                        double ekef = (1 - estepx) * eke, sip1;
                        int lelkef;
                        if (ekef <= E[0]) {
                            sip1 = (E[0] - ekef) / dedx0;
                            ekef = E[0];
                            lelkef = 1;
                        } else {
                            double elkef = std::log(ekef);
                            lelkef = (int)(P.eke1 * elkef + P.eke0);
                            int leip1l = lelkef + 1;
                            double eip1l = (leip1l - P.eke0) / P.eke1;
                            double eip1 = E[leip1l - 1];
                            double aux = (eip1 - ekef) / eip1;
                            double elktmp = 0.5 * (elkef + eip1l + 0.25 * aux * aux * (1 + aux * (1 + 0.875 * aux)));
                            double ektmp = 0.5 * (ekef + eip1);
                            int lelktmp = lelkef;
                            double ed = ev(P.ededx0, P.ededx1, lelktmp, elktmp);
                            aux = P.ededx1[lelktmp - 1] / ed;
                            sip1 = (eip1 - ekef) / ed * (1 + aux * (1 + 2 * aux) * std::pow((eip1 - ekef) / ektmp, 2) / 24);
                        }
                        sip1 = sip1 + r0[lelke - 1] - r0[lelkef];
                        t1[i - 1] = (sip1 - si) * P.eke1;
                        t0[i - 1] = sip1 - t1[i - 1] * elke;
                        si = sip1;
                    }
                    t0[neke - 1] = t0[neke - 2];
                    t1[neke - 1] = t1[neke - 2];
                }
                out.i[K("meke")] = {neke};
                out.f[K("eke0")] = {P.eke0};
                out.f[K("eke1")] = {P.eke1};
                out.f[K("esig0")] = P.esig0; out.f[K("esig1")] = P.esig1;
                out.f[K("psig0")] = P.psig0; out.f[K("psig1")] = P.psig1;
                out.f[K("ededx0")] = P.ededx0; out.f[K("ededx1")] = P.ededx1;
                out.f[K("pdedx0")] = P.pdedx0; out.f[K("pdedx1")] = P.pdedx1;
                out.f[K("ebr10")] = P.ebr10; out.f[K("ebr11")] = P.ebr11;
                out.f[K("pbr10")] = P.pbr10; out.f[K("pbr11")] = P.pbr11;
                out.f[K("pbr20")] = P.pbr20; out.f[K("pbr21")] = P.pbr21;
                out.f[K("tmxs0")] = t0; out.f[K("tmxs1")] = t1;
                out.f[K("range_ep0")] = r0; out.f[K("range_ep1")] = r1;
                out.f[K("e_array")] = E;
                out.f[K("expeke1")] = {expeke1};
                out.f[K("blcc")] = {blcc};
                out.f[K("xcc")] = {xcc};
                out.f[K("teff0")] = {P.teff0};
                out.f[K("esig_e")] = {sigee};
                out.f[K("psig_e")] = {sigep};
                out.i[K("sig_monotone")] = {ise ? 1 : 0, isp ? 1 : 0};
            }
        }
        out.f["esige_max"] = {esige_max};
        out.f["psige_max"] = {psige_max};

        // ---- bound Compton shells (init_compton :2442-2598, eadl_relax)
        {
            Records r(data + "incoh.data");
            if (!r.ok()) throw std::runtime_error("cannot open incoh.data");
            r.skip(18);
            std::vector<int32_t> iz_a(kMXTOTSH), shn_a(kMXTOTSH), ne_a(kMXTOTSH);
            std::vector<double> Jo(kMXTOTSH), be(kMXTOTSH), erfJo(kMXTOTSH);
            int izc = 0;
            for (int j = 0; j < kMXTOTSH; j++) {
                auto v = r.reals(5);
                iz_a[j] = (int32_t)v[0]; shn_a[j] = (int32_t)v[1]; ne_a[j] = (int32_t)v[2];
                Jo[j] = v[3] * 137.;
                be[j] = v[4] * 1e-6 / kRM;
                double aux_erf = 0.70710678119 * (1 + 0.3 * Jo[j]);
                erfJo[j] = 0.82436063535 * (std::erf(aux_erf) - 1);
                if (iz_a[j] != izc) { shn_a[j] = 1; izc = iz_a[j]; }
                else shn_a[j] = shn_a[j - 1] + 1;
                double &bnd = bind[iz_a[j] * kMXSHXSEC + (shn_a[j] - 1)];
                if (bnd > 0) be[j] = bnd / kRM;
                else bnd = be[j] * kRM;
            }
            out.i["cs_iz"] = iz_a; out.i["cs_shn"] = shn_a; out.i["cs_ne"] = ne_a;
            out.f["cs_Jo"] = Jo; out.f["cs_be"] = be; out.f["cs_erfJo"] = erfJo;
            for (int m = 0; m < nmed; m++) {
                auto z = out.F(Blob::mk(m, "z"));
                auto pz = out.F(Blob::mk(m, "pz"));
                std::vector<int32_t> shell;
                std::vector<double> eno;
                double pztot = 0;
                for (size_t i = 0; i < z.size(); i++) {
                    int iz = (int)z[i];
                    for (int j = 0; j < kMXTOTSH; j++)
                        if (iz == iz_a[j]) {
                            shell.push_back(j);
                            double aux = pz[i] * ne_a[j];
                            eno.push_back(aux);
                            pztot += aux;
                        }
                }
                int nsh = (int)shell.size();
                if (nsh == 0) throw std::runtime_error("medium has zero Compton shells");
                // 1-based working copies of the alias setup
                std::vector<double> E(nsh + 2);
                std::vector<int32_t> A(nsh + 2);
                for (int i = 1; i <= nsh; i++) { E[i] = -(eno[i - 1] / pztot); A[i] = i; }
                double atav = 1. / nsh;
                for (int i = 1; i <= nsh - 1; i++) {
                    int jh, jl;
                    for (jh = 1; jh <= nsh - 1; jh++)
                        if (E[jh] < 0 && std::fabs(E[jh]) > atav) break;
                    for (jl = 1; jl <= nsh - 1; jl++)
                        if (E[jl] < 0 && std::fabs(E[jl]) < atav) break;
                    double aux = atav - std::fabs(E[jl]);
                    E[jh] += aux;
                    E[jl] = -E[jl] / atav + jl;
                    A[jl] = jh;
                    if (i == nsh - 1) E[jh] = 1 + jh;
                }
                for (int i = 1; i <= nsh; i++) if (E[i] < 0) E[i] = 1 + i;
                out.i[Blob::mk(m, "n_shell")] = {nsh};
                out.i[Blob::mk(m, "shell")] = shell;       // 0-based shell index
                out.f[Blob::mk(m, "eno")] = std::vector<double>(E.begin() + 1, E.begin() + 1 + nsh);
                out.i[Blob::mk(m, "eno_atbin")] = std::vector<int32_t>(A.begin() + 1, A.begin() + 1 + nsh);
            }
        }
        out.f["binding_energies"] = bind;   // [iz*30 + k], MeV

        // ---- EDGSET (:6312-6436): shell interaction probabilities and
        //      parametrised photo-absorption cross sections
        {
            Records r(data + "photo_relax.data");
            if (!r.ok()) throw std::runtime_error("cannot open photo_relax.data");
            r.skip(100);
            r.skip(1);
            std::vector<double> ip(101 * kMXSHELL, 0.0);
            for (int i = 1; i <= 100; i++) {
                auto v = r.reals(6);
                for (int k = 0; k < 5; k++) ip[i * kMXSHELL + k] = v[1 + k];
                ip[i * kMXSHELL + 5] = 1.01;
            }
            out.f["interaction_prob"] = ip;
            Records c(data + "photo_cs.data");
            if (!c.ok()) throw std::runtime_error("cannot open photo_cs.data");
            std::vector<int32_t> en(101, 0);
            std::vector<double> ea(101 * kMXEDGE, 0), eb(101 * kMXEDGE, 0), ec(101 * kMXEDGE, 0),
                ed(101 * kMXEDGE, 0), ee(101 * kMXEDGE, 0);
            for (int i = 1; i <= 100; i++) {
                auto h = c.ints(2);
                en[i] = (int32_t)h[1];
                if (en[i] > kMXEDGE) throw std::runtime_error("too many edges in photo_cs.data");
                for (int j = 0; j < en[i]; j++) {
                    auto v = c.reals(5);
                    ea[i * kMXEDGE + j] = v[0]; eb[i * kMXEDGE + j] = v[1];
                    ec[i * kMXEDGE + j] = v[2]; ed[i * kMXEDGE + j] = v[3];
                    ee[i * kMXEDGE + j] = v[4];
                }
            }
            out.i["edge_number"] = en;
            out.f["edge_a"] = ea; out.f["edge_b"] = eb; out.f["edge_c"] = ec;
            out.f["edge_d"] = ed; out.f["edge_energies"] = ee;
        }

        // ---- egs_init_relax (:5794-5998): EADL relaxation data of every
        //      element of the media, from relax.data (4-byte records: the
        //      element count, one record per element pointing at its block;
        //      per shell the EADL type, the transition count, the binding
        //      energy, then (final state, probability) per transition, final
        //      states counted from 0 as C arrays do)
        {
            std::ifstream f(data + "relax.data", std::ios::binary);
            if (!f) throw std::runtime_error("cannot open relax.data");
            std::vector<char> raw((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
            const size_t nw = raw.size() / 4;
            auto word = [&](size_t rec) -> int32_t {   // rec: 1-based record
                if (rec < 1 || rec > nw) throw std::runtime_error("relax.data: record out of range");
                int32_t v;
                std::memcpy(&v, raw.data() + 4 * (rec - 1), 4);
                return v;
            };
            auto real = [&](size_t rec) -> double {
                int32_t v = word(rec);
                float x;
                std::memcpy(&x, &v, 4);
                return (double)x;
            };
            const double min_be = 1e-3;   // $RELAX-CUTOFF
            const std::vector<double> &bind = out.F("binding_energies");
            std::vector<int32_t> zfirst(101, -1), znsh(101, 0), shell_num, rfirst, rntran, rstate,
                ratbin;
            std::vector<double> shell_be, rprob;
            for (int m = 0; m < nmed; m++) {
                std::vector<double> zs = out.F(Blob::mk(m, "z"));
                std::sort(zs.begin(), zs.end());
                for (double zz : zs) {
                    const int iZ = (int)(zz + 0.5);
                    if (iZ < 1 || iZ > 100) throw std::runtime_error("relax.data: Z out of range");
                    if (zfirst[iZ] >= 0) continue;
                    size_t rec = (size_t)word((size_t)iZ + 1);
                    const int nshell = word(rec);
                    if (nshell < 1 || nshell > kMXSHXSEC)
                        throw std::runtime_error("relax.data: unexpected shell count");
                    zfirst[iZ] = (int32_t)shell_be.size();
                    znsh[iZ] = nshell;
                    for (int s = 1; s <= nshell; s++) {
                        rec++;                          // EADL shell type (not used)
                        const int ntran0 = word(++rec);
                        double be = real(++rec);
                        // XCOM edge energies where EDGSET has one
                        if (bind[(size_t)iZ * kMXSHXSEC + s - 1] > 0)
                            be = bind[(size_t)iZ * kMXSHXSEC + s - 1];
                        std::vector<int32_t> st(ntran0);
                        std::vector<double> w(ntran0);
                        for (int k = 0; k < ntran0; k++) {
                            int32_t t = word(++rec);
                            w[k] = real(++rec);
                            st[k] = t < 64 ? t + 1 : t + 65;   // 1-based shells
                        }
                        shell_be.push_back(be);
                        shell_num.push_back(s);
                        if (be < min_be || ntran0 < 1) {
                            rfirst.push_back(-1);
                            rntran.push_back(-1);
                            continue;
                        }
                        double sumw = 0;
                        for (double x : w) sumw += x;
                        if (sumw > 1) {
                            for (double &x : w) x /= sumw;
                        } else if (sumw < 1) {   // the rest: no new vacancy
                            st.push_back(-1);
                            w.push_back(1 - sumw);
                        }
                        const int n = (int)w.size();
                        // prepare_alias_histogram (egs_utilities.mortran:1940-1991)
                        std::vector<int32_t> bin(n, -1);
                        double avg = 0;
                        for (double x : w) avg += x;
                        avg /= n;
                        for (int i = 0; i < n - 1; i++) {
                            int jh = -1, jl = -1;
                            for (int j = 0; j < n; j++)
                                if (bin[j] < 0 && w[j] > avg) { jh = j; break; }
                            for (int j = 0; j < n; j++)
                                if (bin[j] < 0 && w[j] < avg) { jl = j; break; }
                            if (jh < 0 || jl < 0) break;   // the rest are at the average
                            w[jh] -= avg - w[jl];
                            w[jl] /= avg;
                            bin[jl] = jh + 1;
                        }
                        for (int j = 0; j < n; j++)
                            if (bin[j] < 0) { bin[j] = j + 1; w[j] = 1; }
                        rfirst.push_back((int32_t)rprob.size());
                        rntran.push_back(n);
                        for (int j = 0; j < n; j++) {
                            rprob.push_back(w[j]);
                            ratbin.push_back(bin[j]);
                            rstate.push_back(st[j]);
                        }
                    }
                }
            }
            out.i["relax_zfirst"] = zfirst;   // first shell of element Z (0-based), -1: none
            out.i["relax_nshell"] = znsh;
            out.f["shell_be"] = shell_be;     // MeV
            out.i["shell_num"] = shell_num;   // shell number within its element (1-based)
            out.i["relax_first"] = rfirst;    // first transition (0-based), -1: none
            out.i["relax_ntran"] = rntran;
            out.f["relax_prob"] = rprob;      // alias-histogram weights
            out.i["relax_atbin"] = ratbin;    // alias bins (1-based)
            out.i["relax_state"] = rstate;    // final state: shell (<= 64) or 64*s1 + s2
        }

        // ---- init_ms_SR (:3267-3306): q^(2+) surface, single precision
        {
            Records r(data + "msnew.data");
            if (!r.ok()) throw std::runtime_error("cannot open msnew.data");
            int n = kMS_L * kMS_Q * kMS_U;
            std::vector<double> ums(n), fms(n), wms(n, 0.0);
            std::vector<int32_t> ims(n, 0);
            for (int i = 0; i < kMS_L; i++)
                for (int j = 0; j < kMS_Q; j++) {
                    int base = (i * kMS_Q + j) * kMS_U;
                    auto u = r.reals(kMS_U);
                    auto f = r.reals(kMS_U);
                    auto w = r.reals(kMS_U - 1);
                    auto im = r.ints(kMS_U - 1);
                    float ff[kMS_U];
                    for (int k = 0; k < kMS_U; k++) { ums[base + k] = (float)u[k]; ff[k] = (float)f[k]; }
                    for (int k = 0; k < kMS_U - 1; k++) {
                        ff[k] = ff[k + 1] / ff[k] - 1.0f;
                        wms[base + k] = (float)w[k];
                        ims[base + k] = (int32_t)im[k] - 1;
                    }
                    ff[kMS_U - 1] = ff[kMS_U - 2];
                    for (int k = 0; k < kMS_U; k++) fms[base + k] = ff[k];
                }
            out.f["ms_u"] = ums; out.f["ms_f"] = fms; out.f["ms_w"] = wms;
            out.i["ms_i"] = ims;
            float llammin = std::log(1.0f), llammax = std::log(1e5f);
            float dllamb = (llammax - llammin) / 63.0f;
            float dllambi = 1.0f / dllamb;
            float dqms = 0.5f / 7.0f, dqmsi = 1.0f / dqms;
            out.f["ms_const"] = {llammin, llammax, dllamb, dllambi, dqms, dqmsi};
        }
        // init_nist_brems runs after fix_brems and mscati (HATCH,
        // egsnrc.mortran:2261-2289): it rewrites their cross sections
        out.i["ibr_nist"] = {0};
        if (o.ibr_nist >= 1) nist_brems_tables(data, o.ibr_nist, out);
        out.i["pair_nrc"] = {0};
        if (o.pair_nrc) nrc_pair_tables(data, out);
        // eii_init follows them (HATCH :2294) and corrects their tables
        out.i["eii"] = {0};
        if (!o.eii.empty()) {
            eii_tables(data, o.eii, out);
            std::vector<int32_t> name;
            for (char c : o.eii) name.push_back(c);
            out.i["eii_name"] = name;
        }
    } catch (const std::exception &e) {
        set_error("table build failed: %s", e.what());
        return EGSGPU_E_FORMAT;
    }
    return EGSGPU_OK;
}

}  // namespace egsgpu
