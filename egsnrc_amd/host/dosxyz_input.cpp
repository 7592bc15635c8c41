// dosxyz_input.cpp -- DOSXYZnrc input records (see dosxyz_input.h).
#include "dosxyz_input.h"

#include <algorithm>
#include <cctype>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace egshost {

namespace {

std::string trim(const std::string &s) {
    size_t a = 0, b = s.size();
    while (a < b && isspace((unsigned char)s[a])) a++;
    while (b > a && isspace((unsigned char)s[b - 1])) b--;
    return s.substr(a, b - a);
}

// Fields of one record: comma-separated pieces, each split further at
// blanks; an empty piece is an empty field (read as zero).
std::vector<std::string> fields(const std::string &line) {
    std::vector<std::string> out;
    std::string piece;
    auto flush = [&](bool keep_empty) {
        std::string t = trim(piece);
        piece.clear();
        if (t.empty()) {
            if (keep_empty) out.push_back("");
            return;
        }
        std::istringstream is(t);
        std::string w;
        while (is >> w) out.push_back(w);
    };
    bool any_comma = false;
    for (char c : line) {
        if (c == ',') { any_comma = true; flush(true); }
        else piece += c;
    }
    flush(false);
    (void)any_comma;
    return out;
}

double num(const std::vector<std::string> &f, size_t i) {
    if (i >= f.size() || f[i].empty()) return 0.0;
    std::string s = f[i];
    for (char &c : s) if (c == 'd' || c == 'D') c = 'e';   // Fortran exponent
    char *end = nullptr;
    double v = strtod(s.c_str(), &end);
    if (end == s.c_str()) throw std::runtime_error("not a number: '" + f[i] + "'");
    return v;
}

int inum(const std::vector<std::string> &f, size_t i) { return (int)num(f, i); }

class Records {
  public:
    explicit Records(const std::string &path) {
        std::ifstream f(path);
        if (!f) throw std::runtime_error("cannot open " + path);
        std::string l;
        while (std::getline(f, l)) {
            if (!l.empty() && l.back() == '\r') l.pop_back();
            lines.push_back(l);
        }
    }
    const std::string &next(const char *what) {
        if (pos >= lines.size())
            throw std::runtime_error(std::string("input ends before the record: ") + what);
        last = what;
        return lines[pos++];
    }
    std::vector<std::string> next_fields(const char *what) { return fields(next(what)); }
    std::string rest() const {
        std::string s;
        for (size_t i = pos; i < lines.size(); i++) s += lines[i] + "\n";
        return s;
    }
    std::vector<std::string> lines;
    size_t pos = 0;
    const char *last = "";
};

// $GET-X-BOUNDARIES etc. (dosxyznrc.mortran:1675-1735): max > 0 planes one
// per record; max < 0 an initial plane, then |max| (width, count) groups
void read_bounds(Records &r, int max, std::vector<double> &b, const char *axis) {
    b.clear();
    if (max > 0) {
        for (int i = 0; i <= max; i++) b.push_back(num(r.next_fields(axis), 0));
    } else {
        b.push_back(num(r.next_fields(axis), 0));
        for (int g = 0; g < -max; g++) {
            auto f = r.next_fields(axis);
            double width = num(f, 0);
            int nn = inum(f, 1);
            if (nn <= 0) nn = 1;
            if (width <= 0) width = 1.0;
            for (int k = 0; k < nn; k++) b.push_back(b.back() + width);
        }
    }
    for (size_t i = 1; i < b.size(); i++)
        if (!(b[i] > b[i - 1]))
            throw std::runtime_error(std::string(axis) + " boundaries out of order");
}

}  // namespace

bool is_dosxyz_input(const std::string &path) {
    std::ifstream f(path);
    if (!f) return false;
    std::string l1, l2;
    if (!std::getline(f, l1) || !std::getline(f, l2)) return false;
    if (l1.find(":start") != std::string::npos || l1.find(":Start") != std::string::npos)
        return false;
    auto fl = fields(l2);
    if (fl.size() != 1) return false;
    char *end = nullptr;
    strtol(fl[0].c_str(), &end, 10);
    return end && *end == 0;
}

DosxyzInput read_dosxyz_input(const std::string &path) {
    Records r(path);
    DosxyzInput d;
    try {
        d.title = trim(r.next("title").substr(0, 80));
        d.nmed = inum(r.next_fields("NMED"), 0);
        if (d.nmed < 0) throw std::runtime_error("NMED < 0");
        if (d.nmed == 0) {
            d.phantom_file = trim(r.next("phantom file name"));
            auto f4 = r.next_fields("ECUTIN, PCUTIN, SMAX");
            d.ecutin = num(f4, 0);
            d.pcutin = num(f4, 1);
            auto f5 = r.next_fields("zeroairdose, doseprint, MAX20");
            d.zeroairdose = inum(f5, 0);
        } else {
            for (int m = 0; m < d.nmed; m++) {
                std::string nm = r.next("medium name");
                d.media.push_back(trim(nm.substr(0, std::min<size_t>(24, nm.size()))));
            }
            auto f4 = r.next_fields("ECUTIN, PCUTIN, ESTEPM, SMAX");
            d.ecutin = num(f4, 0);
            d.pcutin = num(f4, 1);
            auto f5 = r.next_fields("IMAX, JMAX, KMAX, IPHANT");
            int mx = inum(f5, 0), my = inum(f5, 1), mz = inum(f5, 2);
            if (mx == 0) mx = 1;
            if (my == 0) my = 1;
            if (mz == 0) mz = 1;
            read_bounds(r, mx, d.xb, "x boundaries");
            read_bounds(r, my, d.yb, "y boundaries");
            read_bounds(r, mz, d.zb, "z boundaries");
            const int nx = (int)d.xb.size() - 1, ny = (int)d.yb.size() - 1,
                      nz = (int)d.zb.size() - 1;
            d.med.assign((size_t)nx * ny * nz, 1);
            d.rho.assign((size_t)nx * ny * nz, 0.0);
            // media and densities by index ranges until a blank record
            // (dosxyznrc.mortran:1784-1811)
            for (;;) {
                auto f = r.next_fields("IL, IU, JL, JU, KL, KU, MEDIUM, DENSITY");
                int il = inum(f, 0), iu = inum(f, 1), jl = inum(f, 2), ju = inum(f, 3);
                int kl = inum(f, 4), ku = inum(f, 5), m = inum(f, 6);
                double rho = num(f, 7);
                if (il == 0 && iu == 0) break;
                if (il <= 0 || il > nx) il = 1;
                if (iu < il || iu > nx) iu = nx;
                if (jl <= 0 || jl > ny) jl = 1;
                if (ju < jl || ju > ny) ju = ny;
                if (kl <= 0 || kl > nz) kl = 1;
                if (ku < kl || ku > nz) ku = nz;
                if (m < 0 || m > d.nmed) m = 1;
                for (int k = kl; k <= ku; k++)
                    for (int j = jl; j <= ju; j++)
                        for (int i = il; i <= iu; i++) {
                            size_t irl = (size_t)(i - 1) + (size_t)(j - 1) * nx +
                                         (size_t)(k - 1) * nx * ny;
                            d.rho[irl] = rho;
                            d.med[irl] = m;
                        }
            }
            // local ECUT/PCUT groups: a disabled option, read to the blank record
            for (;;) {
                auto f = r.next_fields("IL, IU, JL, JU, KL, KU, ECUTL, PCUTL");
                if (inum(f, 0) == 0 && inum(f, 1) == 0) break;
            }
            // dose print groups (:1838-1860)
            for (;;) {
                auto f = r.next_fields("dose print groups");
                if (inum(f, 0) == 0 && inum(f, 1) == 0) break;
            }
        }
        // source records SC1-SC3 (srcxyznrc.mortran:1010-1060, 1760-1830)
        auto s1 = r.next_fields("iqin, isource, ...");
        d.iqin = inum(s1, 0);
        d.isource = inum(s1, 1);
        for (size_t i = 2; i < s1.size(); i++) d.src.push_back(num(s1, i));
        while (d.src.size() < 12) d.src.push_back(0.0);
        if (d.iqin < -1 || d.iqin > 1) d.iqin = 0;
        if (d.isource != 0 && d.isource != 1 && d.isource != 3)
            throw std::runtime_error("DOSXYZnrc source isource = " + std::to_string(d.isource) +
                                     " is not supported (isource 0, 1 and 3 are)");
        auto s2 = r.next_fields("enflag, mode, medsur, dsurround, dflag");
        d.enflag = inum(s2, 0);
        if (d.enflag == 0) {
            d.ein = num(r.next_fields("ein"), 0);
        } else if (d.enflag == 1) {
            d.spectrum_file = trim(r.next("spectrum file name"));
        } else {
            throw std::runtime_error("enflag = " + std::to_string(d.enflag) +
                                     " needs a phase-space or BEAM source, which is not supported");
        }
        // record 13 (dosxyznrc.mortran:1886)
        auto f = r.next_fields("NCASE, IWATCH, TIMMAX, INSEED1, INSEED2, ...");
        d.ncase = (int64_t)num(f, 0);
        d.iwatch = inum(f, 1);
        d.timmax = num(f, 2);
        d.inseed1 = inum(f, 3);
        d.inseed2 = inum(f, 4);
        d.beam_size = num(f, 5);
        d.ismooth = inum(f, 6);
        d.irestart = inum(f, 7);
        d.idat = inum(f, 8);
        d.ireject = inum(f, 9);
        d.esave_global = num(f, 10);
        d.nrcycl = inum(f, 11);
        d.iparallel = inum(f, 12);
        d.parnum = inum(f, 13);
        d.n_split = inum(f, 14);
        d.ihowfarless = inum(f, 15);
        d.i_phsp_out = inum(f, 16);
        d.i_bindos = inum(f, 17);
        d.transport_text = r.rest();
    } catch (const std::runtime_error &e) {
        throw std::runtime_error(path + ", record " + std::to_string(r.pos) + " (" + r.last +
                                 "): " + e.what());
    }
    return d;
}

void dosxyz_spectrum_cdfinv(const std::string &path, int ngrid, std::vector<double> &c1,
                            std::vector<double> &c2, double *emax) {
    Records r(path);
    r.next("spectrum title");
    auto h = r.next_fields("NENSRC, ENMIN, IMODE");
    int n = inum(h, 0), mode = inum(h, 2);
    double enmin = num(h, 1);
    if (n < 1) throw std::runtime_error(path + ": NENSRC < 1");
    if (mode != 0 && mode != 1)
        throw std::runtime_error(path + ": spectrum MODE must be 0 or 1");
    // the (energy, weight) pairs may share records or take one each
    std::vector<double> vals;
    while ((int)vals.size() < 2 * n) {
        auto f = r.next_fields("spectrum bins");
        for (size_t i = 0; i < f.size(); i++) vals.push_back(num(f, i));
    }
    std::vector<double> e(n), pdf(n), cdf(n);
    for (int i = 0; i < n; i++) { e[i] = vals[2 * i]; pdf[i] = vals[2 * i + 1]; }
    if (mode == 1) {   // counts per MeV -> per bin
        pdf[0] *= e[0] - enmin;
        for (int i = 1; i < n; i++) pdf[i] *= e[i] - e[i - 1];
    }
    cdf[0] = pdf[0];
    for (int i = 1; i < n; i++) cdf[i] = cdf[i - 1] + pdf[i];
    if (!(cdf[n - 1] > 0)) throw std::runtime_error(path + ": spectrum has no weight");
    double fnorm = 1.0 / cdf[n - 1];
    for (int i = 0; i < n; i++) cdf[i] *= fnorm;
    const double gridsz = 1.0 / (double)ngrid;
    c1.assign(ngrid, 0.0);
    c2.assign(ngrid, 0.0);
    for (int k = 1; k <= ngrid; k++) {
        double ak = (double)k * gridsz;
        int i = 0;
        while (i < n - 1 && !(ak <= cdf[i])) i++;
        c1[k - 1] = i > 0 ? e[i - 1] : enmin;
        c2[k - 1] = e[i] - c1[k - 1];
    }
    if (emax) *emax = e[n - 1];
}

}  // namespace egshost
