// egs_phantom.cpp -- readers for .egsphant phantoms and CT ramps.
//
// .egsphant (egs_nd_geometry.cpp:596-710, DOSXYZnrc's phantom format):
// number of media, one medium name per line, one ESTEPE value per medium
// (not used), nx ny nz, the x, y and z plane positions, one character per
// voxel for its medium ('1'..'9', then 'A'..'Z'; z slices of y rows, each
// slice followed by a blank line), then the mass density of every voxel.
// zlib reads compressed and plain files alike.
#include "egs_phantom.h"

#include <zlib.h>

#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace {

const char *kMediumChars = "123456789ABCDEFGHIJKLMNOPQRSTUVWXYZ";

struct GzLines {
    gzFile f;
    explicit GzLines(const std::string &path) : f(gzopen(path.c_str(), "rb")) {
        if (!f) throw std::runtime_error("cannot open " + path);
        gzbuffer(f, 1 << 20);
    }
    ~GzLines() { gzclose(f); }
    bool line(std::string &out) {
        out.clear();
        char buf[65536];
        for (;;) {
            if (!gzgets(f, buf, sizeof(buf))) return !out.empty();
            out += buf;
            if (!out.empty() && out.back() == '\n') {
                out.pop_back();
                if (!out.empty() && out.back() == '\r') out.pop_back();
                return true;
            }
        }
    }
};

}  // namespace

EgsPhantom read_egsphant(const std::string &path) {
    GzLines in(path);
    EgsPhantom ph;
    std::string ln;
    auto need = [&](bool ok, const char *what) {
        if (!ok) throw std::runtime_error(path + ": " + what);
    };
    need(in.line(ln), "missing number of media");
    int nmed = 0;
    need(sscanf(ln.c_str(), "%d", &nmed) == 1 && nmed > 0, "bad number of media");
    for (int i = 0; i < nmed; i++) {
        need(in.line(ln), "missing medium name");
        size_t a = ln.find_first_not_of(" \t"), b = ln.find_last_not_of(" \t");
        ph.media.push_back(a == std::string::npos ? "" : ln.substr(a, b - a + 1));
    }
    // numbers until the planes are complete
    std::vector<double> nums;
    int nx = 0, ny = 0, nz = 0;
    size_t want = (size_t)nmed + 3;
    while (nums.size() < want) {
        need(in.line(ln), "file ends in the header");
        std::istringstream is(ln);
        double x;
        while (is >> x) nums.push_back(x);
        if (want == (size_t)nmed + 3 && nums.size() >= want) {
            nx = (int)nums[nmed]; ny = (int)nums[nmed + 1]; nz = (int)nums[nmed + 2];
            need(nx > 0 && ny > 0 && nz > 0, "bad voxel numbers");
            want += (size_t)nx + ny + nz + 3;
        }
    }
    need(nums.size() == want, "unexpected numbers after the planes");
    size_t k = (size_t)nmed + 3;
    ph.xb.assign(nums.begin() + k, nums.begin() + k + nx + 1); k += nx + 1;
    ph.yb.assign(nums.begin() + k, nums.begin() + k + ny + 1); k += ny + 1;
    ph.zb.assign(nums.begin() + k, nums.begin() + k + nz + 1);
    size_t nreg = (size_t)nx * ny * nz;
    ph.medium.resize(nreg);
    for (int iz = 0; iz < nz; iz++) {
        for (int iy = 0; iy < ny;) {
            need(in.line(ln), "medium matrix ends early");
            size_t a = ln.find_first_not_of(" \t");
            if (a == std::string::npos) continue;
            need(ln.size() - a >= (size_t)nx, "short medium row");
            for (int ix = 0; ix < nx; ix++) {
                const char *p = strchr(kMediumChars, ln[a + ix]);
                need(p && ln[a + ix] != 0, "bad medium character");
                int m = (int)(p - kMediumChars);
                need(m < nmed, "medium index beyond the media list");
                ph.medium[((size_t)iz * ny + iy) * nx + ix] = m;
            }
            iy++;
        }
    }
    ph.density.reserve(nreg);
    while (ph.density.size() < nreg && in.line(ln)) {
        std::istringstream is(ln);
        double x;
        while (is >> x) ph.density.push_back(x);
    }
    need(ph.density.size() >= nreg, "density matrix ends early");
    ph.density.resize(nreg);
    return ph;
}

std::vector<CtRampEntry> read_ct_ramp(const std::string &path) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error("cannot open CT ramp " + path);
    std::vector<CtRampEntry> r;
    CtRampEntry e;
    while (f >> e.name >> e.rho_min >> e.rho_max >> e.rho_def) r.push_back(e);
    if (r.empty()) throw std::runtime_error(path + ": no media in the CT ramp");
    return r;
}
