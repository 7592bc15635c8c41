// egs_nrc_pair.cpp -- the NRC pair cross sections (init_nrc_pair,
// egsnrc.mortran:7256-7352, pair_nrc = 1): positron energy-fraction
// densities per element from data/pair_nrc1.data, mixed per medium with
// the weights n_Z Z^2, and an alias table per photon energy.  Host, FP64.
#include <cmath>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "egs_tables.h"

namespace egsgpu {

int nrc_pair_tables(const std::string &data_dir, Blob &out) {
    // a direct-access file of 65 4-byte words per record: record 1 holds
    // emin, emax (k in units of m c^2), the numbers of energies and of
    // fractions, and a byte that is 0 for little-endian data; element Z's
    // energies are records (Z-1)*ne + 2 ...
    std::ifstream in(data_dir + "pair_nrc1.data", std::ios::binary);
    if (!in) throw std::runtime_error("cannot open pair_nrc1.data");
    const size_t rec = (size_t)kNrcpX * 4;
    std::vector<unsigned char> hdr(rec);
    if (!in.read((char *)hdr.data(), rec)) throw std::runtime_error("pair_nrc1.data: short header");
    const bool swap = hdr[16] != 0;   // this code runs on little-endian hosts
    auto word = [&](const unsigned char *p) {
        unsigned char b[4] = {p[0], p[1], p[2], p[3]};
        if (swap) { std::swap(b[0], b[3]); std::swap(b[1], b[2]); }
        uint32_t u;
        std::memcpy(&u, b, 4);
        return u;
    };
    auto f32 = [&](const unsigned char *p) {
        uint32_t u = word(p);
        float f;
        std::memcpy(&f, &u, 4);
        return (double)f;
    };
    const double emin = f32(&hdr[0]), emax = f32(&hdr[4]);
    const int ne = (int)word(&hdr[8]), nb = (int)word(&hdr[12]);
    if (nb != kNrcpX || ne != kNrcpE) throw std::runtime_error("pair_nrc1.data: unexpected grid");
    // energy-fraction grid: quadratic near 0 and near 1
    std::vector<double> xd(kNrcpX);
    const int nbb = nb / 2;
    const double ddx = std::sqrt(0.5) / nbb;
    for (int ix = 0; ix <= nbb; ix++) { double xx = ddx * ix; xd[ix] = xx * xx; }
    for (int ix = nbb - 1; ix >= 0; ix--) { double xx = ddx * ix; xd[nb - 1 - ix] = 1 - xx * xx; }
    auto element = [&](int iz, int ie, std::vector<double> &row) {
        const size_t irec = (size_t)(iz - 1) * ne + 1 + ie;   // 0-based record
        std::vector<unsigned char> buf(rec);
        in.seekg((std::streamoff)(irec * rec));
        if (!in.read((char *)buf.data(), rec)) throw std::runtime_error("pair_nrc1.data: short read");
        for (int ix = 0; ix < kNrcpX; ix++) row[ix] = f32(&buf[4 * ix]);
    };
    const int nmed = out.i1("nmed");
    std::vector<double> row(kNrcpX);
    for (int m = 0; m < nmed; m++) {
        auto K = [&](const char *f) { return Blob::mk(m, f); };
        const int nne = out.i1(K("nne"));
        const std::vector<double> &z = out.F(K("z")), &pz = out.F(K("pz"));
        std::vector<double> fd((size_t)kNrcpE * kNrcpX, 0.0), wd((size_t)kNrcpE * (kNrcpX - 1));
        std::vector<int32_t> id((size_t)kNrcpE * (kNrcpX - 1));
        for (int j = 0; j < nne; j++) {
            const int iz = (int)(z[j] + 0.5);
            const double w = pz[j] * z[j] * z[j];
            for (int ie = 0; ie < kNrcpE; ie++) {
                element(iz, ie, row);
                for (int ix = 0; ix < kNrcpX; ix++) fd[(size_t)ie * kNrcpX + ix] += w * row[ix];
            }
        }
        for (int ie = 0; ie < kNrcpE; ie++)
            alias_table(kNrcpX - 1, xd.data(), &fd[(size_t)ie * kNrcpX],
                        &wd[(size_t)ie * (kNrcpX - 1)], &id[(size_t)ie * (kNrcpX - 1)]);
        out.f[K("nrcp_fdata")] = fd;
        out.f[K("nrcp_wdata")] = wd;
        out.i[K("nrcp_idata")] = id;
    }
    out.f["nrcp_xdata"] = xd;
    out.f["nrcp_const"] = {emin, emax, (ne - 1) / std::log((emax - 2) / (emin - 2))};
    out.i["pair_nrc"] = {1};
    return 0;
}

}  // namespace egsgpu
