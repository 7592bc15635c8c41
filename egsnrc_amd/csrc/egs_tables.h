// egs_tables.h -- cross-section table blob: builder (from PEGS4 + EGSnrc
// data files) and reader.
//
// The blob is the state HATCH leaves behind for the media of a run
// (src/egsnrc.mortran:1828-2320): PEGS4 electron data scaled to cm,
// photon cross sections rebuilt on the 2000-point log grid from the XCOM
// element files (egs_init_user_photon :7973-8284), Rayleigh form-factor
// sampling tables (egs_init_rayleigh/prepare_rayleigh_data :8284-8720),
// bound Compton shells (init_compton :2442-2598), photo-absorption edges
// (EDGSET :6312-6436), brems/pair screening fits (fix_brems :2322-2420),
// the multiple-scattering step tables (mscati :2750-2981) and the screened
// Rutherford q-surface (init_ms_SR :3267-3306, data/msnew.data).
//
// Format: "EGSGPUT1", uint32 version, uint32 nrec, then nrec records of
// {char name[48]; uint32 type (1=f64, 2=i32); uint32 pad; uint64 count;
//  payload padded to 8 bytes}.  Keys are "<field>" for global data and
// "m<k>.<field>" for medium k.
#pragma once
#include <cstdint>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace egsgpu {

struct Blob {
    std::map<std::string, std::vector<double>> f;
    std::map<std::string, std::vector<int32_t>> i;

    bool has(const std::string &k) const { return f.count(k) || i.count(k); }
    const std::vector<double> &F(const std::string &k) const;
    const std::vector<int32_t> &I(const std::string &k) const;
    double f1(const std::string &k) const { return F(k).at(0); }
    int32_t i1(const std::string &k) const { return I(k).at(0); }
    static std::string mk(int m, const char *field) {
        return "m" + std::to_string(m) + "." + field;
    }
};

// Returns 0 or a negative EGSGPU_E_* code (message through set_error).
int blob_write(const Blob &b, const std::string &path);
int blob_read(const std::string &path, Blob &b);

struct BuildOptions {
    double estepe = 0.25;     // ESTEPE
    double ximax = 0.5;       // XIMAX
    int bound_compton = 3;    // IBCMP
    int rayleigh = 1;         // IRAYLR
    int ibr_nist = 0;         // 0 Bethe-Heitler, 1 NIST, 2 NRC brems cross sections
    int pair_nrc = 0;         // 1: add the NRC pair sampling tables
    std::string photon_xsec = "xcom";   // photon data set: <prefix>_photo.data, ...
    std::string eii;          // "" off, else the EII data set (eii_<name>.data)
};

int tables_build(const std::string &hen_house, const std::string &pegs_file,
                 const std::vector<std::string> &media, const BuildOptions &o,
                 Blob &out);

// ------------------------------------------------------- record reading
// Fortran list-directed / E-format records: every READ starts on a new
// line and drops what is left of its last line.
struct Records {
    std::ifstream in;
    std::string path;
    explicit Records(const std::string &p) : in(p), path(p) {}
    bool ok() const { return (bool)in; }
    bool line(std::string &s) { return (bool)std::getline(in, s); }
    void skip(int n = 1) { std::string s; for (int k = 0; k < n; k++) std::getline(in, s); }
    std::vector<double> reals(size_t n) {
        std::vector<double> v;
        std::string s;
        while (v.size() < n) {
            if (!std::getline(in, s)) throw std::runtime_error("unexpected end of " + path);
            for (auto &c : s) if (c == ',' || c == 'D' || c == 'd') c = (c == ',') ? ' ' : 'E';
            std::istringstream ls(s);
            std::string tok;
            while (v.size() < n && ls >> tok) v.push_back(std::stod(tok));
        }
        return v;
    }
    std::vector<long> ints(size_t n) {
        std::vector<double> r = reals(n);
        std::vector<long> v(r.size());
        for (size_t k = 0; k < r.size(); k++) v[k] = (long)std::llround(r[k]);
        return v;
    }
};

// NIST bremsstrahlung (init_nist_brems, egsnrc.mortran:6893-7226,
// ibr_nist = 1): replaces the total brems cross sections (esig, ebr1,
// psig, pbr1, pbr2, esig_e, psig_e) of every medium of a built blob by the
// integrals of data_dir/nist_brems.data and adds the photon-energy alias
// tables nb_xdata/nb_fdata/nb_wdata/nb_idata ([kNbE][kNbX+1] / [kNbE][kNbX])
// with nb_emin, nb_emax, nb_lemin, nb_dlei; sets "ibr_nist".
int nist_brems_tables(const std::string &data_dir, int ibr_nist, Blob &out);
// NRC pair cross sections (init_nrc_pair, egsnrc.mortran:7256-7352,
// pair_nrc = 1): from data_dir/pair_nrc1.data, per medium the positron
// energy-fraction densities at kNrcpE photon energies on the kNrcpX-point
// grid nrcp_xdata, their alias tables nrcp_fdata/nrcp_wdata/nrcp_idata and
// the grid constants nrcp_emin, nrcp_emax, nrcp_dlei (k in units of m c^2).
int nrc_pair_tables(const std::string &data_dir, Blob &out);
constexpr int kNrcpE = 84;      // photon energies ($NRC-PAIR-NEE)
constexpr int kNrcpX = 65;      // energy-fraction points ($NRC-PAIR-NXX)

// prepare_alias_table (egs_utilities.mortran:1846-1899) for the linear
// density through (xs[i], fs[i]), i = 0..n: ws[n] and 1-based ibin[n].
void alias_table(int n, const double *xs, const double *fs, double *ws, int32_t *ibin);

// Electron impact ionisation (eii_init, egsnrc.mortran:7434-7758): the
// shell cross sections of data_dir/eii_<name>.data for the K..LIII shells
// above the problem's lowest threshold ("eii_*" keys) and the corrected
// esig, ebr1, ededx, esig_e and sig_monotone of the media with such shells;
// sets "eii" (0 when no shell qualifies).
int eii_tables(const std::string &data_dir, const std::string &name, Blob &out);
constexpr int kEiiBins = 250;        // $N_EII_BINS
constexpr int kEiiMaxShells = 40;    // $MAX_EII_SHELLS

constexpr int kNbE = 100;       // energies of the brems sampling grid ($MXBRES)
constexpr int kNbX = 50;        // bins per alias table ($MXBRXS)
constexpr int kNbGauss = 64;    // Gauss-Legendre points ($MXGAUSS)

// mscati's default minimum step for PRESTA-I boundary crossing
// (egsnrc.mortran:2790-2817): BLCMIN = min over `media` of log(t / log t),
// t = blcc (E^2 - m^2) / xcc^2 at the lowest ECUT E, with the PEGS4 blcc and
// xcc; the skin depth is exp(BLCMIN) elastic mean free paths.
double bca_blcmin(const Blob &b, const std::vector<int> &media, double ecutmn);

// Dimensions fixed by the reference data layout.
constexpr int kMXGE = 2000;      // photon log-energy grid ($MXGE)
constexpr int kMXRAYFF = 100;    // Rayleigh form factor points
constexpr int kRAYCDF = 100;     // Rayleigh CDF bins ($RAYCDFSIZE)
constexpr int kMXTOTSH = 1538;   // Compton shells Z=1..100
constexpr int kMXSHXSEC = 30;    // binding energies per element
constexpr int kMXEDGE = 16;      // photo-absorption fit edges
constexpr int kMXSHELL = 6;
constexpr int kMS_L = 64, kMS_Q = 8, kMS_U = 32;   // msnew.data surface
constexpr double kRM = 0.5109989461;               // egsnrc.macros:1140

}  // namespace egsgpu
