// oracle_tables.h -- the oracle's own reader of the cross-section table blob
// (record layout documented in egsnrc_amd/csrc/egs_tables.h) and its
// per-medium view.  TEST INFRASTRUCTURE ONLY.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace oracle {

struct Blob {
    std::map<std::string, std::vector<double>> f;
    std::map<std::string, std::vector<int32_t>> i;
    const std::vector<double> &F(const std::string &k) const {
        auto it = f.find(k);
        if (it == f.end()) throw std::runtime_error("blob lacks " + k);
        return it->second;
    }
    const std::vector<int32_t> &I(const std::string &k) const {
        auto it = i.find(k);
        if (it == i.end()) throw std::runtime_error("blob lacks " + k);
        return it->second;
    }
    bool has(const std::string &k) const { return f.count(k) || i.count(k); }
    void read(const std::string &path) {
        std::ifstream in(path, std::ios::binary);
        if (!in) throw std::runtime_error("cannot open " + path);
        char magic[8];
        uint32_t ver, nrec;
        in.read(magic, 8);
        in.read((char *)&ver, 4);
        in.read((char *)&nrec, 4);
        if (memcmp(magic, "EGSGPUT1", 8) != 0) throw std::runtime_error("not a blob");
        for (uint32_t r = 0; r < nrec; r++) {
            char nm[48];
            uint32_t type, pad;
            uint64_t count;
            in.read(nm, 48);
            in.read((char *)&type, 4);
            in.read((char *)&pad, 4);
            in.read((char *)&count, 8);
            nm[47] = 0;
            size_t bytes;
            if (type == 1) {
                std::vector<double> v(count);
                in.read((char *)v.data(), 8 * count);
                f[nm] = v;
                bytes = 8 * count;
            } else {
                std::vector<int32_t> v(count);
                in.read((char *)v.data(), 4 * count);
                i[nm] = v;
                bytes = 4 * count;
            }
            if (bytes % 8) in.ignore(8 - bytes % 8);
        }
    }
};

// Per-medium data with the Fortran names of COMIN PHOTIN/ELECIN/THRESH/...
struct Medium {
    std::string name;
    double rho, ae, ap, ue, up, te, thmoll;
    int nne;
    std::vector<double> z, pz, wa;
    // photons
    double ge0, ge1;
    std::vector<double> gmfp0, gmfp1, gbr10, gbr11, gbr20, gbr21, cohe0, cohe1;
    // the same rows with the Klein-Nishina Compton cross section (empty
    // when the blob's main rows already are Klein-Nishina)
    std::vector<double> kn_gmfp0, kn_gmfp1, kn_gbr10, kn_gbr11, kn_gbr20, kn_gbr21, kn_cohe0,
        kn_cohe1;
    bool has_ray;
    std::vector<double> ray_x, ray_fcum, ray_b, ray_c, pmax0, pmax1;
    std::vector<int32_t> ray_i;
    int n_shell;
    std::vector<double> eno;
    std::vector<int32_t> eno_atbin, shell;
    double dl[6][8], bpar[2], delcm, zbrang, lzbrang, log_ap;
    // electrons
    int meke;
    double eke0, eke1, expeke1, blcc, xcc, esig_e, psig_e;
    int sig_mono[2];
    std::vector<double> esig0, esig1, psig0, psig1, ededx0, ededx1, pdedx0,
        pdedx1, ebr10, ebr11, pbr10, pbr11, pbr20, pbr21, tmxs0, tmxs1,
        range_ep0, range_ep1, e_array;
    // NIST brems sampling (init_nist_brems): 100 grid energies of 50-bin
    // alias tables, x/f [j*51 + k], w/bin [j*50 + k]
    double nb_emin = 0, nb_lemin = 0, nb_dlei = 0;
    std::vector<double> nb_x, nb_f, nb_w;
    std::vector<int32_t> nb_i;
    // NRC pair sampling (init_nrc_pair): 84 photon energies of 64-bin
    // alias tables over the positron energy fraction, f [ie*65 + ix],
    // w/bin [ie*64 + ix]
    std::vector<double> nrcp_f, nrcp_w;
    std::vector<int32_t> nrcp_i;
    // electron impact ionisation (eii_init): shells per element and the
    // first (1-based) in the global list, rho N_A / sum(PZ A)
    int eii_nsh = 0;
    std::vector<int32_t> eii_no, eii_first;
    double eii_cons = 0;
};

struct Tables {
    int nmed;
    int ibr_nist = 0;   // 1 / 2: the blob holds NIST / NRC brems cross sections
    int ibcmp = 3;      // the bound-Compton option the main photon rows were built for

    // a run with bound Compton off (IBCMP 0, 1) uses Klein-Nishina photon
    // rows (egs_init_user_photon, egsnrc.mortran:8028-8274)
    void select_compton_rows(int run_ibcmp) {
        if (run_ibcmp > 1 && ibcmp <= 1)
            throw std::runtime_error("bound Compton needs tables built with it");
        if (run_ibcmp > 1 || ibcmp <= 1) return;
        for (auto &M : med) {
            if (M.kn_gmfp0.empty()) throw std::runtime_error("the blob lacks Klein-Nishina rows");
            M.gmfp0 = M.kn_gmfp0; M.gmfp1 = M.kn_gmfp1;
            M.gbr10 = M.kn_gbr10; M.gbr11 = M.kn_gbr11;
            M.gbr20 = M.kn_gbr20; M.gbr21 = M.kn_gbr21;
            M.cohe0 = M.kn_cohe0; M.cohe1 = M.kn_cohe1;
        }
        ibcmp = run_ibcmp;
    }
    int pair_nrc = 0;   // 1: the blob holds the NRC pair tables
    int eii = 0;        // 1: the blob holds EII shells (and corrected tables)
    std::vector<double> eii_a, eii_b, eii_xa, eii_xb;   // per shell; 250 bins each
    std::vector<double> nrcp_x;   // positron energy-fraction grid (65)
    double nrcp_emin = 0, nrcp_emax = 0, nrcp_dlei = 0;
    std::vector<Medium> med;
    // Compton shells (init_compton)
    std::vector<int32_t> cs_iz, cs_shn;
    std::vector<double> cs_Jo, cs_be, cs_erfJo;
    // EDGE common
    std::vector<double> bind, iprob, ea, eb, ec, ed, ee;
    std::vector<int32_t> enumb;
    // EADL relaxation (egs_init_relax): shells of the media's elements and
    // their transitions as alias histograms
    std::vector<int32_t> relax_zfirst, relax_nshell, shell_num, relax_first, relax_ntran,
        relax_atbin, relax_state;
    std::vector<double> shell_be, relax_prob;
    // MS data
    std::vector<double> ms_u, ms_f, ms_w, ms_const;
    std::vector<int32_t> ms_i;
    double esige_max, psige_max;

    void load(const std::string &path) {
        Blob b;
        b.read(path);
        nmed = b.I("nmed")[0];
        med.resize(nmed);
        for (int m = 0; m < nmed; m++) {
            auto K = [&](const char *s) { return "m" + std::to_string(m) + "." + s; };
            Medium &M = med[m];
            for (int c : b.I(K("name"))) M.name.push_back((char)c);
            M.rho = b.F(K("rho"))[0]; M.ae = b.F(K("ae"))[0]; M.ap = b.F(K("ap"))[0];
            M.ue = b.F(K("ue"))[0]; M.up = b.F(K("up"))[0]; M.te = b.F(K("te"))[0];
            M.thmoll = b.F(K("thmoll"))[0];
            M.nne = b.I(K("nne"))[0];
            M.z = b.F(K("z")); M.pz = b.F(K("pz")); M.wa = b.F(K("wa"));
            M.ge0 = b.F(K("ge0"))[0]; M.ge1 = b.F(K("ge1"))[0];
            M.gmfp0 = b.F(K("gmfp0")); M.gmfp1 = b.F(K("gmfp1"));
            M.gbr10 = b.F(K("gbr10")); M.gbr11 = b.F(K("gbr11"));
            M.gbr20 = b.F(K("gbr20")); M.gbr21 = b.F(K("gbr21"));
            M.cohe0 = b.F(K("cohe0")); M.cohe1 = b.F(K("cohe1"));
            if (b.has(K("gmfp0_kn"))) {
                M.kn_gmfp0 = b.F(K("gmfp0_kn")); M.kn_gmfp1 = b.F(K("gmfp1_kn"));
                M.kn_gbr10 = b.F(K("gbr10_kn")); M.kn_gbr11 = b.F(K("gbr11_kn"));
                M.kn_gbr20 = b.F(K("gbr20_kn")); M.kn_gbr21 = b.F(K("gbr21_kn"));
                M.kn_cohe0 = b.F(K("cohe0_kn")); M.kn_cohe1 = b.F(K("cohe1_kn"));
            }
            M.has_ray = b.has(K("ray_x"));
            if (M.has_ray) {
                M.ray_x = b.F(K("ray_x")); M.ray_fcum = b.F(K("ray_fcum"));
                M.ray_b = b.F(K("ray_b")); M.ray_c = b.F(K("ray_c"));
                M.ray_i = b.I(K("ray_i"));
                M.pmax0 = b.F(K("ray_pmax0")); M.pmax1 = b.F(K("ray_pmax1"));
            }
            M.n_shell = b.I(K("n_shell"))[0];
            M.eno = b.F(K("eno")); M.eno_atbin = b.I(K("eno_atbin")); M.shell = b.I(K("shell"));
            auto dl = b.F(K("dl"));
            for (int k = 0; k < 6; k++) for (int l = 0; l < 8; l++) M.dl[k][l] = dl[k * 8 + l];
            M.bpar[0] = b.F(K("bpar"))[0]; M.bpar[1] = b.F(K("bpar"))[1];
            M.delcm = b.F(K("delcm"))[0];
            M.zbrang = b.F(K("zbrang"))[0]; M.lzbrang = b.F(K("lzbrang"))[0];
            M.log_ap = b.F(K("log_ap"))[0];
            M.meke = b.I(K("meke"))[0];
            M.eke0 = b.F(K("eke0"))[0]; M.eke1 = b.F(K("eke1"))[0];
            M.expeke1 = b.F(K("expeke1"))[0]; M.blcc = b.F(K("blcc"))[0];
            M.xcc = b.F(K("xcc"))[0]; M.esig_e = b.F(K("esig_e"))[0];
            M.psig_e = b.F(K("psig_e"))[0];
            M.sig_mono[0] = b.I(K("sig_monotone"))[0]; M.sig_mono[1] = b.I(K("sig_monotone"))[1];
            M.esig0 = b.F(K("esig0")); M.esig1 = b.F(K("esig1"));
            M.psig0 = b.F(K("psig0")); M.psig1 = b.F(K("psig1"));
            M.ededx0 = b.F(K("ededx0")); M.ededx1 = b.F(K("ededx1"));
            M.pdedx0 = b.F(K("pdedx0")); M.pdedx1 = b.F(K("pdedx1"));
            M.ebr10 = b.F(K("ebr10")); M.ebr11 = b.F(K("ebr11"));
            M.pbr10 = b.F(K("pbr10")); M.pbr11 = b.F(K("pbr11"));
            M.pbr20 = b.F(K("pbr20")); M.pbr21 = b.F(K("pbr21"));
            M.tmxs0 = b.F(K("tmxs0")); M.tmxs1 = b.F(K("tmxs1"));
            M.range_ep0 = b.F(K("range_ep0")); M.range_ep1 = b.F(K("range_ep1"));
            M.e_array = b.F(K("e_array"));
            if (b.has("ibr_nist") && b.I("ibr_nist")[0] >= 1) {
                M.nb_emin = b.F(K("nb_emin"))[0];
                M.nb_lemin = b.F(K("nb_lemin"))[0];
                M.nb_dlei = b.F(K("nb_dlei"))[0];
                M.nb_x = b.F(K("nb_xdata")); M.nb_f = b.F(K("nb_fdata"));
                M.nb_w = b.F(K("nb_wdata")); M.nb_i = b.I(K("nb_idata"));
            }
        }
        ibr_nist = b.has("ibr_nist") ? b.I("ibr_nist")[0] : 0;
        ibcmp = b.I("ibcmp")[0];
        pair_nrc = b.has("pair_nrc") ? b.I("pair_nrc")[0] : 0;
        eii = b.has("eii") ? b.I("eii")[0] : 0;
        if (eii) {
            eii_a = b.F("eii_a"); eii_b = b.F("eii_b");
            eii_xa = b.F("eii_xa"); eii_xb = b.F("eii_xb");
            for (int m = 0; m < nmed; m++) {
                auto K = [&](const char *f) { return "m" + std::to_string(m) + "." + f; };
                med[m].eii_nsh = b.I(K("eii_nsh"))[0];
                med[m].eii_no = b.I(K("eii_no"));
                med[m].eii_first = b.I(K("eii_first"));
                med[m].eii_cons = b.F(K("eii_cons"))[0];
            }
        }
        if (pair_nrc) {
            nrcp_x = b.F("nrcp_xdata");
            nrcp_emin = b.F("nrcp_const")[0];
            nrcp_emax = b.F("nrcp_const")[1];
            nrcp_dlei = b.F("nrcp_const")[2];
            for (int m = 0; m < nmed; m++) {
                auto K = [&](const char *f) { return "m" + std::to_string(m) + "." + f; };
                med[m].nrcp_f = b.F(K("nrcp_fdata"));
                med[m].nrcp_w = b.F(K("nrcp_wdata"));
                med[m].nrcp_i = b.I(K("nrcp_idata"));
            }
        }
        cs_iz = b.I("cs_iz"); cs_shn = b.I("cs_shn");
        cs_Jo = b.F("cs_Jo"); cs_be = b.F("cs_be"); cs_erfJo = b.F("cs_erfJo");
        bind = b.F("binding_energies"); iprob = b.F("interaction_prob");
        ea = b.F("edge_a"); eb = b.F("edge_b"); ec = b.F("edge_c");
        ed = b.F("edge_d"); ee = b.F("edge_energies"); enumb = b.I("edge_number");
        relax_zfirst = b.I("relax_zfirst"); relax_nshell = b.I("relax_nshell");
        shell_num = b.I("shell_num"); relax_first = b.I("relax_first");
        relax_ntran = b.I("relax_ntran"); relax_atbin = b.I("relax_atbin");
        relax_state = b.I("relax_state"); shell_be = b.F("shell_be");
        relax_prob = b.F("relax_prob");
        ms_u = b.F("ms_u"); ms_f = b.F("ms_f"); ms_w = b.F("ms_w");
        ms_i = b.I("ms_i"); ms_const = b.F("ms_const");
        esige_max = b.F("esige_max")[0]; psige_max = b.F("psige_max")[0];
    }
};

}  // namespace oracle
