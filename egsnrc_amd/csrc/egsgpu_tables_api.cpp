// egsgpu_tables_api.cpp -- C ABI for the cross-section table blob.
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/egsgpu.h"
#include "egs_tables.h"
#include "egsgpu_common.h"
#include "egsgpu_engine.h"

using namespace egsgpu;

extern "C" {

int egsgpu_tables_build(const char *hen_house, const char *pegs_file,
                        int32_t nmedia, const char *const *media,
                        const EgsTransport *tp, const char *out_path) {
    EGS_ARG(hen_house && pegs_file && media && out_path, "paths and media");
    EGS_ARG(nmedia > 0, "nmedia > 0");
    std::vector<std::string> names;
    for (int i = 0; i < nmedia; i++) {
        EGS_ARG(media[i], "medium name");
        names.push_back(media[i]);
    }
    BuildOptions o;
    if (tp) {
        if (tp->estepe > 0 && tp->estepe < 1) o.estepe = tp->estepe;
        if (tp->ximax > 0 && tp->ximax < 1) o.ximax = tp->ximax;
        o.bound_compton = tp->bound_compton;
        o.rayleigh = tp->rayleigh;
        EGS_ARG(tp->brems_xsec >= 0 && tp->brems_xsec <= 2,
                "brems_xsec: 0 Bethe-Heitler, 1 NIST or 2 NRC");
        o.ibr_nist = tp->brems_xsec;
        EGS_ARG(tp->pair_xsec == 0 || tp->pair_xsec == 1, "pair_xsec: 0 Bethe-Heitler or 1 NRC");
        o.pair_nrc = tp->pair_xsec;
        std::string eii(tp->eii_xsec, strnlen(tp->eii_xsec, sizeof tp->eii_xsec));
        for (auto &c : eii) c = (char)std::tolower((unsigned char)c);
        if (eii == "on") eii = "ik";   // setEIIData, egs_advanced_application.cpp:841-859
        if (eii == "off") eii.clear();
        EGS_ARG(eii.empty() || eii == "ik" || eii == "casnati" || eii == "kolbenstvedt" ||
                    eii == "gryzinski" || eii == "penelope",
                "eii_xsec: off, on, ik, casnati, kolbenstvedt, gryzinski or penelope");
        o.eii = eii;
        const std::string px(tp->photon_xsec, strnlen(tp->photon_xsec, sizeof tp->photon_xsec));
        if (!px.empty()) {
            EGS_ARG(px == "xcom" || px == "epdl" || px == "si",
                    "photon_xsec: xcom, epdl or si");
            o.photon_xsec = px;
        }
    }
    Blob b;
    int rc = tables_build(hen_house, pegs_file, names, o, b);
    if (rc) return rc;
    return blob_write(b, out_path);
}

int egsgpu_tables_load(const char *path, egsgpu_tables **out) {
    EGS_ARG(path && out, "path and out");
    auto *t = new egsgpu_tables();
    int rc = blob_read(path, t->blob);
    if (rc) { delete t; return rc; }
    if (!t->blob.has("nmed")) {
        delete t;
        set_error("%s has no medium count", path);
        return EGSGPU_E_FORMAT;
    }
    *out = t;
    return EGSGPU_OK;
}

int egsgpu_tables_free(egsgpu_tables *t) {
    delete t;
    return EGSGPU_OK;
}

int egsgpu_tables_nmedia(const egsgpu_tables *t, int32_t *n) {
    EGS_ARG(t && n, "tables and n");
    *n = t->blob.i1("nmed");
    return EGSGPU_OK;
}

int egsgpu_tables_medium_info(const egsgpu_tables *t, int32_t i, char *name,
                              size_t n, double *rho) {
    EGS_ARG(t, "tables");
    int nmed = t->blob.i1("nmed");
    EGS_ARG(i >= 0 && i < nmed, "medium index");
    if (name && n > 0) {
        std::string s;
        for (int c : t->blob.I(Blob::mk(i, "name"))) s.push_back((char)c);
        snprintf(name, n, "%s", s.c_str());
    }
    if (rho) *rho = t->blob.f1(Blob::mk(i, "rho"));
    return EGSGPU_OK;
}

int egsgpu_tables_bca_skindepth(const egsgpu_tables *t, double ecut, int32_t nmed,
                                const int32_t *media, double *blcmin, double *skindepth) {
    EGS_ARG(t && blcmin && skindepth, "tables, blcmin and skindepth");
    EGS_ARG(ecut > kRM, "ecut is a total energy (MeV) above the electron rest mass");
    const int all = t->blob.i1("nmed");
    std::vector<int> use;
    if (media) {
        EGS_ARG(nmed > 0, "nmed > 0");
        for (int i = 0; i < nmed; i++) {
            EGS_ARG(media[i] >= 0 && media[i] < all, "medium index out of range");
            use.push_back(media[i]);
        }
    } else {
        for (int m = 0; m < all; m++) use.push_back(m);
    }
    *blcmin = bca_blcmin(t->blob, use, ecut);
    *skindepth = std::exp(*blcmin);
    return EGSGPU_OK;
}

}  // extern "C"
