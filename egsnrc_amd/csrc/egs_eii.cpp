// egs_eii.cpp -- electron impact ionisation (eii_init,
// egsnrc.mortran:7434-7758): K, LI, LII and LIII shells whose binding
// energies exceed the lowest AE - m or AP of the problem get the cross
// sections of data/eii_<name>.data (250 log-spaced bins from the binding
// energy to emax); the Moller part of the electron cross section and the
// restricted stopping power of every medium with such shells are
// corrected for them.  Host, FP64, after the other tables.
#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "egs_tables.h"

namespace egsgpu {

int eii_tables(const std::string &data_dir, const std::string &name, Blob &out) {
    const int nmed = out.i1("nmed");
    const std::vector<double> &bind = out.F("binding_energies");   // [iz*30 + k]
    auto U = [&](int ish, int iz) { return bind[(size_t)iz * kMXSHXSEC + ish - 1]; };
    // the lowest threshold of the problem
    double emin = 1e30;
    for (int m = 0; m < nmed; m++) {
        emin = std::min(emin, out.f1(Blob::mk(m, "ae")) - kRM);
        emin = std::min(emin, out.f1(Blob::mk(m, "ap")));
    }
    // shells per element that take part, in the order of Z
    std::vector<int> nshells(101, 0);
    for (int m = 0; m < nmed; m++) {
        const std::vector<double> &z = out.F(Blob::mk(m, "z"));
        for (double zz : z) {
            const int iz = (int)(zz + 0.5);
            if (nshells[iz]) continue;
            int n = 0;
            for (int ish = 1; ish <= 4; ish++)
                if (U(ish, iz) > emin) n++;
            nshells[iz] = n;
        }
    }
    int nsh_tot = 0;
    std::vector<int> first_of(101, 0);   // shells of Z before this one
    for (int iz = 1; iz <= 100; iz++) {
        first_of[iz] = nsh_tot;
        nsh_tot += nshells[iz];
    }
    out.i["eii"] = {nsh_tot > 0 ? 1 : 0};
    if (nsh_tot == 0) return 0;   // no shell above the thresholds: EII off
    if (nsh_tot > kEiiMaxShells) throw std::runtime_error("eii: too many shells");

    // ---- the data file
    Records r(data_dir + "eii_" + name + ".data");
    if (!r.ok()) throw std::runtime_error("cannot open eii_" + name + ".data");
    const int nskip = (int)r.ints(1)[0];
    r.skip(nskip);
    auto hdr = r.reals(2);
    const double emax = hdr[0];
    if ((int)hdr[1] != kEiiBins) throw std::runtime_error("eii data: unexpected bin count");
    std::vector<double> ea(nsh_tot), eb(nsh_tot), xa((size_t)nsh_tot * kEiiBins),
        xb((size_t)nsh_tot * kEiiBins);
    std::vector<int32_t> shz(nsh_tot), shtype(nsh_tot);
    int ii = 0;
    for (int j = 1; j <= 100 && ii < nsh_tot; j++) {
        auto zn = r.ints(2);
        const int iz = (int)zn[0], nsh = (int)zn[1];
        if (iz < 1 || iz > 100) throw std::runtime_error("eii data: bad element");
        if (nsh < nshells[iz]) throw std::runtime_error("eii data: too few shells for an element");
        for (int ish = 1; ish <= nsh; ish++) {
            const double fmax = r.reals(1)[0];   // L shells scaled by eii_L_factor = 1
            const std::vector<double> aux = r.reals(kEiiBins);
            if (ish > nshells[iz]) continue;
            shz[ii] = iz;
            shtype[ii] = ish;
            ea[ii] = kEiiBins / std::log(emax / U(ish, iz));
            eb[ii] = 1 - ea[ii] * std::log(U(ish, iz));
            for (int k = 1; k <= kEiiBins; k++) {
                const double sigo = k > 1 ? fmax * aux[k - 2] : 0;
                const double loge = (k - eb[ii]) / ea[ii];
                const size_t i = (size_t)ii * kEiiBins + k - 1;
                xa[i] = (fmax * aux[k - 1] - sigo) * ea[ii];
                xb[i] = sigo - xa[i] * loge;
            }
            ii++;
        }
    }
    if (ii != nsh_tot) throw std::runtime_error("eii data: shells missing");
    out.f["eii_a"] = ea;
    out.f["eii_b"] = eb;
    out.f["eii_xa"] = xa;
    out.f["eii_xb"] = xb;
    out.i["eii_z"] = shz;
    out.i["eii_sh"] = shtype;

    // ---- per medium: shells of each element, and the corrected
    // Moller cross section and restricted stopping power
    const double cons = 0.153536;   // 2 pi r_e^2 m / u
    const int occn[5] = {0, 2, 2, 2, 4};
    for (int m = 0; m < nmed; m++) {
        auto K = [&](const char *f) { return Blob::mk(m, f); };
        const std::vector<double> &z = out.F(K("z")), &pz = out.F(K("pz")), &wa = out.F(K("wa"));
        const int nne = out.i1(K("nne"));
        std::vector<int32_t> no(nne), first(nne);
        int nshm = 0;
        double sum_a = 0;
        for (int e = 0; e < nne; e++) {
            const int iz = (int)(z[e] + 0.5);
            no[e] = nshells[iz];
            first[e] = nshells[iz] > 0 ? first_of[iz] + 1 : 0;   // 1-based
            nshm += nshells[iz];
            sum_a += pz[e] * wa[e];
        }
        const double rho = out.f1(K("rho")), ae = out.f1(K("ae")), ap = out.f1(K("ap"));
        const double con_med = rho / 1.6605655 / sum_a;
        out.i[K("eii_no")] = no;
        out.i[K("eii_first")] = first;
        out.i[K("eii_nsh")] = {nshm};
        out.f[K("eii_cons")] = {con_med};
        if (nshm == 0) continue;
        const double Ec = ae - kRM, Ecc = std::min(Ec, ap);
        const int meke = out.i1(K("meke"));
        const double eke0 = out.f1(K("eke0")), eke1 = out.f1(K("eke1"));
        std::vector<double> &ed0 = out.f[K("ededx0")], &ed1 = out.f[K("ededx1")];
        std::vector<double> &es0 = out.f[K("esig0")], &es1 = out.f[K("esig1")];
        std::vector<double> &eb0 = out.f[K("ebr10")], &eb1 = out.f[K("ebr11")];
        auto pw = [](const std::vector<double> &a0, const std::vector<double> &a1, int L,
                     double x) { return a1[L - 1] * x + a0[L - 1]; };
        std::vector<double> v_dedx(meke + 1), v_sig(meke + 1), v_wb(meke + 1);
        bool monotone = true;
        double sigma_max = 0;
        for (int j = 1; j <= meke; j++) {
            const double loge = (j - eke0) / eke1, e = std::exp(loge);
            const double tau = e / kRM, beta2 = tau * (tau + 2) / ((tau + 1) * (tau + 1));
            const double p2 = 2 * kRM * tau * (tau + 2);
            double dedx = pw(ed0, ed1, j, loge);
            const double sig = (e > ap || e > 2 * Ec) ? pw(es0, es1, j, loge) : 0;
            double sigm, wbrem;
            if (e > 2 * Ec) {
                wbrem = pw(eb0, eb1, j, loge);
                sigm = sig * (1 - wbrem);
            } else {
                sigm = 0;
                wbrem = 1;
            }
            double sum_occn = 0, sum_sigma = 0, sum_dedx = 0;
            for (int el = 0; el < nne; el++) {
                const int iz = (int)(z[el] + 0.5);
                double sum_sh = 0;
                for (int ish = 1; ish <= no[el]; ish++) {
                    const int jj = first[el] + ish - 1;          // global shell (1-based)
                    const int jjj = shtype[jj - 1];              // K = 1, LI = 2, ...
                    const double Ub = U(jjj, iz), Wmax = (e + Ub) / 2, uwm = Ub / Wmax;
                    if (!(Ub < e && Ub > Ecc)) continue;
                    sum_sh += occn[jjj];
                    const double ss_0 = 2 * (std::log(p2 / Ub) - uwm * uwm * uwm * std::log(p2 / Wmax) -
                                             (beta2 + 0.833333) * (1 - uwm * uwm * uwm)) / 3 / Ub;
                    const double sh_0 = ((1 - uwm) * (1 + uwm / (2 - uwm)) +
                                         Ub * (Wmax - Ub) / ((e + kRM) * (e + kRM)) -
                                         (2 * tau + 1) / ((tau + 1) * (tau + 1)) * uwm / 2 *
                                             std::log((2 - uwm) / uwm)) / Ub;
                    const double ss_1 = std::log(p2 / Ub) - uwm * uwm * std::log(p2 / Wmax) -
                                        (beta2 + 1) * (1 - uwm * uwm);
                    const double sh_1 = std::log(Wmax / Ub / (2 - uwm)) +
                                        2 * (Wmax - Ub) / (2 * Wmax - Ub) +
                                        (Wmax * Wmax - Ub * Ub) / ((e + kRM) * (e + kRM)) / 2 -
                                        (2 * tau + 1) / ((tau + 1) * (tau + 1)) *
                                            std::log((2 * Wmax - Ub) / Wmax);
                    const double av_e = (ss_1 + sh_1) / (ss_0 + sh_0);
                    // the reference indexes the energy grid by the shell type
                    // (eii_a(jjj)) and the data by the shell (jj); kept so
                    const int i = (int)(ea[jjj - 1] * loge + eb[jjj - 1]);
                    const long flat = (long)(jj - 1) * kEiiBins + i;     // 1-based
                    if (flat < 1 || flat > (long)xa.size())
                        throw std::runtime_error("eii: cross-section index outside the table");
                    double sig_j = xa[flat - 1] * loge + xb[flat - 1];
                    sig_j = sig_j * pz[el] * con_med;
                    sum_sigma += sig_j;
                    sum_dedx += sig_j * av_e;
                }
                sum_occn += sum_sh * pz[el];
            }
            sigm += sum_sigma;
            dedx -= sum_dedx;
            const double aux = Ec / e;
            if (e > 2 * Ec) {
                const double t1 = tau / (tau + 1), c2 = (2 * tau + 1) / ((tau + 1) * (tau + 1));
                const double sigo = cons * sum_occn * rho / (beta2 * Ec) *
                                    ((1 - 2 * aux) * (1 + aux / (1 - aux) + t1 * t1 * aux / 2) -
                                     c2 * aux * std::log((1 - aux) / aux)) / sum_a;
                const double de = cons * sum_occn * rho / beta2 *
                                  (std::log(0.25 / aux / (1 - aux)) + (1 - 2 * aux) / (1 - aux) +
                                   t1 * t1 * (1 - 4 * aux * aux) / 8 -
                                   c2 * std::log(2 * (1 - aux))) / sum_a;
                sigm -= sigo;
                dedx += de;
            }
            const double sigma = sigm + wbrem * sig;
            sigma_max = std::max(sigma_max, sigma / dedx);
            wbrem = sigma > 0 ? wbrem * sig / sigma : 1;
            v_dedx[j] = dedx;
            v_sig[j] = sigma;
            v_wb[j] = wbrem;
            if (j > 1 && sigma / dedx < v_sig[j - 1] / v_dedx[j - 1]) monotone = false;
        }
        auto refit = [&](std::vector<double> &a0, std::vector<double> &a1,
                         const std::vector<double> &v) {
            for (int j = 2; j <= meke; j++) {
                const double loge = (j - eke0) / eke1;
                a1[j - 2] = (v[j] - v[j - 1]) * eke1;
                a0[j - 2] = v[j] - a1[j - 2] * loge;
            }
            a1[meke - 1] = a1[meke - 2];
            a0[meke - 1] = a0[meke - 2];
        };
        refit(ed0, ed1, v_dedx);
        refit(es0, es1, v_sig);
        refit(eb0, eb1, v_wb);
        std::vector<int32_t> &mono = out.i[K("sig_monotone")];
        mono[0] = monotone ? 1 : 0;
        out.f[K("esig_e")] = {sigma_max};
    }
    return 0;
}

}  // namespace egsgpu
