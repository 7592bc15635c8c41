// egs_nist_brems.cpp -- the NIST or NRC bremsstrahlung cross sections
// (init_nist_brems, egsnrc.mortran:6893-7226, ibr_nist = 1 or 2): the total
// brems cross sections of a table blob are replaced by integrals of
// Seltzer and Berger's scaled cross sections (data/nist_brems.data) or of
// Tessier's (data/nrc_brems.data, the same layout on another grid), and
// per medium a 100-energy grid of 50-bin alias tables samples the photon
// energy.  FP64 on the host, at initialisation.
#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "egs_tables.h"

namespace egsgpu {

namespace {

// Cubic spline through (x[i], f[i]) with zero second derivative at both
// ends (set_spline, egs_utilities.mortran:1762-1802): on interval i the
// value is a[i] + q (b[i] + q (c[i] + q d[i])), q = s - x[i].  x may be
// increasing or decreasing.
struct Spline {
    std::vector<double> x, a, b, c, d;

    Spline(const std::vector<double> &xs, const std::vector<double> &f) : x(xs) {
        const int n = (int)x.size();
        a.assign(n, 0); b.assign(n, 0); c.assign(n, 0); d.assign(n, 0);
        // 0-based restatement of the tridiagonal sweep
        double s = 0;
        for (int m = 0; m < n - 1; m++) {
            d[m] = x[m + 1] - x[m];
            double r = (f[m + 1] - f[m]) / d[m];
            c[m] = r - s;
            s = r;
        }
        double r = 0;
        s = 0;
        c[0] = 0;
        c[n - 1] = 0;
        for (int m = 1; m < n - 1; m++) {
            c[m] = c[m] + r * c[m - 1];
            b[m] = 2 * (x[m - 1] - x[m + 1]) - r * s;
            s = d[m];
            r = s / b[m];
        }
        for (int m = n - 2; m >= 1; m--) c[m] = (d[m] * c[m + 1] - c[m]) / b[m];
        for (int m = 0; m < n - 1; m++) {
            s = d[m];
            r = c[m + 1] - c[m];
            d[m] = r / s;
            c[m] = 3 * c[m];
            b[m] = (f[m + 1] - f[m]) / s - (c[m] + r) * s;
            a[m] = f[m];
        }
    }

    // spline (egs_utilities.mortran:1804-1838): the end intervals extend
    // beyond the data
    double operator()(double s) const {
        const int n = (int)x.size();
        int m;
        if (x[0] > x[n - 1]) {   // decreasing abscissas
            if (s >= x[0]) m = 0;
            else if (s <= x[n - 1]) m = n - 2;
            else {
                int lo = 0, hi = n - 1;   // x[lo] > s >= x[hi]
                while (hi - lo > 1) {
                    int mid = (lo + hi) / 2;
                    if (s < x[mid]) lo = mid; else hi = mid;
                }
                m = lo;
            }
        } else {
            if (s >= x[n - 1]) m = n - 2;
            else if (s <= x[0]) m = 0;
            else {
                int lo = 0, hi = n - 1;   // x[lo] <= s < x[hi]
                while (hi - lo > 1) {
                    int mid = (lo + hi) / 2;
                    if (s < x[mid]) hi = mid; else lo = mid;
                }
                m = lo;
            }
        }
        const double q = s - x[m];
        return a[m] + q * (b[m] + q * (c[m] + q * d[m]));
    }
};

// n-point Gauss-Legendre abscissas and weights on [0, 1] (Newton's
// iteration on the Legendre polynomial from the Chebyshev-like start).
void gauss_legendre01(int n, std::vector<double> &x, std::vector<double> &w) {
    x.assign(n, 0);
    w.assign(n, 0);
    const double pi = 3.14159265358979323846;
    for (int i = 0; i < (n + 1) / 2; i++) {
        double z = std::cos(pi * (i + 0.75) / (n + 0.5)), pp = 0;
        for (int it = 0; it < 100; it++) {
            double p1 = 1, p2 = 0;
            for (int j = 1; j <= n; j++) {
                double p3 = p2;
                p2 = p1;
                p1 = ((2 * j - 1) * z * p2 - (j - 1) * p3) / j;
            }
            pp = n * (z * p1 - p2) / (z * z - 1);
            double dz = p1 / pp;
            z -= dz;
            if (std::fabs(dz) < 1e-15) break;
        }
        x[i] = 0.5 * (1 - z);
        x[n - 1 - i] = 0.5 * (1 + z);
        w[i] = w[n - 1 - i] = 1 / ((1 - z * z) * pp * pp);
    }
}

}  // namespace

// Alias table of the piecewise-linear density through (xs[i], fs[i]),
// i = 0..n (prepare_alias_table, egs_utilities.mortran:1846-1899): bin i
// (1..n) keeps its own mass with probability ws[i-1] and otherwise
// passes to bin ibin[i-1].
void alias_table(int n, const double *xs, const double *fs, double *ws, int32_t *ibin) {
    double mean = 0;
    for (int i = 1; i <= n; i++) {
        double aux = 0.5 * (fs[i] + fs[i - 1]) * (xs[i] - xs[i - 1]);
        if (aux < 1e-30) aux = 1e-30;
        ws[i - 1] = -aux;   // negative: not yet settled
        ibin[i - 1] = 1;
        mean += aux;
    }
    mean /= n;
    for (int it = 1; it <= n - 1; it++) {
        int jh = n, jl = n;
        for (int j = 1; j <= n; j++)
            if (ws[j - 1] < 0 && std::fabs(ws[j - 1]) > mean) { jh = j; break; }
        for (int j = 1; j <= n; j++)
            if (ws[j - 1] < 0 && std::fabs(ws[j - 1]) < mean) { jl = j; break; }
        double aux = mean - std::fabs(ws[jl - 1]);
        ws[jh - 1] += aux;
        ws[jl - 1] = -ws[jl - 1] / mean;
        ibin[jl - 1] = jh;
        if (it == n - 1) ws[jh - 1] = 1;
    }
}

int nist_brems_tables(const std::string &data_dir, int ibr_nist, Blob &out) {
    if (ibr_nist != 1 && ibr_nist != 2)
        throw std::runtime_error("ibr_nist: 1 NIST or 2 NRC brems cross sections");
    // the same layout in both files (egsnrc.mortran:6933-6964)
    const std::string file = ibr_nist == 1 ? "nist_brems.data" : "nrc_brems.data";
    Records r(data_dir + file);
    if (!r.ok()) throw std::runtime_error("cannot open " + file);
    r.skip(1);
    auto hdr = r.ints(2);
    const int nmix = (int)hdr[0], kmix = (int)hdr[1];
    if (nmix > 57 || kmix > 54) throw std::runtime_error(file + ": grid too large");
    const std::vector<double> energy = r.reals(nmix);
    const std::vector<double> xk = r.reals(kmix);
    r.skip(1);
    // cs[iz][k * nmix + n] for Z = iz (1..100): (beta^2/Z^2) k dsigma/dk (mb)
    std::vector<std::vector<double>> cs_el(101);
    for (int iz = 1; iz <= 100; iz++) cs_el[iz] = r.reals((size_t)nmix * kmix);
    std::vector<double> xi_k(kmix);
    for (int k = 0; k < kmix; k++) xi_k[k] = std::log(1 - xk[k] + 1e-6);
    std::vector<double> xg, wg;
    gauss_legendre01(kNbGauss, xg, wg);
    const double amu = 1660.5655;   // mb per atom -> cm^2/g for A in g/mol

    const int nmed = out.i1("nmed");
    double esige_max = out.f1("esige_max"), psige_max = out.f1("psige_max");
    for (int m = 0; m < nmed; m++) {
        auto K = [&](const char *f) { return Blob::mk(m, f); };
        const double ae = out.f1(K("ae")), ap = out.f1(K("ap")), ue = out.f1(K("ue"));
        const double rho = out.f1(K("rho"));
        const int nne = out.i1(K("nne"));
        const std::vector<double> &z = out.F(K("z")), &pz = out.F(K("pz")), &wa = out.F(K("wa"));
        // energies of the data that cover [max(AE - m, AP), UE - m]
        const double emin = std::max(ae - kRM, ap);
        int ifirst = 0;
        while (ifirst < nmix && energy[ifirst] < emin) ifirst++;
        int ilast = nmix - 1;
        while (ilast >= 0 && !(energy[ilast] < ue - kRM)) ilast--;
        ilast = ilast + 1;
        if (ifirst >= nmix) ifirst = nmix - 1;
        if (ilast > nmix - 1) ilast = nmix - 1;
        if (ilast < 0) ilast = 0;
        const int nener = ilast - ifirst + 1;
        if (nener < 2) throw std::runtime_error("nist_brems: fewer than two data energies");
        double sumA = 0;
        for (int j = 0; j < nne; j++) sumA += pz[j] * wa[j];
        sumA *= amu;
        // the medium's scaled cross section per gram at the data energies
        std::vector<std::vector<double>> cs(nener, std::vector<double>(kmix));
        std::vector<double> ele(nener), cse(nener);
        for (int ii = 0; ii < nener; ii++) {
            const int i = ifirst + ii;
            ele[ii] = std::log(energy[i]);
            std::vector<double> lcs(kmix);
            for (int k = 0; k < kmix; k++) {
                double s = 0;
                for (int j = 0; j < nne; j++) {
                    const int iz = (int)(z[j] + 0.1);
                    s += pz[j] * z[j] * z[j] / sumA * cs_el[iz][(size_t)k * nmix + i];
                }
                cs[ii][k] = s;
                lcs[k] = std::log(s);
            }
            Spline sp(xi_k, lcs);
            // integral over u = log(k/AP) / log(E/AP) in [0, 1]
            const double aux = std::log(energy[i] / ap);
            double sum = 0;
            for (int g = 0; g < kNbGauss; g++) {
                double xi = std::log(1 - ap / energy[i] * std::exp(xg[g] * aux) + 1e-6);
                sum += wg[g] * std::exp(sp(xi));
            }
            cse[ii] = sum;
        }
        Spline sp_e(ele, cse);
        // ---- total cross sections on the electron energy grid
        const int neke = out.i1(K("meke"));
        const double eke0 = out.f1(K("eke0")), eke1 = out.f1(K("eke1"));
        std::vector<double> &esig0 = out.f[K("esig0")], &esig1 = out.f[K("esig1")];
        std::vector<double> &psig0 = out.f[K("psig0")], &psig1 = out.f[K("psig1")];
        std::vector<double> &ebr10 = out.f[K("ebr10")], &ebr11 = out.f[K("ebr11")];
        std::vector<double> &pbr10 = out.f[K("pbr10")], &pbr11 = out.f[K("pbr11")];
        std::vector<double> &pbr20 = out.f[K("pbr20")], &pbr21 = out.f[K("pbr21")];
        const std::vector<double> &ededx0 = out.F(K("ededx0")), &ededx1 = out.F(K("ededx1"));
        const std::vector<double> &pdedx0 = out.F(K("pdedx0")), &pdedx1 = out.F(K("pdedx1"));
        auto pw = [](const std::vector<double> &a0, const std::vector<double> &a1, int L, double x) {
            return a1[L - 1] * x + a0[L - 1];   // L is 1-based
        };
        // new node values first (the PWLF coefficients are rewritten from
        // them, bin L spanning nodes L and L+1)
        std::vector<double> ne_sig(neke + 1), ne_br1(neke + 1), np_sig(neke + 1), np_br1(neke + 1),
            np_br2(neke + 1);
        double sigee = 1e-15, sigep = 1e-15;
        for (int i = 1; i <= neke; i++) {
            const double eil = ((double)i - eke0) / eke1, ei = std::exp(eil);
            const double beta2 = ei * (ei + 2 * kRM) / ((ei + kRM) * (ei + kRM));
            double sigb;
            if (ei <= ap) sigb = 1e-30;
            else sigb = sp_e(eil) * std::log(ei / ap) / beta2 * rho;
            double sigt = pw(esig0, esig1, i, eil), br1 = pw(ebr10, ebr11, i, eil);
            if (sigt < 0) sigt = 0;
            br1 = std::min(1.0, std::max(0.0, br1));
            ne_sig[i] = sigt * (1 - br1) + sigb;
            ne_br1[i] = sigb / ne_sig[i];
            sigt = pw(psig0, psig1, i, eil);
            br1 = pw(pbr10, pbr11, i, eil);
            double br2 = pw(pbr20, pbr21, i, eil);
            if (sigt < 0) sigt = 0;
            br1 = std::min(1.0, std::max(0.0, br1));
            br2 = std::min(1.0, std::max(0.0, br2));
            double sig_bhabha = std::max(0.0, sigt * (br2 - br1));
            np_sig[i] = sigt * (1 - br1) + sigb;
            np_br1[i] = sigb / np_sig[i];
            np_br2[i] = (sigb + sig_bhabha) / np_sig[i];
            sigee = std::max(sigee, ne_sig[i] / pw(ededx0, ededx1, i, eil));
            sigep = std::max(sigep, np_sig[i] / pw(pdedx0, pdedx1, i, eil));
        }
        auto refit = [&](std::vector<double> &a0, std::vector<double> &a1,
                         const std::vector<double> &v) {
            for (int i = 2; i <= neke; i++) {
                const double eil = ((double)i - eke0) / eke1;
                a1[i - 2] = (v[i] - v[i - 1]) * eke1;
                a0[i - 2] = v[i] - a1[i - 2] * eil;
            }
            a1[neke - 1] = a1[neke - 2];
            a0[neke - 1] = a0[neke - 2];
        };
        refit(esig0, esig1, ne_sig);
        refit(ebr10, ebr11, ne_br1);
        refit(psig0, psig1, np_sig);
        refit(pbr10, pbr11, np_br1);
        refit(pbr20, pbr21, np_br2);
        out.f[K("esig_e")] = {sigee};
        out.f[K("psig_e")] = {sigep};
        esige_max = std::max(esige_max, sigee);
        psige_max = std::max(psige_max, sigep);

        // ---- photon-energy sampling tables on a log grid of kNbE energies
        double nb_emin = energy[ifirst];
        if (nb_emin <= ap) nb_emin = energy[ifirst + 1];
        const double nb_emax = energy[ilast];
        const double lemin = std::log(nb_emin), lemax = std::log(nb_emax);
        const double dle = (lemax - lemin) / (kNbE - 1);
        std::vector<double> xd((size_t)kNbE * (kNbX + 1)), fd(xd.size()), wd((size_t)kNbE * kNbX);
        std::vector<int32_t> id((size_t)kNbE * kNbX);
        for (int i = 0; i < kNbE; i++) {
            const double eil = lemin + i * dle, ei = std::exp(eil);
            int ii = 0;   // data interval [ii, ii+1] holding ei (clamped)
            while (ii < nener && !(ei < std::exp(ele[ii]))) ii++;
            ii = ii - 1;
            if (ii < 0) ii = 0;
            if (ii > nener - 2) ii = nener - 2;
            const double ple = (eil - ele[ii]) / (ele[ii + 1] - ele[ii]), qle = 1 - ple;
            std::vector<double> csx(kmix);
            for (int k = 0; k < kmix; k++) csx[k] = std::log(qle * cs[ii][k] + ple * cs[ii + 1][k]);
            Spline sp(xi_k, csx);
            double *X = &xd[(size_t)i * (kNbX + 1)], *F = &fd[(size_t)i * (kNbX + 1)];
            const double x0 = ap / ei, aux = -std::log(x0);
            auto fval = [&](double u) {
                return std::exp(sp(std::log(1 - ap / ei * std::exp(u * aux) + 1e-6)));
            };
            X[0] = 0;
            F[0] = std::exp(sp(std::log(1 - x0 + 1e-6)));
            int k = 0;   // first data point strictly above AP/E (last if none)
            while (k < kmix && !(xk[k] > x0)) k++;
            if (k > kmix - 1) k = kmix - 1;
            if (kmix - 1 - k > kNbX)
                // the data points above AP/E alone exceed the 50 bins (the NRC
                // grid's 54 fractions reach down to 1e-6): the reference then
                // writes past the row and samples from its first 50 bins only
                throw std::runtime_error(file + ": more data points above AP/E than the 50 "
                                         "sampling bins hold (raise AP or lower UE)");
            int ndat = 0;
            for (int j = k + 1; j <= kmix - 2; j++) {
                ndat++;
                X[ndat] = std::log(xk[j] / x0) / aux;
                F[ndat] = std::exp(csx[j]);
            }
            ndat++;
            X[ndat] = 1;
            F[ndat] = std::exp(csx[kmix - 1]);
            // add the midpoint of largest relative linear-interpolation
            // error until the table holds kNbX bins
            while (ndat < kNbX) {
                double emax = 0, xm = 0, fm = 0;
                int km = 0;
                for (int q = 0; q < ndat; q++) {
                    double xmid = 0.5 * (X[q] + X[q + 1]), flin = 0.5 * (F[q] + F[q + 1]);
                    double fs = fval(xmid);
                    double err = std::fabs(1 - flin / fs);
                    if (err > emax) { emax = err; xm = xmid; fm = fs; km = q; }
                }
                ndat++;
                for (int q = ndat; q >= km + 2; q--) { X[q] = X[q - 1]; F[q] = F[q - 1]; }
                X[km + 1] = xm;
                F[km + 1] = fm;
            }
            alias_table(kNbX, X, F, &wd[(size_t)i * kNbX], &id[(size_t)i * kNbX]);
        }
        out.f[K("nb_emin")] = {nb_emin};
        out.f[K("nb_emax")] = {nb_emax};
        out.f[K("nb_lemin")] = {lemin};
        out.f[K("nb_dlei")] = {1 / dle};
        out.f[K("nb_xdata")] = xd;
        out.f[K("nb_fdata")] = fd;
        out.f[K("nb_wdata")] = wd;
        out.i[K("nb_idata")] = id;
    }
    out.f["esige_max"] = {esige_max};
    out.f["psige_max"] = {psige_max};
    out.i["ibr_nist"] = {ibr_nist};
    return 0;
}

}  // namespace egsgpu
