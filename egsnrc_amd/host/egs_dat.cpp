// egs_dat.cpp -- .egsdat state files (see egs_dat.h).
#include "egs_dat.h"

#include "../csrc/ranmar_seed.h"

#include <fstream>
#include <iomanip>
#include <stdexcept>

namespace egshost {

namespace {

void put_i64(std::ostream &o, int64_t n) {
    o << " " << (int)(n / 1000000000) << " " << (int)(n % 1000000000);
}

int64_t get_i64(std::istream &in, const char *what) {
    int64_t a, b;
    if (!(in >> a >> b)) throw std::runtime_error(std::string("reading ") + what);
    return a * 1000000000 + b;
}

template <typename T>
T get(std::istream &in, const char *what) {
    T v;
    if (!(in >> v)) throw std::runtime_error(std::string("reading ") + what);
    return v;
}

}  // namespace

void ranmar_initial_state(int ixx, int jxx, EgsDat &d) {
    if (ixx <= 0 || ixx >= 31328) ixx = 1802;
    if (jxx <= 0 || jxx >= 30081) jxx = 9373;
    int u[97];
    ranmar_seed_words(ixx, jxx, u);
    d.u.assign(u, u + 97);
    d.c = kRanmarC0;
    d.ix = kRanmarIx0;
    d.jx = kRanmarJx0;
    d.iseed1 = ixx;
    d.iseed2 = jxx;
    d.high_res = 0;
    d.rng_count = 0;
    d.rarray.assign(128, 0.0);
    d.rng_ip = 128;
}

void write_egsdat(const std::string &path, const EgsDat &d) {
    std::ofstream o(path);
    if (!o) throw std::runtime_error("cannot write " + path);
    o << std::setprecision(17);
    put_i64(o, d.ndone);
    o << "  " << d.cpu << "\n";
    put_i64(o, d.current_case);
    o << "\n";
    put_i64(o, d.rng_count);
    o << " " << d.rarray.size() << "  " << d.rng_ip << "\n";
    for (double r : d.rarray) o << r << " ";
    o << "\n";
    o << d.ix << " " << d.jx << " " << d.c << " " << d.iseed1 << " " << d.iseed2 << " "
      << d.high_res << "\n";
    for (double v : d.u) o << v << " ";
    o << "\n";
    put_i64(o, d.src_count);
    put_i64(o, d.spec_count);
    o << " " << d.sum_E << " " << d.sum_E2 << "\n";
    put_i64(o, d.last_case);
    o << "\n";
    for (const EgsScoringArray &a : d.arrays) {
        o << a.nreg << "  " << a.current_ncase_short << "\n";
        put_i64(o, a.current_ncase);
        put_i64(o, a.current_ncase_65536);
        o << "\n";
        for (int j = 0; j < a.nreg; j++)
            o << a.last[j] << "  " << a.sum[j] << "  " << a.sum2[j] << "\n";
    }
    if (!o) throw std::runtime_error("writing " + path);
}

EgsDat read_egsdat(const std::string &path, int narrays) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("cannot open " + path);
    EgsDat d;
    try {
        d.ndone = get_i64(in, "the run control state");
        d.cpu = get<double>(in, "the cpu time");
        d.current_case = get_i64(in, "the current case");
        d.rng_count = get_i64(in, "the RNG count");
        int np = get<int>(in, "the RNG buffer size");
        d.rng_ip = get<int>(in, "the RNG buffer position");
        if (np < 1) throw std::runtime_error("RNG buffer size < 1");
        d.rarray.resize(np);
        for (double &r : d.rarray) r = get<double>(in, "the RNG buffer");
        d.ix = get<int>(in, "ranmar ix");
        d.jx = get<int>(in, "ranmar jx");
        d.c = get<int>(in, "ranmar c");
        d.iseed1 = get<int>(in, "ranmar iseed1");
        d.iseed2 = get<int>(in, "ranmar iseed2");
        d.high_res = get<int>(in, "ranmar high_res");
        d.u.resize(97);
        for (double &v : d.u) v = get<double>(in, "ranmar u");
        d.src_count = get_i64(in, "the source count");
        d.spec_count = get_i64(in, "the spectrum count");
        d.sum_E = get<double>(in, "the spectrum sum_E");
        d.sum_E2 = get<double>(in, "the spectrum sum_E2");
        d.last_case = get_i64(in, "the dose scoring state");
        for (int k = 0; k < narrays; k++) {
            EgsScoringArray a;
            a.nreg = get<int>(in, "a scoring array");
            a.current_ncase_short = get<int>(in, "a scoring array");
            if (a.nreg < 1) throw std::runtime_error("scoring array with nreg < 1");
            a.current_ncase = get_i64(in, "a scoring array");
            a.current_ncase_65536 = get_i64(in, "a scoring array");
            a.last.resize(a.nreg);
            a.sum.resize(a.nreg);
            a.sum2.resize(a.nreg);
            for (int j = 0; j < a.nreg; j++) {
                a.last[j] = get<int64_t>(in, "a region score");
                a.sum[j] = get<double>(in, "a region score");
                a.sum2[j] = get<double>(in, "a region score");
            }
            d.arrays.push_back(std::move(a));
        }
    } catch (const std::runtime_error &e) {
        throw std::runtime_error(path + ": " + e.what());
    }
    return d;
}

void add_egsdat(EgsDat &acc, const EgsDat &d) {
    if (acc.arrays.empty()) { acc = d; return; }
    if (acc.arrays.size() != d.arrays.size())
        throw std::runtime_error("state files with different scoring arrays");
    acc.ndone += d.ndone;
    acc.cpu += d.cpu;
    acc.current_case += d.current_case;
    int64_t count = acc.rng_count + d.rng_count;
    acc.rarray = d.rarray; acc.rng_ip = d.rng_ip; acc.u = d.u;
    acc.ix = d.ix; acc.jx = d.jx; acc.c = d.c;
    acc.iseed1 = d.iseed1; acc.iseed2 = d.iseed2; acc.high_res = d.high_res;
    acc.rng_count = count;
    acc.src_count += d.src_count;
    acc.spec_count += d.spec_count;
    acc.sum_E += d.sum_E;
    acc.sum_E2 += d.sum_E2;
    acc.last_case += d.last_case;
    for (size_t k = 0; k < acc.arrays.size(); k++) {
        EgsScoringArray &a = acc.arrays[k];
        const EgsScoringArray &b = d.arrays[k];
        if (a.nreg != b.nreg) throw std::runtime_error("scoring arrays of different sizes");
        for (int j = 0; j < a.nreg; j++) {
            a.sum[j] += b.sum[j];
            a.sum2[j] += b.sum2[j];
        }
        a.current_ncase += b.current_ncase;
    }
}

}  // namespace egshost
