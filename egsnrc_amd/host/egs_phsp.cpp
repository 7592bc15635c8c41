// egs_phsp.cpp -- EGSnrc phase-space file reader (see egs_phsp.h).
#include "egs_phsp.h"

#include <cstring>
#include <fstream>
#include <stdexcept>

namespace {

template <typename T>
T swap32(T v) {
    uint32_t b;
    memcpy(&b, &v, 4);
    b = (b >> 24) | ((b >> 8) & 0xff00u) | ((b << 8) & 0xff0000u) | (b << 24);
    memcpy(&v, &b, 4);
    return v;
}

}  // namespace

EgsPhsp read_egsphsp(const std::string &path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open phase-space file " + path);
    char mode[5];
    f.read(mode, 5);
    EgsPhsp p;
    if (!f || (memcmp(mode, "MODE0", 5) && memcmp(mode, "MODE2", 5)))
        throw std::runtime_error(path + " is not a MODE0 or MODE2 phase-space file");
    p.mode2 = mode[4] == '2';
    const int recl = p.mode2 ? 32 : 28;
    int32_t n, nphot;
    float emax, emin, pinc;
    f.read((char *)&n, 4); f.read((char *)&nphot, 4);
    f.read((char *)&emax, 4); f.read((char *)&emin, 4); f.read((char *)&pinc, 4);
    if (!f) throw std::runtime_error(path + ": truncated header");
    bool swap = false;
    auto bad = [&]() {
        return n <= 0 || nphot < 0 || nphot > n || emin < 0 || emax < 0 || emax < emin || pinc < 0;
    };
    if (bad()) {   // the other byte order (openFile, :131-170)
        n = swap32(n); nphot = swap32(nphot);
        emax = swap32(emax); emin = swap32(emin); pinc = swap32(pinc);
        swap = true;
        if (bad()) throw std::runtime_error(path + ": meaningless phase-space header");
    }
    p.n = n; p.nphot = nphot; p.emax = emax; p.emin = emin; p.pinc = pinc;
    p.latch.resize(n); p.e.resize(n); p.x.resize(n); p.y.resize(n);
    p.u.resize(n); p.v.resize(n); p.wt.resize(n); p.q.resize(n); p.new_history.resize(n);
    std::vector<char> rec(recl);
    for (int64_t k = 0; k < n; k++) {
        // record k+1 (the header fills record 0)
        f.seekg((std::streamoff)(k + 1) * recl, std::ios::beg);
        f.read(rec.data(), recl);
        if (!f) throw std::runtime_error(path + ": truncated at particle " + std::to_string(k));
        int32_t latch;
        float v[6];
        memcpy(&latch, &rec[0], 4);
        memcpy(v, &rec[4], 24);
        if (swap) {
            latch = swap32(latch);
            for (float &a : v) a = swap32(a);
        }
        const float RM = 0.5109989461f;
        int8_t q = (latch & 1073741824) ? -1 : (latch & 536870912) ? 1 : 0;
        float E = v[0];
        p.new_history[k] = E < 0;
        if (E < 0) E = -E;
        if (q) E -= RM;
        p.latch[k] = latch; p.e[k] = E; p.x[k] = v[1]; p.y[k] = v[2];
        p.u[k] = v[3]; p.v[k] = v[4]; p.wt[k] = v[5]; p.q[k] = q;
    }
    return p;
}
