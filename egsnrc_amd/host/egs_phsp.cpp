// egs_phsp.cpp -- EGSnrc phase-space file reader (see egs_phsp.h).
#include "egs_phsp.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
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

namespace {

std::string trim_iaea(const std::string &s) {
    size_t a = s.find_first_not_of(" \t\r");
    if (a == std::string::npos) return "";
    size_t b = s.find_last_not_of(" \t\r");
    return s.substr(a, b - a + 1);
}

// the value lines of each "$NAME:" block (blank lines skipped)
std::vector<std::pair<std::string, std::vector<std::string>>> iaea_blocks(const std::string &p) {
    std::ifstream f(p);
    if (!f) throw std::runtime_error("cannot open IAEA header " + p);
    std::vector<std::pair<std::string, std::vector<std::string>>> out;
    std::string l;
    while (std::getline(f, l)) {
        std::string t = trim_iaea(l);
        if (t.empty()) continue;
        if (t[0] == '$') {
            size_t c = t.find(':');
            out.push_back({t.substr(1, c == std::string::npos ? std::string::npos : c - 1), {}});
            std::string rest = c == std::string::npos ? "" : trim_iaea(t.substr(c + 1));
            if (!rest.empty()) out.back().second.push_back(rest);
        } else if (!out.empty()) {
            out.back().second.push_back(t);
        }
    }
    return out;
}

const std::vector<std::string> *iaea_block(
    const std::vector<std::pair<std::string, std::vector<std::string>>> &b, const char *name) {
    for (auto &x : b)
        if (x.first == name) return &x.second;
    return nullptr;
}

}  // namespace

EgsPhsp read_iaea_phsp(const std::string &base, double &zplane) {
    const std::string hp = base + ".IAEAheader", dp = base + ".IAEAphsp";
    auto blocks = iaea_blocks(hp);
    auto need = [&](const char *name) {
        const std::vector<std::string> *v = iaea_block(blocks, name);
        if (!v || v->empty()) throw std::runtime_error(hp + ": no " + name + " block");
        return v;
    };
    const std::vector<std::string> *rc = need("RECORD_CONTENTS");
    int contents[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = 0; i < 9 && i < (int)rc->size(); i++) contents[i] = atoi((*rc)[i].c_str());
    const int nxf = contents[7], nxl = contents[8];
    float constant[7] = {0, 0, 0, 0, 0, 0, 1};
    const std::vector<std::string> *cs = iaea_block(blocks, "RECORD_CONSTANT");
    for (int i = 0, k = 0; i < 7; i++) {
        if (contents[i] > 0) continue;
        if (!cs || k >= (int)cs->size())
            throw std::runtime_error(hp + ": RECORD_CONSTANT misses a constant");
        constant[i] = (float)atof((*cs)[k++].c_str());
    }
    const int order = atoi((*need("BYTE_ORDER"))[0].c_str());
    if (order != 1234 && order != 4321)
        throw std::runtime_error(hp + ": BYTE_ORDER must be 1234 or 4321");
    const bool swap = order == 4321;
    const int nstored = 1 + (contents[0] > 0) + (contents[1] > 0) + (contents[2] > 0) +
                        (contents[3] > 0) + (contents[4] > 0) + (contents[6] > 0) + nxf;
    const int reclen = 1 + 4 * nstored + 4 * nxl;
    const int hlen = atoi((*need("RECORD_LENGTH"))[0].c_str());
    if (hlen != reclen)
        throw std::runtime_error(hp + ": RECORD_LENGTH " + std::to_string(hlen) +
                                 " does not match RECORD_CONTENTS (" + std::to_string(reclen) + ")");
    std::ifstream f(dp, std::ios::binary | std::ios::ate);
    if (!f) throw std::runtime_error("cannot open IAEA phase space " + dp);
    const int64_t bytes = (int64_t)f.tellg();
    if (bytes % reclen) throw std::runtime_error(dp + ": size is not a whole number of records");
    f.seekg(0);
    std::vector<char> buf(bytes);
    f.read(buf.data(), bytes);
    EgsPhsp p;
    p.n = bytes / reclen;
    if (const std::vector<std::string> *ph = iaea_block(blocks, "PARTICLES"))
        if (atoll((*ph)[0].c_str()) != p.n)
            throw std::runtime_error(hp + ": PARTICLES does not match the size of " + dp);
    if (const std::vector<std::string> *oh = iaea_block(blocks, "ORIG_HISTORIES"))
        p.pinc = (float)atof((*oh)[0].c_str());
    p.e.resize(p.n); p.x.resize(p.n); p.y.resize(p.n); p.u.resize(p.n); p.v.resize(p.n);
    p.wt.resize(p.n); p.q.resize(p.n); p.latch.assign(p.n, 0); p.new_history.resize(p.n);
    p.emin = 1e30f; p.emax = 0;
    bool zset = false;
    for (int64_t k = 0; k < p.n; k++) {
        const char *r = buf.data() + k * reclen;
        int type = (signed char)r[0];
        float w_sign = type < 0 ? -1.0f : 1.0f;
        type = type < 0 ? -type : type;
        float fl[16];
        for (int i = 0; i < nstored && i < 16; i++) {
            uint32_t b;
            memcpy(&b, r + 1 + 4 * i, 4);
            if (swap) b = __builtin_bswap32(b);
            memcpy(&fl[i], &b, 4);
        }
        float val[7];
        for (int i = 0, j = 1; i < 7; i++) {
            if (i == 5) { val[i] = constant[5]; continue; }   // w is never stored
            val[i] = contents[i] > 0 ? fl[j++] : constant[i];
        }
        if (type == 1) p.q[k] = 0;
        else if (type == 2) p.q[k] = -1;
        else if (type == 3) p.q[k] = 1;
        else throw std::runtime_error(dp + ": particle type " + std::to_string(type) +
                                      " (only photons, electrons and positrons)");
        p.new_history[k] = fl[0] < 0;
        p.e[k] = fabsf(fl[0]);
        p.x[k] = val[0]; p.y[k] = val[1];
        p.u[k] = val[3]; p.v[k] = val[4];
        // the sign of w: the type's sign when w follows from u and v,
        // else the constant's
        float ws = contents[5] > 0 ? w_sign : (constant[5] < 0 ? -1.0f : 1.0f);
        p.wt[k] = ws * val[6];
        if (!zset) { zplane = val[2]; zset = true; }
        else if (fabs(val[2] - zplane) > 1e-5 * (1 + fabs(zplane)))
            throw std::runtime_error(dp + ": particles on more than one z plane");
        if (p.q[k] == 0) p.nphot++;
        p.emin = std::min(p.emin, p.e[k]);
        p.emax = std::max(p.emax, p.e[k]);
    }
    return p;
}
