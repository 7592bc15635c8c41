// oracle_rng.h -- CPU restatement of the EGSnrc random number generators.
//
// TEST INFRASTRUCTURE ONLY: the oracle is the checker for the HIP engine.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
// load it.  It is never part of the product path.
//
// ranmar: EGS_Ranmar::setState / fillArray   egs++/egs_rndm.cpp:303-398
//         (same stream as ranmar.mortran init_ranmar/ranmar_get)
// ranlux: ranlux / init_ranlux                src/ranlux.mortran (whole file)
// Pinned by the ranmar_test / ranlux_test known sums
// (user_codes/ranmar_test/ranmar_test.mortran, ranlux_test/ranlux_test.mortran).
#pragma once
#include <cstdint>

#include "../egsnrc_amd/csrc/ranmar_seed.h"

namespace oracle {

// Seed pair of stream s: (ixx, jxx+s) as createRNG(input, s) builds it
// (egs_rndm.cpp:438-446), the excess above 30080 carried into ixx
// (ranmar_seed.h, shared with the engine).
inline void stream_seeds(int ixx, int jxx, int64_t s, int &is, int &js) {
    ranmar_stream_seeds(ixx, jxx, s, is, js);
}

struct Ranmar {
    int u[97];
    int c, ix, jx;
    // egs_rndm.cpp:303-344
    void set_state(int ixx, int jxx) {
        ranmar_seed_words(ixx, jxx, u);
        c = kRanmarC0; ix = kRanmarIx0; jx = kRanmarJx0;
    }
    // one step of egs_rndm.cpp:338-370; returns the 24-bit integer
    int next_int() {
        int r = u[ix] - u[jx--];
        if (r < 0) r += 16777216;
        u[ix--] = r;
        if (ix < 0) ix = 96;
        if (jx < 0) jx = 96;
        c -= 7654321;
        if (c < 0) c += 16777213;
        r -= c;
        if (r < 0) r += 16777216;
        return r;
    }
    double uniform() { return next_int() * (1.0 / 16777216.0); }
};

struct Ranlux {
    int seeds[24];
    int carry, i24, j24, nskip;
    int next[24];
    // init_ranlux (ranlux.mortran) with 0-based indices: Fortran index f
    // is stored at f-1.
    void init(int luxury, int seed) {
        static const int nskipll[5] = {0, 24, 73, 199, 365};
        int jseed = seed;
        if (jseed <= 0) jseed = 314159265;
        if (luxury < 0 || luxury > 4) luxury = 1;
        nskip = nskipll[luxury];
        for (int j = 1; j <= 24; j++) {
            int k = jseed / 53668;
            jseed = 40014 * (jseed - k * 53668) - k * 12211;
            if (jseed < 0) jseed += 2147483563;
            seeds[j - 1] = jseed % 16777216;
            next[j - 1] = j - 1;             // next(j) = j-1 (Fortran)
        }
        next[0] = 24;                        // next(1) = 24
        i24 = 24; j24 = 10; carry = 0;
        if (seeds[23] == 0) carry = 1;
    }
    // ranlux(rng_array): 24 outputs then nskip discarded steps.  Each
    // output is formed in single precision exactly as the Fortran
    // expression (real*4 twom24) does.
    void fill24(float out[24]) {
        const float twom24 = 1.0f / 16777216.0f;
        for (int j = 0; j < 24; j++) {
            int uni = seeds[j24 - 1] - seeds[i24 - 1] - carry;
            if (uni < 0) { uni += 16777216; carry = 1; } else carry = 0;
            seeds[i24 - 1] = uni;
            i24 = next[i24 - 1]; j24 = next[j24 - 1];
            if (uni >= 4096) {
                out[j] = (float)uni * twom24;
            } else {
                volatile float a = (float)uni * twom24;
                volatile float b = (float)seeds[j24 - 1] * twom24 * twom24;
                out[j] = a + b;
            }
        }
        for (int j = 0; j < nskip; j++) {
            int uni = seeds[j24 - 1] - seeds[i24 - 1] - carry;
            if (uni < 0) { uni += 16777216; carry = 1; } else carry = 0;
            seeds[i24 - 1] = uni;
            i24 = next[i24 - 1]; j24 = next[j24 - 1];
        }
    }
};

// $RANDOMSET over ranlux: 24-value buffer refilled on demand
// (ranlux.macros:71-78).
struct RanluxStream {
    Ranlux g;
    float buf[24];
    int ip = 24;
    void init(int lux, int seed) { g.init(lux, seed); ip = 24; }
    double uniform() {
        if (ip >= 24) { g.fill24(buf); ip = 0; }
        return buf[ip++];
    }
};

}  // namespace oracle
