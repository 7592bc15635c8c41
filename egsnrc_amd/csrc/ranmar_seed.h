// ranmar_seed.h -- the seeding of the ranmar generator, shared by the
// device engine, the host application and the CPU oracle, so that every
// part of the repository builds the same 97-word lag table from a seed
// pair (the integer recurrences must be identical for bit-exact streams).
//
//  ranmar_seed_words: EGS_Ranmar::setState (egs++/egs_rndm.cpp:303-336),
//      the same table as init_ranmar of src/ranmar.mortran;
//  ranmar_stream_seeds: the seed pair of stream s, (ixx, jxx + s) as
//      EGS_RandomGenerator::createRNG(input, s) builds it for parallel job s
//      (egs_rndm.cpp:438-446); beyond jxx + s = 30080 the excess is carried
//      into ixx, so that every stream keeps a distinct valid pair
//      (setState would otherwise reset them all to 9373, egs_rndm.cpp:307-309).
//
// Plain C++ usable from host and device code (no library calls).
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define EGS_RANMAR_HD __host__ __device__
#else
#define EGS_RANMAR_HD
#endif

// Words of the lag table for seeds (ixx, jxx); an out-of-range seed takes
// the default (1802 or 9373), as setState does.  After seeding the state
// words are c = 362436, ix = 96, jx = 32 (kRanmarC0, kRanmarIx0, kRanmarJx0).
EGS_RANMAR_HD inline void ranmar_seed_words(int ixx, int jxx, int *u) {
    if (ixx <= 0 || ixx >= 31328) ixx = 1802;
    if (jxx <= 0 || jxx >= 30081) jxx = 9373;
    int i = (ixx / 177) % 177 + 2;
    int j = ixx % 177 + 2;
    int k = (jxx / 169) % 178 + 1;
    int l = jxx % 169;
    for (int ii = 0; ii < 97; ii++) {
        int s = 0, t = 8388608;
        for (int jj = 0; jj < 24; jj++) {
            const int m = (((i * j) % 179) * k) % 179;
            i = j;
            j = k;
            k = m;
            l = (53 * l + 1) % 169;
            if ((l * m) % 64 >= 32) s += t;
            t >>= 1;
        }
        u[ii] = s;
    }
}
constexpr int kRanmarC0 = 362436, kRanmarIx0 = 96, kRanmarJx0 = 32;

EGS_RANMAR_HD inline void ranmar_stream_seeds(int ixx, int jxx, int64_t s, int &is, int &js) {
    if (ixx <= 0 || ixx >= 31328) ixx = 1802;
    if (jxx <= 0 || jxx >= 30081) jxx = 9373;
    const int64_t t = (int64_t)jxx + s;   // >= 1
    const int64_t carry = (t - 1) / 30080;
    js = (int)(1 + (t - 1) % 30080);
    is = (int)(1 + ((int64_t)ixx - 1 + carry) % 31327);
}
