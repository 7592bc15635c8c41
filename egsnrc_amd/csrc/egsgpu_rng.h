// egsgpu_rng.h -- device random number generators (gfx950).
//
// ranmar: the Marsaglia-Zaman generator of EGS_Ranmar
//         (egs++/egs_rndm.cpp:303-398, src/ranmar.mortran), one stream per
//         lane.  The 97-word lag table lives in HBM (one 400 B record per
//         stream, so a refill touches two short contiguous runs of it);
//         the outputs are produced RNG_NB at a time into an LDS ring, the
//         device analogue of the 128-entry rng_array of $RANDOMSET
//         (interface/egs_c_interface2.macros:73-80).
// ranlux: Luescher's generator as in src/ranlux.mortran; the 24-word state
//         is kept rotated so that i24 is always word 0 and the per-call
//         update is a fully unrolled register computation.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ranmar_seed.h"

namespace egsgpu {

constexpr int kRNB = 16;          // ranmar outputs per refill
constexpr int kRing = 32;         // LDS ring of random numbers per lane
constexpr int kRingLow = 12;      // top the ring up below this many

// Layout of one ranmar stream record (int32 words).
struct RanmarRec {
    int u[97];
    int c, ix, jx;
};
static_assert(sizeof(RanmarRec) == 400, "ranmar record is 400 bytes");

// Seed pair of stream s and the lag table of a seed pair: ranmar_seed.h
// (createRNG(input, s), egs_rndm.cpp:438-446; EGS_Ranmar::setState,
// egs_rndm.cpp:303-336).
__host__ __device__ inline void stream_seeds(int ixx, int jxx, int64_t s, int &is, int &js) {
    ranmar_stream_seeds(ixx, jxx, s, is, js);
}
__device__ inline void ranmar_seed(RanmarRec *r, int ixx, int jxx) {
    ranmar_seed_words(ixx, jxx, r->u);
    r->c = kRanmarC0;
    r->ix = kRanmarIx0;
    r->jx = kRanmarJx0;
}

// Produce NB consecutive outputs of the stream (EGS_Ranmar::fillArray,
// egs_rndm.cpp:338-370) as exact floats r*2^-24 into buf[k*stride].
//
// ix and jx step down together and always differ by 33 (mod 97): the NB
// words updated, u[ix-k], and the NB words read, u[jx-k], k < NB, are
// disjoint sets for NB <= 33.  So all 2*NB loads are issued before any
// arithmetic (one memory round trip per refill instead of NB dependent
// ones), and the updated words are stored at the end.
// The outputs go to buf[((start + k) & mask) * stride] (a ring in LDS).
template <int NB>
__device__ inline void ranmar_refill(RanmarRec *__restrict__ r,
                                     float *__restrict__ buf, int stride,
                                     int start = 0, int mask = NB - 1) {
    static_assert(NB >= 1 && NB <= 33, "ranmar refill runs must not overlap");
    int c = r->c, ix = r->ix, jx = r->jx;
    const float twom24 = 1.0f / 16777216.0f;
    int a[NB], b[NB];
#pragma unroll
    for (int k = 0; k < NB; k++) {
        int i = ix - k, j = jx - k;
        i += (i < 0) ? 97 : 0;
        j += (j < 0) ? 97 : 0;
        a[k] = r->u[i];
        b[k] = r->u[j];
    }
#pragma unroll
    for (int k = 0; k < NB; k++) {
        int v = a[k] - b[k];
        v += (v < 0) ? 16777216 : 0;
        a[k] = v;
        c -= 7654321;
        c += (c < 0) ? 16777213 : 0;
        v -= c;
        v += (v < 0) ? 16777216 : 0;
        buf[((start + k) & mask) * stride] = (float)v * twom24;
    }
#pragma unroll
    for (int k = 0; k < NB; k++) {
        int i = ix - k;
        i += (i < 0) ? 97 : 0;
        r->u[i] = a[k];
    }
    ix -= NB; ix += (ix < 0) ? 97 : 0;
    jx -= NB; jx += (jx < 0) ? 97 : 0;
    r->c = c; r->ix = ix; r->jx = jx;
}

// The next single output of the stream, straight from its record.
__device__ __forceinline__ float ranmar_next(RanmarRec *__restrict__ r) {
    int c = r->c, ix = r->ix, jx = r->jx;
    int v = r->u[ix] - r->u[jx];
    v += (v < 0) ? 16777216 : 0;
    r->u[ix] = v;
    r->ix = (ix == 0) ? 96 : ix - 1;
    r->jx = (jx == 0) ? 96 : jx - 1;
    c -= 7654321;
    c += (c < 0) ? 16777213 : 0;
    r->c = c;
    v -= c;
    v += (v < 0) ? 16777216 : 0;
    return (float)v * (1.0f / 16777216.0f);
}

// ---- ranlux ------------------------------------------------------------
// Stream record: w[k] = seeds((i24-1-k) mod 24 + 1) (Fortran indices), so
// that the lag-14 partner of word k is word (k+14) mod 24 and the word
// read for the low bits after step k is (k+15) mod 24.
struct RanluxRec {
    int w[24];
    int carry;
    int lux;
    int pad[2];
};
static_assert(sizeof(RanluxRec) == 112, "ranlux record is 112 bytes");

// init_ranlux (ranlux.mortran): i24 = 24 at start, so w[k] = seeds(24-k).
__device__ inline void ranlux_seed(RanluxRec *r, int luxury, int seed) {
    int jseed = seed <= 0 ? 314159265 : seed;
    if (luxury < 0 || luxury > 4) luxury = 1;
    int seeds[24];
    for (int j = 0; j < 24; j++) {
        int k = jseed / 53668;
        jseed = 40014 * (jseed - k * 53668) - k * 12211;
        if (jseed < 0) jseed += 2147483563;
        seeds[j] = jseed % 16777216;
    }
    for (int k = 0; k < 24; k++) r->w[k] = seeds[23 - k];
    r->carry = seeds[23] == 0 ? 1 : 0;
    r->lux = luxury;
}

// Steps k0..k0+23 of the subtract-with-borrow recurrence on the rotated
// registers s[]; OUT selects whether outputs are written, output k going
// to buf[((start + k) & mask) * stride].
template <bool OUT>
__device__ inline void ranlux_block(int (&s)[24], int &carry,
                                    float *__restrict__ buf, int stride,
                                    int start = 0, int mask = -1) {
    const float twom24 = 1.0f / 16777216.0f;
#pragma unroll
    for (int k = 0; k < 24; k++) {
        int uni = s[(k + 14) % 24] - s[k] - carry;
        carry = uni < 0 ? 1 : 0;
        uni += uni < 0 ? 16777216 : 0;
        s[k] = uni;
        if (OUT) {
            float v;
            if (uni >= 4096) {
                v = (float)uni * twom24;
            } else {
                float a = (float)uni * twom24;
                float b = (float)s[(k + 15) % 24] * twom24 * twom24;
                v = __fadd_rn(a, b);
            }
            buf[((start + k) & mask) * stride] = v;
        }
    }
}

template <int REM>
__device__ inline void ranlux_tail(int (&s)[24], int &carry) {
#pragma unroll
    for (int k = 0; k < REM; k++) {
        int uni = s[(k + 14) % 24] - s[k] - carry;
        carry = uni < 0 ? 1 : 0;
        uni += uni < 0 ? 16777216 : 0;
        s[k] = uni;
    }
    // rotate so that the next i24 is word 0 again
    int t[24];
#pragma unroll
    for (int k = 0; k < 24; k++) t[k] = s[(k + REM) % 24];
#pragma unroll
    for (int k = 0; k < 24; k++) s[k] = t[k];
}

// ranlux(rng_array): 24 outputs then nskip = {0,24,73,199,365}[lux]
// further steps (ranlux.mortran).  nskip = 24*q + rem.  Output k goes to
// buf[((start + k) & mask) * stride].
__device__ inline void ranlux_refill(RanluxRec *__restrict__ r,
                                     float *__restrict__ buf, int stride,
                                     int start = 0, int mask = -1) {
    int s[24];
#pragma unroll
    for (int k = 0; k < 24; k++) s[k] = r->w[k];
    int carry = r->carry;
    ranlux_block<true>(s, carry, buf, stride, start, mask);
    switch (r->lux) {
        case 0: break;
        case 1: ranlux_block<false>(s, carry, buf, stride); break;
        case 2:
            for (int q = 0; q < 3; q++) ranlux_block<false>(s, carry, buf, stride);
            ranlux_tail<1>(s, carry);
            break;
        case 3:
            for (int q = 0; q < 8; q++) ranlux_block<false>(s, carry, buf, stride);
            ranlux_tail<7>(s, carry);
            break;
        default:
            for (int q = 0; q < 15; q++) ranlux_block<false>(s, carry, buf, stride);
            ranlux_tail<5>(s, carry);
            break;
    }
#pragma unroll
    for (int k = 0; k < 24; k++) r->w[k] = s[k];
    r->carry = carry;
}

// ranlux_refill as a called function, for the rare refill from inside a
// sampling loop (a ring that ran dry): one copy of the unrolled block code
// per kernel instead of one per draw site.  Only the ranlux builds of the
// kernels contain the call.
__device__ __noinline__ void ranlux_refill_call(RanluxRec *r, float *buf, int stride, int start,
                                                int mask) {
    ranlux_refill(r, buf, stride, start, mask);
}

// ---- per-lane ring -------------------------------------------------------
// A stream's next numbers wait in a ring of kRing floats per lane in LDS
// (ring[i * stride], positions pos .. pos + avail - 1 mod kRing), the
// device analogue of rng_array.  ring_topup is called once per loop turn
// with the wave converged: when any lane runs low, the lanes with room add
// a block (16 ranmar or 24 ranlux outputs).  ring_next takes the next
// number; a lane whose ring ran dry inside a sampling loop takes it
// straight from its record (ranmar: one output, inline; ranlux: a whole
// block, through ranlux_refill_call).  Either way the lane consumes its stream in
// order, whenever the refills happen.
template <bool LUX>
struct RingRec;
template <>
struct RingRec<false> { using type = RanmarRec; };
template <>
struct RingRec<true> { using type = RanluxRec; };

template <bool LUX>
__device__ __forceinline__ void ring_topup(typename RingRec<LUX>::type *rec, float *ring,
                                           int stride, int pos, int &avail) {
    if (__any(avail < kRingLow)) {
        if constexpr (LUX) {
            if (avail <= kRing - 24) {
                ranlux_refill(rec, ring, stride, pos + avail, kRing - 1);
                avail += 24;
            }
        } else {
            if (avail <= kRing - kRNB) {
                ranmar_refill<kRNB>(rec, ring, stride, pos + avail, kRing - 1);
                avail += kRNB;
            }
        }
    }
}

template <bool LUX>
__device__ __forceinline__ float ring_next(typename RingRec<LUX>::type *rec, float *ring,
                                           int stride, int &pos, int &avail) {
    if (avail == 0) {
        if constexpr (LUX) {
            ranlux_refill_call(rec, ring, stride, pos, kRing - 1);
            avail = 24;
        } else {
            return ranmar_next(rec);
        }
    }
    avail--;
    return ring[((pos++) & (kRing - 1)) * stride];
}

}  // namespace egsgpu
