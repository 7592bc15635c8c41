// egsgpu_transport.h -- the per-lane shower loop and interaction samplers.
//
// One lane runs one ranmar stream (an egs-parallel job in miniature) and
// simulates its histories one after the other; the particle on top of the
// lane's stack is held in registers, the rest of the stack in HBM.
// Routines restate src/egsnrc.mortran (PHOTON :6439-6643, COMPT :528-806,
// PHOTO :5050-5196, PAIR :4518-4731, UPHI :6723-6847, ANNIH_AT_REST
// :186-224, egs_rayleigh_sampling :8559-8590, the ECUT discard of ELECTR
// :1773-1800) with the C-interface callbacks
// (egs_advanced_application.cpp:1351-1441) and dose scoring
// (egs_dose_scoring.h:203-229, egs_scoring.h:63-125).
#pragma once
#include "egsgpu_device.h"

namespace egsgpu {

// A history taking more steps than this is abandoned and counted as an
// overflow; it bounds every lane's work even on corrupted input.
constexpr unsigned kMaxStepsPerHistory = 4000000u;

template <typename R>
struct Part {
    R e, x, y, z, u, v, w, wt, dnear;
    int iq, ir;
};

template <typename R>
__device__ __forceinline__ R dlog(R x) { return log(x); }
template <typename R>
__device__ __forceinline__ R dexp(R x) { return exp(x); }
template <typename R>
__device__ __forceinline__ R dsqrt(R x) { return sqrt(x); }

template <typename R>
struct Shower {
    const DevTables<R> &T;
    const GeomView<R> &G;
    const uint8_t *__restrict__ vmed;
    const R *__restrict__ vrho;
    const DevParams &P;
    const DevLanes<R> &L;
    int lane;
    // RNG
    RanmarRec *rrec;
    float *rbuf;
    int rip;
    // particle stack: top in registers
    Part<R> p;
    int np;
    // UPHI saved state
    R A, B, C, cosphi, sinphi;
    // history tally hash
    int hused;
    int *hkey;
    double *hval;
    int *hlist;
    int cvox;          // run-length cache: voxel and energy
    double cval;
    const DevTally &tally;
    // counters
    unsigned long long c_psteps, c_cross, c_dep, c_flush, c_esteps, c_ovf, c_hovf;
    unsigned hsteps;   // steps of the current history (runaway guard)

    __device__ Shower(const DevTables<R> &t, const GeomView<R> &g,
                      const uint8_t *vm, const R *vr, const DevParams &pp,
                      const DevLanes<R> &l, const DevTally &ta, int ln,
                      float *rb)
        : T(t), G(g), vmed(vm), vrho(vr), P(pp), L(l), lane(ln), rbuf(rb),
          tally(ta) {
        rrec = L.rng + lane;
        rip = kRNB;
        np = 0;
        A = B = C = 0; cosphi = 1; sinphi = 0;
        hused = 0;
        hkey = L.hkey + (size_t)lane * L.hcap;
        hval = L.hval + (size_t)lane * L.hcap;
        hlist = L.hlist + (size_t)lane * L.hcap;
        cvox = -1; cval = 0;
        c_psteps = c_cross = c_dep = c_flush = c_esteps = c_ovf = c_hovf = 0;
        hsteps = 0;
    }

    __device__ __forceinline__ R rnd() {
        if (rip >= kRNB) {
            ranmar_refill<kRNB>(rrec, rbuf, blockDim.x);
            rip = 0;
        }
        return (R)rbuf[(rip++) * blockDim.x];
    }

    // ---------------------------------------------------------- scoring
    __device__ void hash_add(int vox, double e) {
        unsigned key = (unsigned)vox + 1u;
        unsigned mask = (unsigned)L.hcap - 1u;
        for (int attempt = 0; attempt < 2; attempt++) {
            unsigned s = (key * 2654435761u) & mask;
            for (int probe = 0; probe < L.hcap; probe++) {
                int k = hkey[s];
                if (k == (int)key) { hval[s] += e; return; }
                if (k == 0) {
                    if (hused * 4 >= L.hcap * 3) break;
                    hkey[s] = (int)key;
                    hval[s] = e;
                    hlist[hused++] = (int)s;
                    return;
                }
                s = (s + 1) & mask;
            }
            // table full: close the partial history early (counted)
            c_hovf++;
            flush_table();
        }
    }
    __device__ void flush_table() {
        for (int k = 0; k < hused; k++) {
            int s = hlist[k];
            int vox = hkey[s] - 1;
            double e = hval[s];
            atomicAdd(tally.sum + vox, e);
            atomicAdd(tally.sum2 + vox, e * e);
            hkey[s] = 0;
        }
        c_flush += hused;
        hused = 0;
    }
    // ausgab with iarg <= 4: score edep*wt in region ir (>= 0)
    __device__ __forceinline__ void score(int ir, R edep, R wt) {
        if (ir < 0 || !(edep > 0)) return;
        c_dep++;
        double e = (double)edep * (double)wt;
        if (ir == cvox) { cval += e; return; }
        if (cvox >= 0) hash_add(cvox, cval);
        cvox = ir;
        cval = e;
    }
    __device__ void end_history() {
        hsteps = 0;
        if (cvox >= 0) hash_add(cvox, cval);
        cvox = -1;
        cval = 0;
        flush_table();
    }

    // ------------------------------------------------------------ stack
    __device__ __forceinline__ void store(int k, const Part<R> &q) {
        size_t i = (size_t)k * L.nlanes + lane;
        L.se[i] = q.e; L.sx[i] = q.x; L.sy[i] = q.y; L.sz[i] = q.z;
        L.su[i] = q.u; L.sv[i] = q.v; L.sw[i] = q.w; L.swt[i] = q.wt;
        L.sdn[i] = q.dnear; L.siq[i] = q.iq; L.sir[i] = q.ir;
    }
    __device__ __forceinline__ void load(int k, Part<R> &q) {
        size_t i = (size_t)k * L.nlanes + lane;
        q.e = L.se[i]; q.x = L.sx[i]; q.y = L.sy[i]; q.z = L.sz[i];
        q.u = L.su[i]; q.v = L.sv[i]; q.w = L.sw[i]; q.wt = L.swt[i];
        q.dnear = L.sdn[i]; q.iq = L.siq[i]; q.ir = L.sir[i];
    }
    // push a copy of the top ($TRANSFER PROPERTIES); the copy is the new top
    __device__ __forceinline__ bool push_copy() {
        if (np >= L.mxstk) { c_ovf++; return false; }
        store(np - 1, p);
        np++;
        return true;
    }
    __device__ __forceinline__ void pop() {
        np--;
        if (np > 0) load(np - 1, p);
    }
    // the particle just below the top
    __device__ __forceinline__ void below(Part<R> &q) { load(np - 2, q); }
    __device__ __forceinline__ void set_below_iq(int iq) {
        L.siq[(size_t)(np - 2) * L.nlanes + lane] = iq;
    }

    // ------------------------------------------------------------- UPHI
    __device__ void azimuth(R &cphi, R &sphi) {
        R xphi, xphi2, yphi, yphi2, rhophi2;
        do {
            xphi = rnd();
            xphi = 2 * xphi - 1;
            xphi2 = xphi * xphi;
            yphi = rnd();
            yphi2 = yphi * yphi;
            rhophi2 = xphi2 + yphi2;
        } while (rhophi2 > 1);
        rhophi2 = 1 / rhophi2;
        cphi = (xphi2 - yphi2) * rhophi2;
        sphi = 2 * xphi * yphi * rhophi2;
    }
    __device__ __forceinline__ void rotate(R costhe, R sinthe) {
        R sinps2 = A * A + B * B;
        if (sinps2 < (R)1.0e-20) {
            p.u = sinthe * cosphi;
            p.v = sinthe * sinphi;
            p.w = C * costhe;
        } else {
            R sinpsi = dsqrt(sinps2);
            R us = sinthe * cosphi, vs = sinthe * sinphi;
            R sindel = B / sinpsi, cosdel = A / sinpsi;
            p.u = C * cosdel * us - sindel * vs + A * costhe;
            p.v = C * sindel * us + cosdel * vs + B * costhe;
            p.w = -sinpsi * us + C * costhe;
        }
    }
    __device__ __forceinline__ void uphi21(R costhe, R sinthe) {
        azimuth(cosphi, sinphi);
        A = p.u; B = p.v; C = p.w;
        rotate(costhe, sinthe);
    }
    __device__ __forceinline__ void uphi32(R costhe, R sinthe) { rotate(costhe, sinthe); }

    // ----------------------------------------------------------- PHOTON
    __device__ void photon() {
        const R pcut = (R)P.pcut;
        R peig = p.e, eig = peig;
        int medium = vmed[p.ir];
        R rhof = vrho[p.ir];
        if (eig <= pcut) { score(p.ir, peig, p.wt); pop(); return; }
        for (;;) {   // :PNEWENERGY:
            if (p.wt == 0) { pop(); return; }
            R gle = dlog(eig);
            R rnno35 = rnd();
            if (rnno35 == 0) rnno35 = (R)1.e-30;
            R dpmfp = -dlog(rnno35);
            R cohfac = 1;
            int lgle = 0;
            bool interact = false;
            while (!interact) {   // :PNEWMEDIUM:
                const MedScalars<R> &M = T.ms[medium];
                lgle = (int)(M.ge1 * gle + M.ge0);
                lgle = lgle < 1 ? 1 : (lgle > kGE ? kGE : lgle);
                int li = medium * kGE + lgle - 1;
                R gmfpr0 = T.gmfp1[li] * gle + T.gmfp0[li];
                if (P.rayleigh) cohfac = T.cohe1[li] * gle + T.cohe0[li];
                for (;;) {   // :PTRANS:
                    c_psteps++;
                    if (++hsteps > kMaxStepsPerHistory) { c_ovf++; np = 0; return; }
                    R gmfp = gmfpr0 / rhof;
                    if (P.rayleigh) gmfp = gmfp * cohfac;
                    R ustep = gmfp * dpmfp;
                    int irnew = p.ir, idisc = 0;
                    if (ustep > p.dnear || p.wt <= 0) {
                        int inew = geom_howfar_in(G, p.ir, p.x, p.y, p.z, p.u, p.v, p.w, ustep);
                        if (inew != p.ir) { irnew = inew; if (inew < 0) idisc = -1; }
                    }
                    p.x = p.x + p.u * ustep;
                    p.y = p.y + p.v * ustep;
                    p.z = p.z + p.w * ustep;
                    p.dnear = p.dnear - ustep;
                    R d = dpmfp - ustep / gmfp;
                    dpmfp = d > 0 ? d : (R)0;
                    int medold = medium;
                    if (irnew != p.ir) {
                        c_cross++;
                        p.ir = irnew;
                        if (irnew >= 0) { medium = vmed[irnew]; rhof = vrho[irnew]; }
                        else medium = -1;
                    }
                    if (idisc < 0) { pop(); return; }   // left the grid: ausgab(3), not scored
                    if (medium != medold) break;
                    if (dpmfp <= (R)1.e-8) { interact = true; break; }
                }
            }
            const MedScalars<R> &M = T.ms[medium];
            int li = medium * kGE + lgle - 1;
            if (P.rayleigh) {
                R rnno37 = rnd();
                if (rnno37 <= (R)1.0 - cohfac) {
                    R costhe, sinthe;
                    rayleigh_sample(medium, p.e, gle, li, costhe, sinthe);
                    uphi21(costhe, sinthe);
                    continue;
                }
            }
            R rnno36 = rnd();
            R gbr1 = T.gbr11[li] * gle + T.gbr10[li];
            if (rnno36 <= gbr1 && p.e > (R)(2 * 0.5109989461)) {
                pair(medium, M);
                if (p.iq != 0) return;
            } else {
                R gbr2 = T.gbr21[li] * gle + T.gbr20[li];
                if (rnno36 < gbr2) {
                    compt(medium, M);
                    if (p.iq != 0) return;
                } else {
                    photo(medium, M, gle);
                    if (p.iq != 0) return;
                }
            }
            peig = p.e;
            eig = peig;
            if (eig < pcut) { score(p.ir, peig, p.wt); pop(); return; }
        }
    }

    __device__ void rayleigh_sample(int m, R e, R gle, int li, R &costhe, R &sinthe) {
        const double hc_i = 80.65506856998, twice_hc2 = 0.000307444456;
        const double *fcum = T.ray_fcum + m * 100;
        const double *xg = T.ray_x + m * 100;
        const double *bb = T.ray_b + m * 100;
        const double *cc = T.ray_c + m * 100;
        const int *ia = T.ray_i + m * 100;
        R pmax = T.pmax1[li] * gle + T.pmax0[li];
        R xmax = (R)hc_i * e, xv, csqthe, rnnray1;
        do {
            rnnray1 = rnd();
            do {
                double r0 = (double)rnd() * (double)pmax;
                int ibin = (int)(1 + r0 * 99.0);
                ibin = ibin > 99 ? 99 : ibin;
                int ib = ia[ibin - 1];
                if (ia[ibin] > ib) {
                    while (ib < 99 && !(r0 < fcum[ib])) ib++;
                }
                r0 = (r0 - fcum[ib - 1]) * cc[ib - 1];
                xv = (R)(xg[ib - 1] * exp(log(1 + r0) * bb[ib - 1]));
            } while (!(xv < xmax));
            xv = xv / e;
            costhe = 1 - (R)twice_hc2 * xv * xv;
            csqthe = costhe * costhe;
        } while (!(2 * rnnray1 < 1 + csqthe));
        sinthe = dsqrt(1 - csqthe);
    }

    // ------------------------------------------------------------ COMPT
    __device__ void compt(int m, const MedScalars<R> &M) {
        const R RM = (R)0.5109989461;
        R peig = p.e;
        R ko = peig / RM;
        R broi = 1 + 2 * ko;
        int ibcmpl = P.ibcmp;
        bool first_time = true;
        R Uj = 0, Jo = 0;
        int j = 0;
        R broi2 = 0, alph1 = 0, bro = 0, alpha = 0, bro1 = 0, rejmax = 0;
        R br = 0, temp = 0, sinthe = 0, costhe = 0, pz = 0, af = 0, Fmax = 0, fpz = 0, qc = 0, qc2 = 0, br2 = 0;
        for (;;) {   // :resample-compton:
            if (ibcmpl > 0) {
                R rnno17 = rnd();
                rnno17 = 1 + rnno17 * M.n_shell;
                int i = (int)rnno17;
                i = i > M.n_shell ? M.n_shell : i;
                if (rnno17 > T.eno[m * kMaxShMed + i - 1]) i = T.eno_atbin[m * kMaxShMed + i - 1];
                j = T.shell[m * kMaxShMed + i - 1];
                Uj = T.cs_be[j];
                if (ko <= Uj) {
                    if (ibcmpl == 1) return;
                    continue;
                }
                Jo = T.cs_Jo[j];
            }
            for (;;) {   // :RESAMPLE:
                if (ko > 2) {
                    if (first_time) {
                        broi2 = broi * broi;
                        alph1 = dlog(broi);
                        bro = 1 / broi;
                        R alph2 = ko * (broi + 1) * bro * bro;
                        alpha = alph1 + alph2;
                    }
                    R aux, rejf3, rnno19;
                    do {
                        R rnno15 = rnd(), rnno16 = rnd();
                        if (rnno15 * alpha < alph1) br = dexp(alph1 * rnno16) * bro;
                        else br = dsqrt(rnno16 * broi2 + (1 - rnno16)) * bro;
                        temp = (1 - br) / (ko * br);
                        sinthe = temp * (2 - temp);
                        if (sinthe < 0) sinthe = 0;
                        aux = 1 + br * br;
                        rejf3 = aux - br * sinthe;
                        rnno19 = rnd();
                    } while (!(rnno19 * aux <= rejf3));
                } else {
                    if (first_time) {
                        bro = 1 / broi;
                        bro1 = 1 - bro;
                        rejmax = broi + bro;
                    }
                    R rejf3, rnno16;
                    do {
                        R rnno15 = rnd();
                        rnno16 = rnd();
                        br = bro + bro1 * rnno15;
                        temp = (1 - br) / (ko * br);
                        sinthe = temp * (2 - temp);
                        if (sinthe < 0) sinthe = 0;
                        rejf3 = 1 + br * br - br * sinthe;
                    } while (!(rnno16 * br * rejmax <= rejf3));
                }
                first_time = false;
                if (!(br < bro || br > 1)) break;
            }
            costhe = 1 - temp;
            if (P.ibcmp == 0) { Uj = 0; break; }
            br2 = br * br;
            R aux = ko * (ko - Uj) * temp;
            R aux1 = 2 * aux + Uj * Uj;
            R pzmax = aux - Uj;
            if (pzmax < 0 && pzmax * pzmax >= aux1) {
                if (ibcmpl == 1) return;
                continue;
            }
            pzmax = pzmax / dsqrt(aux1);
            qc2 = 1 + br * br - 2 * br * costhe;
            qc = dsqrt(qc2);
            bool retry_direct = false;
            if (pzmax > 1) {
                pzmax = 1; af = 0; Fmax = 1; fpz = 1;
                retry_direct = true;
            } else {
                R aux3 = 1 + 2 * Jo * fabs(pzmax);
                R aux4 = (R)0.5 * (1 - aux3 * aux3);
                fpz = (R)0.5 * dexp(aux4);
                af = qc * (1 + br * (br - costhe) / qc2);
                if (af < 0) {
                    if (pzmax > 0) fpz = 1 - fpz;
                    R eta_incoh = rnd();
                    if (eta_incoh > fpz) {
                        if (ibcmpl == 1) return;
                        continue;
                    }
                    af = 0; Fmax = 1;
                    retry_direct = true;
                } else {
                    R fpz1;
                    if (pzmax < (R)-0.15) {
                        Fmax = 1 - af * (R)0.15;
                        fpz1 = fpz * Fmax * Jo;
                    } else if (pzmax < (R)0.15) {
                        Fmax = 1 + af * pzmax;
                        aux3 = 1 / (1 + (R)0.33267252734 * aux3);
                        aux4 = fpz * aux3 * ((R)0.3480242 + aux3 * ((R)-0.0958798 + aux3 * (R)0.7478556)) + T.cs_erfJo[j];
                        if (pzmax > 0) {
                            fpz1 = (1 - Fmax * fpz) * Jo - (R)0.62665706866 * af * aux4;
                            fpz = 1 - fpz;
                        } else {
                            fpz1 = Fmax * fpz * Jo - (R)0.62665706866 * af * aux4;
                        }
                    } else {
                        Fmax = 1 + af * (R)0.15;
                        fpz1 = (1 - Fmax * fpz) * Jo;
                        fpz = 1 - fpz;
                    }
                    R eta_incoh = rnd();
                    if (eta_incoh * Jo > fpz1) {
                        if (ibcmpl == 1) return;
                        continue;
                    }
                }
            }
            (void)retry_direct;
            // :RETRY-PZ:
            if (ibcmpl != 2) {
                for (;;) {
                    R rnno18 = rnd();
                    rnno18 = rnno18 * fpz;
                    if (rnno18 < (R)0.5) {
                        rnno18 = 2 * rnno18;
                        if (rnno18 < (R)1e-30) rnno18 = (R)1e-30;
                        pz = (R)0.5 * (1 - dsqrt(1 - 2 * dlog(rnno18))) / Jo;
                    } else {
                        rnno18 = 2 * (1 - rnno18);
                        pz = (R)0.5 * (dsqrt(1 - 2 * dlog(rnno18)) - 1) / Jo;
                    }
                    if (fabs(pz) > 1) continue;
                    if (pz < (R)0.15) {
                        R frej = pz < (R)-0.15 ? 1 - af * (R)0.15 : 1 + af * pz;
                        R eta = rnd();
                        if (eta * Fmax > frej) continue;
                    }
                    break;
                }
            } else {
                pz = 0; Uj = 0;
            }
            R pz2 = pz * pz;
            if (fabs(pz) < (R)0.01) {
                br = br * (1 + pz * (qc + (br2 - costhe) * pz));
            } else {
                R a = 1 - pz2 * br * costhe, a1 = 1 - pz2 * br2;
                R a2 = qc2 - br2 * pz2 * sinthe;
                if (a2 > (R)1e-10) br = br / a1 * (a + pz * dsqrt(a2));
            }
            Uj = Uj * RM;
            break;
        }
        // :FINISHED-COMPTON-SAMPLING:
        R pesg = br * peig;
        R pese = peig - pesg - Uj + RM;
        sinthe = dsqrt(sinthe);
        uphi21(costhe, sinthe);
        p.e = pesg;
        R aux = 1 + br * br - 2 * br * costhe;
        if (aux > (R)1e-8) {
            costhe = (1 - br * costhe) / dsqrt(aux);
            sinthe = (1 - costhe) * (1 + costhe);
            sinthe = sinthe > 0 ? -dsqrt(sinthe) : (R)0;
        } else {
            costhe = 0;
            sinthe = -1;
        }
        if (!push_copy()) return;
        uphi32(costhe, sinthe);
        p.e = pese;
        p.iq = -1;
        if (ibcmpl == 1 || ibcmpl == 3) {
            // vacancy energy absorbed locally (relaxation products below cut-offs)
            if (Uj > 0) score(p.ir, Uj, p.wt);
        }
    }

    // ------------------------------------------------------------ PHOTO
    __device__ void photo(int m, const MedScalars<R> &M, R gle) {
        const R RM = (R)0.5109989461;
        R peig = p.e;
        if (peig < T.ee[1 * 16 + 1]) { p.iq = -1; p.e = peig + RM; return; }
        bool do_relax = false;
        R e_vac = 0;
        if (P.iedgfl) {
            int iZ;
            if (M.nne == 1) {
                iZ = (int)(T.z[m * kMaxNE] + (R)0.5);
            } else {
                R aux = peig * peig, aux1 = aux * peig;
                aux = aux * dsqrt(peig);
                R sigtot = 0, probs[kMaxNE];
                for (int k = 0; k < M.nne; k++) {
                    int iz = (int)(T.z[m * kMaxNE + k] + (R)0.5);
                    R sigma;
                    if (peig > T.ee[iz * 16]) {
                        sigma = (T.ea[iz * 16] + T.eb[iz * 16] / peig + T.ec[iz * 16] / aux +
                                 T.ed[iz * 16] / aux1) / peig;
                    } else {
                        int jj;
                        int ne = T.enumb[iz];
                        for (jj = 2; jj <= ne; jj++)
                            if (peig >= T.ee[iz * 16 + jj - 1]) break;
                        if (jj > ne) jj = ne;
                        int q = iz * 16 + jj - 1;
                        sigma = T.ea[q] + gle * (T.eb[q] + gle * (T.ec[q] + gle * T.ed[q]));
                        sigma = dexp(sigma);
                    }
                    sigma = sigma * T.pz[m * kMaxNE + k];
                    sigtot += sigma;
                    probs[k] = sigma;
                }
                R br = rnd();
                br = br * sigtot;
                int k;
                for (k = 0; k < M.nne; k++) {
                    br -= probs[k];
                    if (br <= 0) break;
                }
                if (k >= M.nne) k = M.nne - 1;
                iZ = (int)(T.z[m * kMaxNE + k] + (R)0.5);
            }
            if (peig <= T.bind[iZ * 30 + 5]) {
                p.iq = -1;
                p.e = peig + RM;
            } else {
                R br = rnd();
                int k;
                for (k = 1; k <= 5; k++) {
                    if (peig > T.bind[iZ * 30 + k - 1]) {
                        R ip = T.iprob[iZ * 6 + k - 1];
                        if (br < ip) break;
                        br = (br - ip) / (1 - ip);
                    }
                }
                if (k > 4) {
                    p.iq = -1;
                    p.e = peig + RM;
                } else {
                    e_vac = T.bind[iZ * 30 + k - 1];
                    p.e = peig - e_vac + RM;
                    do_relax = true;
                    p.iq = -1;
                }
            }
        } else {
            p.e = peig + RM;
            p.iq = -1;
        }
        if (P.iphter) photoelectron_direction();
        if (do_relax) score(p.ir, e_vac, p.wt);
    }

    __device__ void photoelectron_direction() {
        const R RM = (R)0.5109989461;
        R eelec = p.e;
        if (!(eelec > (R)P.ecut)) return;
        R beta = dsqrt((eelec - RM) * (eelec + RM)) / eelec;
        R gamma = eelec / RM;
        R alpha = (R)0.5 * gamma - (R)0.5 + 1 / gamma;
        R ratio = beta / alpha;
        R costhe, sinth2, xi, rnpht2;
        do {
            R rnpht = rnd();
            rnpht = 2 * rnpht - 1;
            if (ratio <= (R)0.2) {
                R fkappa = rnpht + (R)0.5 * ratio * (1 - rnpht) * (1 + rnpht);
                if (gamma < 100) {
                    costhe = (beta + fkappa) / (1 + beta * fkappa);
                } else {
                    if (fkappa > 0) {
                        R g1 = gamma - 1;
                        costhe = 1 - (1 - fkappa) * (gamma - 3) / (2 * (1 + fkappa) * g1 * g1 * g1);
                    } else {
                        costhe = (beta + fkappa) / (1 + beta * fkappa);
                    }
                }
                xi = (1 + beta * fkappa) * gamma * gamma;
            } else {
                xi = gamma * gamma * (1 + alpha * (dsqrt(1 + ratio * (2 * rnpht + ratio)) - 1));
                costhe = (1 - 1 / xi) / beta;
            }
            sinth2 = (1 - costhe) * (1 + costhe);
            if (sinth2 < 0) sinth2 = 0;
            rnpht2 = rnd();
        } while (rnpht2 > (R)0.5 * (1 + gamma) * sinth2 * xi / gamma);
        uphi21(costhe, dsqrt(sinth2));
    }

    // ------------------------------------------------------------- PAIR
    __device__ void pair(int m, const MedScalars<R> &M) {
        const R RM = (R)0.5109989461, RMT2 = 2 * RM;
        R peig = p.e, eig = peig;
        R pese1, pese2;
        int iq1, iq2;
        const R *dl = T.dl + m * 48;   // dl[k*8 + l]
        if (eig <= (R)2.1) {
            R rnno30 = rnd(), rnno34 = rnd();
            pese2 = RM + (R)0.5 * rnno30 * (peig - 2 * RM);
            pese1 = peig - pese2;
            if (rnno34 < (R)0.5) { iq1 = -1; iq2 = 1; } else { iq1 = 1; iq2 = -1; }
        } else {
            int l, l1;
            R Amax, Bmax, aux1;
            if (eig < 50) {
                l = 5; l1 = l + 1;
                R delta = 4 * M.delcm / eig;
                if (delta < 1) {
                    Amax = dl[l - 1] + delta * (dl[8 + l - 1] + delta * dl[16 + l - 1]);
                    Bmax = dl[l1 - 1] + delta * (dl[8 + l1 - 1] + delta * dl[16 + l1 - 1]);
                } else {
                    R aux2 = dlog(delta + dl[40 + l - 1]);
                    Amax = dl[24 + l - 1] + dl[32 + l - 1] * aux2;
                    Bmax = dl[24 + l1 - 1] + dl[32 + l1 - 1] * aux2;
                }
                aux1 = 1 - RMT2 / eig;
                aux1 = aux1 * aux1;
                aux1 = aux1 * Amax / 3;
                aux1 = aux1 / (Bmax + aux1);
            } else {
                l = 7;
                Amax = dl[l - 1];
                Bmax = dl[l];
                aux1 = M.bpar1 * (1 - M.bpar0 * RM / eig);
            }
            R del0 = eig * M.delcm;
            R Eavail = eig - RMT2;
            R Eminus, rejf, rejmax, rnno34;
            do {
                R rnno30 = rnd(), rnno31 = rnd();
                rnno34 = rnd();
                R br;
                if (rnno30 > aux1) {
                    br = (R)0.5 * rnno31;
                    rejmax = Bmax;
                    l1 = l + 1;
                } else {
                    R rnno32 = rnd(), rnno33 = rnd();
                    R mx = rnno31 > rnno32 ? rnno31 : rnno32;
                    mx = mx > rnno33 ? mx : rnno33;
                    br = (R)0.5 * (1 - mx);
                    rejmax = Amax;
                    l1 = l;
                }
                Eminus = br * Eavail + RM;
                R Eplus = eig - Eminus;
                R delta = del0 / (Eminus * Eplus);
                if (delta < 1)
                    rejf = dl[l1 - 1] + delta * (dl[8 + l1 - 1] + delta * dl[16 + l1 - 1]);
                else
                    rejf = dl[24 + l1 - 1] + dl[32 + l1 - 1] * dlog(delta + dl[40 + l1 - 1]);
            } while (!(rnno34 * rejmax <= rejf));
            pese2 = Eminus;
            pese1 = peig - pese2;
            rnno34 = rnd();
            if (rnno34 < (R)0.5) { iq1 = -1; iq2 = 1; } else { iq1 = 1; iq2 = -1; }
        }
        p.e = pese1;
        if (P.iprdst > 0) {
            for (int ichrg = 1; ichrg <= 2; ichrg++) {
                R ese = ichrg == 1 ? pese1 : pese2;
                R pse = (ese - RM) * (ese + RM);
                pse = pse > 0 ? dsqrt(pse) : (R)0;
                R costhe = rnd();
                costhe = 1 - 2 * costhe;
                R sinthe = RM * dsqrt((1 - costhe) * (1 + costhe)) / (pse * costhe + ese);
                costhe = (ese * costhe + pse) / (pse * costhe + ese);
                if (ichrg == 1) {
                    uphi21(costhe, sinthe);
                } else {
                    if (!push_copy()) return;
                    p.e = pese2;
                    uphi32(costhe, -sinthe);
                }
            }
        } else {
            if (!push_copy()) return;
            p.e = pese2;
        }
        p.iq = iq2;
        set_below_iq(iq1);
    }

    // ---------------------------------------------------- charged (kerma)
    __device__ void electron() {
        const R RM = (R)0.5109989461;
        int lelec = p.iq;
        R peie = p.e;
        // :ECUT-DISCARD: (every charged particle is below ECUT in kerma mode)
        c_esteps++;
        R edep = peie - RM;
        score(p.ir, edep, p.wt);
        if (lelec > 0 && edep < peie) { annih_at_rest(); return; }
        pop();
    }

    __device__ void annih_at_rest() {
        const R RM = (R)0.5109989461;
        R costhe = rnd();
        costhe = 2 * costhe - 1;
        R s2 = (1 - costhe) * (1 + costhe);
        R sinthe = dsqrt(s2 > 0 ? s2 : (R)0);
        R cphi, sphi;
        azimuth(cphi, sphi);
        p.e = RM; p.iq = 0;
        p.u = sinthe * cphi; p.v = sinthe * sphi; p.w = costhe;
        if (!push_copy()) return;
        p.u = -p.u; p.v = -p.v; p.w = -p.w;
    }
};

}  // namespace egsgpu
