// oracle_transport.cpp -- FP64 CPU restatement of the EGSnrc shower loop in
// XYZ voxel geometry.  TEST INFRASTRUCTURE ONLY: the checker of the HIP
// engine and the timed CPU baseline; never part of the product path.
//
// Control flow follows the Mortran routines it names (HEN_HOUSE/src/
// egsnrc.mortran, egsnrc.macros, interface/egs_c_interface2.macros) with
// the egs++ callbacks (egs_advanced_application.cpp:1351-1441,
// egs_application.cpp:919-959, EGS_DoseScoring egs_dose_scoring.h:203-229).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <memory>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "oracle_geom.h"
#include "oracle_rng.h"
#include "oracle_tables.h"

namespace oracle {

static const double RM = 0.5109989461, PRM = RM, RMT2 = 2 * RM, RMSQ = RM * RM;

struct Particle {
    double e, x, y, z, u, v, w, wt, dnear;
    int iq, ir;
};

struct Source {
    int charge, spectrum, nbins, shape;
    double energy, x0, x1, y0, y1, z, ux, uy, uz;
    // alias table (EGS_AliasTable, egs_alias_table.cpp)
    std::vector<double> xi, fi, wi;
    std::vector<int> bin;
    int np = 0, type = 0;
};

struct Params {
    double ecut, pcut;
    int rayleigh, ibcmp, iphter, iedgfl, iprdst, ibrdst, spin, exact_bca;
    double estepe, ximax, skindepth;
    int kerma;
    int presta1;   // transport_algorithm PRESTA-I (msdist_pI)
    int ibr_nist;  // 1 / 2: NIST / NRC brems photon energies
    int pair_nrc;  // 1: NRC pair energy sharing
    int eii;       // 1: electron impact ionisation in MOLLER
};

struct Problem {
    Tables T;
    XYZ<double> geo;
    std::vector<int> med;       // per voxel
    std::vector<double> rhor;   // per voxel
    Source src;
    Params P;
    int nvox;
};

// ------------------------------------------------------------- Shower
struct Shower {
    const Problem &pb;
    const Tables &T;
    Ranmar rng;
    std::vector<Particle> stk;    // stk[np-1] is the top (Fortran np)
    // history-by-history scoring (EGS_ScoringSingle semantics)
    std::vector<double> &tmp;
    std::vector<int> &touched;
    // UPHI saved state
    double A = 0, B = 0, Cc = 0, cosphi = 1, sinphi = 0;
    // score-function event log (density-gradient replay): deposits
    // (type 0, weight e*wt) and log-likelihood terms d log p / d rho_k
    // (type 1) with the stack depth of the particle that made them
    struct Ev { int type, vox, depth; double val; };
    std::vector<Ev> *evlog = nullptr;
    int64_t n_overflow = 0;
    int64_t c_esteps = 0;
    int64_t c_relax_out = 0;   // relaxation products put on the stack
    int relax_depth = 0;       // largest vacancy-stack depth of the cascades

    Shower(const Problem &p, std::vector<double> &t, std::vector<int> &tl)
        : pb(p), T(p.T), tmp(t), touched(tl) {}

    double rnd() { return rng.uniform(); }

    void score(int ir, double edep, double wt) {
        // EGS_DoseScoring::processEvent: iarg <= 4, ir >= 0, edep > 0
        if (evlog && ir >= 0) evlog->push_back({0, ir, (int)stk.size(), edep > 0 ? edep * wt : 0.0});
        if (ir < 0 || !(edep > 0)) return;
        if (tmp[ir] == 0) touched.push_back(ir);
        tmp[ir] += edep * wt;
    }

    Particle &top() { return stk.back(); }

    // $SELECT-AZIMUTHAL-ANGLE (egsnrc.macros:2491-2508)
    void azimuth(double &cphi, double &sphi) {
        double xphi, xphi2, yphi, yphi2, rhophi2;
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

    // UPHI (egsnrc.mortran:6723-6847): entry 2 samples phi and rotates the
    // top particle (lvl 1); entry 3 reuses phi and the saved (A,B,C) to set
    // the direction of a new particle copied from the one below (lvl 2).
    void rotate(Particle &p, double costhe, double sinthe) {
        double sinps2 = A * A + B * B;
        if (sinps2 < 1.0e-20) {
            p.u = sinthe * cosphi;
            p.v = sinthe * sinphi;
            p.w = Cc * costhe;
        } else {
            double sinpsi = std::sqrt(sinps2);
            double us = sinthe * cosphi, vs = sinthe * sinphi;
            double sindel = B / sinpsi, cosdel = A / sinpsi;
            p.u = Cc * cosdel * us - sindel * vs + A * costhe;
            p.v = Cc * sindel * us + cosdel * vs + B * costhe;
            p.w = -sinpsi * us + Cc * costhe;
        }
    }
    void uphi21(double costhe, double sinthe) {
        azimuth(cosphi, sinphi);
        Particle &p = top();
        A = p.u; B = p.v; Cc = p.w;
        rotate(p, costhe, sinthe);
    }
    void uphi32(double costhe, double sinthe) {
        // caller has pushed the new particle (properties transferred)
        rotate(top(), costhe, sinthe);
    }
    void push_copy() {
        Particle p = top();
        stk.push_back(p);
        if (stk.size() > 100000) throw std::runtime_error("stack overflow");
    }

    // ------------------------------------------------------------ PHOTON
    // egsnrc.mortran:6439-6643
    void photon() {
        Particle &p0 = top();
        double peig = p0.e, eig = peig;
        int irl = p0.ir;
        int medium = pb.med[irl];
        double rhof = pb.rhor[irl];
        const double pcut = pb.P.pcut;
        if (eig <= pcut) { pcut_discard(medium, peig); return; }
        for (;;) {   // :PNEWENERGY:
            Particle &p = top();
            if (p.wt == 0.0) { user_discard_photon(); return; }
            double gle = std::log(eig);
            double rnno35 = rnd();
            if (rnno35 == 0.0) rnno35 = 1.e-30;
            double dpmfp = -std::log(rnno35);
            double cohfac = 1, gmfp = 0;
            int lgle = 0;
            const Medium *M = nullptr;
            bool interact = false;
            while (!interact) {   // :PNEWMEDIUM:
                M = &T.med[medium];
                lgle = (int)(M->ge1 * gle + M->ge0);
                double gmfpr0 = M->gmfp1[lgle - 1] * gle + M->gmfp0[lgle - 1];
                for (;;) {   // :PTRANS:
                    Particle &q = top();
                    gmfp = gmfpr0 / rhof;
                    if (pb.P.rayleigh) {
                        cohfac = M->cohe1[lgle - 1] * gle + M->cohe0[lgle - 1];
                        gmfp = gmfp * cohfac;
                    }
                    double tstep = gmfp * dpmfp;
                    int irnew = q.ir, idisc = 0;
                    double ustep = tstep;
                    if (ustep > q.dnear || q.wt <= 0) {
                        // egsHowfar (egs_advanced_application.cpp:1351-1396)
                        int inew = pb.geo.howfar(q.ir, q.x, q.y, q.z, q.u, q.v, q.w, ustep);
                        if (inew != q.ir) {
                            irnew = inew;
                            if (inew < 0) idisc = -1;
                        }
                    }
                    double vstep = ustep;
                    if (evlog)   // survival over the step: -ustep/(gmfp rho_rel)
                        evlog->push_back({1, q.ir, (int)stk.size(), -ustep / (gmfp * rhof)});
                    q.x = q.x + q.u * vstep;
                    q.y = q.y + q.v * vstep;
                    q.z = q.z + q.w * vstep;
                    q.dnear = q.dnear - ustep;
                    dpmfp = std::max(0., dpmfp - ustep / gmfp);
                    int medold = medium;
                    if (irnew != q.ir) {   // $photon_region_change
                        q.ir = irnew;
                        irl = irnew;
                        if (irnew >= 0) {
                            medium = pb.med[irnew];
                            rhof = pb.rhor[irnew];
                        } else {
                            medium = -1;
                        }
                    }
                    if (eig <= pcut) { pcut_discard(medium, peig); return; }
                    if (idisc < 0) { user_discard_photon(); return; }
                    if (medium != medold) break;   // EXIT :PTRANS:
                    if (dpmfp <= 1.e-8) { interact = true; break; }
                }
            }
            // ---- interaction (:6545-6612)
            Particle &p1 = top();
            if (evlog)   // interaction density proportional to rho_rel
                evlog->push_back({1, p1.ir, (int)stk.size(), 1.0 / rhof});
            if (pb.P.rayleigh) {
                double rnno37 = rnd();
                if (rnno37 <= (1.0 - cohfac)) {
                    double costhe, sinthe;
                    rayleigh_sample(*M, p1.e, gle, lgle, costhe, sinthe);
                    uphi21(costhe, sinthe);
                    continue;   // :PNEWENERGY:
                }
            }
            double rnno36 = rnd();
            double gbr1 = M->gbr11[lgle - 1] * gle + M->gbr10[lgle - 1];
            if (rnno36 <= gbr1 && p1.e > RMT2) {
                pair(*M, medium, gle);
                if (top().iq != 0) return;
            } else {
                double gbr2 = M->gbr21[lgle - 1] * gle + M->gbr20[lgle - 1];
                if (rnno36 < gbr2) {
                    compt(*M);
                    if (top().iq != 0) return;
                } else {
                    photo(*M, gle);
                    if (top().iq != 0) return;
                }
            }
            peig = top().e;
            eig = peig;
            if (eig < pcut) { pcut_discard(medium, peig); return; }
        }
    }

    void pcut_discard(int medium, double peig) {
        // :PCUT-DISCARD: edep = peig, ausgab(1 or 2)
        Particle &p = top();
        score(p.ir, peig, p.wt);
        stk.pop_back();
    }
    void user_discard_photon() {
        // :USER-PHOTON-DISCARD: ausgab(3) -- outside the grid: not scored
        Particle &p = top();
        score(p.ir, p.e, p.wt);
        stk.pop_back();
    }

    // egs_rayleigh_sampling (egsnrc.mortran:8559-8590)
    void rayleigh_sample(const Medium &M, double e, double gle, int lgle,
                         double &costhe, double &sinthe) {
        const double hc_i = 80.65506856998, twice_hc2 = 0.000307444456;
        const double dwi = 100 - 1;
        double pmax = M.pmax1[lgle - 1] * gle + M.pmax0[lgle - 1];
        double xmax = hc_i * e, xv, csqthe, rnnray1;
        do {
            rnnray1 = rnd();
            do {
                double rnnray0 = rnd();
                rnnray0 = rnnray0 * pmax;
                int ibin = (int)(1 + rnnray0 * dwi);
                int ib = M.ray_i[ibin - 1];
                if (M.ray_i[ibin] > ib) {
                    while (!(rnnray0 < M.ray_fcum[ib])) ib++;   // fcum(ib+1), 1-based
                }
                rnnray0 = (rnnray0 - M.ray_fcum[ib - 1]) * M.ray_c[ib - 1];
                xv = M.ray_x[ib - 1] * std::exp(std::log(1 + rnnray0) * M.ray_b[ib - 1]);
            } while (!(xv < xmax));
            xv = xv / e;
            costhe = 1 - twice_hc2 * xv * xv;
            csqthe = costhe * costhe;
        } while (!(2 * rnnray1 < 1 + csqthe));
        sinthe = std::sqrt(1.0 - csqthe);
    }

    // ------------------------------------------------------------- COMPT
    // egsnrc.mortran:528-806 (ibcmp 0: Klein-Nishina; 3: bound, no
    // rejection of the interaction; 1: rejection)
    void compt(const Medium &M) {
        Particle &p = top();
        double peig = p.e;
        double ko = peig / RM;
        double broi = 1 + 2 * ko;
        int ibcmpl = pb.P.ibcmp;
        bool first_time = true;
        double Uj = 0, Jo = 0;
        int j = 0;
        double broi2 = 0, alph1 = 0, bro = 0, alph2 = 0, alpha = 0, bro1 = 0, rejmax = 0;
        double br, temp, sinthe, costhe, pz, af = 0, Fmax = 0, fpz = 0, qc = 0, qc2 = 0, br2 = 0;
    resample_compton:
        if (ibcmpl > 0) {
            double rnno17 = rnd();
            rnno17 = 1 + rnno17 * M.n_shell;
            int i = (int)rnno17;
            if (rnno17 > M.eno[i - 1]) i = M.eno_atbin[i - 1];
            j = M.shell[i - 1];
            Uj = T.cs_be[j];
            if (ko <= Uj) {
                if (ibcmpl == 1) return;
                goto resample_compton;
            }
            Jo = T.cs_Jo[j];
        }
    resample:
        if (ko > 2) {
            if (first_time) {
                broi2 = broi * broi;
                alph1 = std::log(broi);
                bro = 1 / broi;
                alph2 = ko * (broi + 1) * bro * bro;
                alpha = alph1 + alph2;
            }
            double aux, rejf3, rnno19;
            do {
                double rnno15 = rnd(), rnno16 = rnd();
                if (rnno15 * alpha < alph1) br = std::exp(alph1 * rnno16) * bro;
                else br = std::sqrt(rnno16 * broi2 + (1 - rnno16)) * bro;
                temp = (1 - br) / (ko * br);
                sinthe = std::max(0., temp * (2 - temp));
                aux = 1 + br * br;
                rejf3 = aux - br * sinthe;
                rnno19 = rnd();
            } while (!(rnno19 * aux <= rejf3));
        } else {
            if (first_time) {
                bro = 1. / broi;
                bro1 = 1 - bro;
                rejmax = broi + bro;
            }
            double rejf3, rnno16;
            do {
                double rnno15 = rnd();
                rnno16 = rnd();
                br = bro + bro1 * rnno15;
                temp = (1 - br) / (ko * br);
                sinthe = std::max(0., temp * (2 - temp));
                rejf3 = 1 + br * br - br * sinthe;
            } while (!(rnno16 * br * rejmax <= rejf3));
        }
        first_time = false;
        if (br < bro || br > 1) goto resample;
        costhe = 1 - temp;
        if (pb.P.ibcmp == 0) { Uj = 0; goto finished; }
        {
            br2 = br * br;
            double aux = ko * (ko - Uj) * temp;
            double aux1 = 2 * aux + Uj * Uj;
            double pzmax = aux - Uj;
            if (pzmax < 0 && pzmax * pzmax >= aux1) {
                if (ibcmpl == 1) return;
                goto resample_compton;
            }
            pzmax = pzmax / std::sqrt(aux1);
            qc2 = 1 + br * br - 2 * br * costhe;
            qc = std::sqrt(qc2);
            if (pzmax > 1) {
                pzmax = 1; af = 0; Fmax = 1; fpz = 1;
                goto retry_pz;
            }
            double aux3 = 1 + 2 * Jo * std::fabs(pzmax);
            double aux4 = 0.5 * (1 - aux3 * aux3);
            fpz = 0.5 * std::exp(aux4);
            af = qc * (1 + br * (br - costhe) / qc2);
            if (af < 0) {
                if (pzmax > 0) fpz = 1 - fpz;
                double eta_incoh = rnd();
                if (eta_incoh > fpz) {
                    if (ibcmpl == 1) return;
                    goto resample_compton;
                }
                af = 0; Fmax = 1;
                goto retry_pz;
            }
            double fpz1;
            if (pzmax < -0.15) {
                Fmax = 1 - af * 0.15;
                fpz1 = fpz * Fmax * Jo;
            } else if (pzmax < 0.15) {
                Fmax = 1 + af * pzmax;
                aux3 = 1 / (1 + 0.33267252734 * aux3);
                aux4 = fpz * aux3 * (0.3480242 + aux3 * (-0.0958798 + aux3 * 0.7478556)) + T.cs_erfJo[j];
                if (pzmax > 0) {
                    fpz1 = (1 - Fmax * fpz) * Jo - 0.62665706866 * af * aux4;
                    fpz = 1 - fpz;
                } else {
                    fpz1 = Fmax * fpz * Jo - 0.62665706866 * af * aux4;
                }
            } else {
                Fmax = 1 + af * 0.15;
                fpz1 = (1 - Fmax * fpz) * Jo;
                fpz = 1 - fpz;
            }
            double eta_incoh = rnd();
            if (eta_incoh * Jo > fpz1) {
                if (ibcmpl == 1) return;
                goto resample_compton;
            }
        }
    retry_pz:
        if (ibcmpl != 2) {
            double rnno18 = rnd();
            rnno18 = rnno18 * fpz;
            if (rnno18 < 0.5) {
                rnno18 = std::max(1e-30, 2 * rnno18);
                pz = 0.5 * (1 - std::sqrt(1 - 2 * std::log(rnno18))) / Jo;
            } else {
                rnno18 = 2 * (1 - rnno18);
                pz = 0.5 * (std::sqrt(1 - 2 * std::log(rnno18)) - 1) / Jo;
            }
            if (std::fabs(pz) > 1) goto retry_pz;
            if (pz < 0.15) {
                double frej = pz < -0.15 ? 1 - af * 0.15 : 1 + af * pz;
                double eta = rnd();
                if (eta * Fmax > frej) goto retry_pz;
            }
        } else {
            pz = 0; Uj = 0;
        }
        {
            double pz2 = pz * pz;
            if (std::fabs(pz) < 0.01) {
                br = br * (1 + pz * (qc + (br2 - costhe) * pz));
            } else {
                double aux = 1 - pz2 * br * costhe, aux1 = 1 - pz2 * br2;
                double aux2 = qc2 - br2 * pz2 * sinthe;
                if (aux2 > 1e-10) br = br / aux1 * (aux + pz * std::sqrt(aux2));
            }
            Uj = Uj * PRM;
        }
    finished:
        {
            double pesg = br * peig;
            double pese = peig - pesg - Uj + PRM;
            sinthe = std::sqrt(sinthe);
            uphi21(costhe, sinthe);
            top().e = pesg;
            double aux = 1 + br * br - 2 * br * costhe;
            if (aux > 1e-8) {
                costhe = (1 - br * costhe) / std::sqrt(aux);
                sinthe = (1 - costhe) * (1 + costhe);
                if (sinthe > 0) sinthe = -std::sqrt(sinthe);
                else sinthe = 0;
            } else {
                costhe = 0;
                sinthe = -1;
            }
            push_copy();
            uphi32(costhe, sinthe);
            top().e = pese;
            top().iq = -1;
            if (ibcmpl == 1 || ibcmpl == 3) {
                // shell vacancy (:778-794): above 1 keV filled by RELAX
                // (egs_eadl_relax), its sub-threshold products absorbed on
                // the spot; below, Uj is absorbed (ausgab iarg = 4)
                double edep = Uj > 1e-3 ? eadl_relax(T.cs_iz[j], T.cs_shn[j]) : Uj;
                if (edep > 0) score(top().ir, edep, top().wt);
            }
        }
    }

    // ------------------------------------------------------------- PHOTO
    // egsnrc.mortran:5050-5196 with eadl_relax (EDGSET data)
    void photo(const Medium &M, double gle) {
        Particle &p = top();
        double peig = p.e;
        if (peig < T.ee[1 * 16 + 1]) {   // edge_energies(2,1)
            p.iq = -1;
            p.e = peig + PRM;
            return;
        }
        int iZ = 0, vac_shell = 0;
        bool do_relax = false;
        double e_vac = 0;
        if (pb.P.iedgfl) {
            if (M.nne == 1) {
                iZ = (int)(M.z[0] + 0.5);
            } else {
                double aux = peig * peig, aux1 = aux * peig;
                aux = aux * std::sqrt(peig);
                double sigtot = 0, probs[32];
                for (int k = 0; k < M.nne; k++) {
                    int iz = (int)(M.z[k] + 0.5);
                    double sigma;
                    if (peig > T.ee[iz * 16 + 0]) {
                        sigma = (T.ea[iz * 16] + T.eb[iz * 16] / peig + T.ec[iz * 16] / aux +
                                 T.ed[iz * 16] / aux1) / peig;
                    } else {
                        int j;
                        for (j = 2; j <= T.enumb[iz]; j++)
                            if (peig >= T.ee[iz * 16 + j - 1]) break;
                        if (j > T.enumb[iz]) j = T.enumb[iz];
                        int q = iz * 16 + j - 1;
                        sigma = T.ea[q] + gle * (T.eb[q] + gle * (T.ec[q] + gle * T.ed[q]));
                        sigma = std::exp(sigma);
                    }
                    sigma = sigma * M.pz[k];
                    sigtot += sigma;
                    probs[k] = sigma;
                }
                double br = rnd();
                br = br * sigtot;
                int k;
                for (k = 0; k < M.nne; k++) {
                    br -= probs[k];
                    if (br <= 0) break;
                }
                if (k >= M.nne) k = M.nne - 1;
                iZ = (int)(M.z[k] + 0.5);
            }
            if (peig <= T.bind[iZ * 30 + 5]) {   // binding_energies($MXSHELL,iZ)
                p.iq = -1;
                p.e = peig + PRM;
            } else {
                double br = rnd();
                int k;
                for (k = 1; k <= 5; k++) {
                    if (peig > T.bind[iZ * 30 + k - 1]) {
                        double ip = T.iprob[iZ * 6 + k - 1];
                        if (br < ip) break;
                        br = (br - ip) / (1 - ip);
                    }
                }
                if (k > 4) {
                    p.iq = -1;
                    p.e = peig + PRM;
                } else {
                    e_vac = T.bind[iZ * 30 + k - 1];
                    p.e = peig - e_vac + PRM;
                    do_relax = true;
                    vac_shell = k;
                    p.iq = -1;
                }
            }
        } else {
            p.e = peig + PRM;
            p.iq = -1;
        }
        if (pb.P.iphter) photoelectron_direction();
        if (do_relax) {   // RELAX (:5191-5193): the cascade; what stays is absorbed here
            double edep = eadl_relax(iZ, vac_shell);
            if (edep > 0) score(top().ir, edep, top().wt);
        }
    }

    // egs_eadl_relax (egsnrc.mortran:6020-6216): fill a vacancy in shell
    // shell_egs of element iZ by fluorescent photons and Auger and Coster-
    // Kronig electrons (EADL transition data as alias histograms).  Products
    // above PCUT / ECUT go on the stack, with the position, region and
    // weight of the particle on top when the cascade starts and isotropic
    // directions; the energy of the others and of vacancies below 1 keV is
    // returned for local absorption.  Vacancies are filled last in, first
    // out.  A transition drawn from the bin that completes probabilities
    // summing to less than 1 gives its whole energy to an electron.
    double eadl_relax(int iZ, int shell_egs) {
        if (iZ < 1 || iZ > 100 || T.relax_zfirst[iZ] < 0 || shell_egs < 1 ||
            shell_egs > T.relax_nshell[iZ])
            return 0;   // unknown vacancy: nothing is done (:6108)
        const int shell = T.relax_zfirst[iZ] + shell_egs - 1;
        const double Ec = pb.P.ecut - RM, Pc = pb.P.pcut, min_E = 1e-3;
        double Evac = T.shell_be[shell];
        if (shell_egs > 4) return Evac;   // below L3: absorbed (:6126-6139)
        const Particle save = top();
        double edep = 0;
        int vacs[128];
        int nvac = 0, vac = shell;
        for (;;) {
            if (Evac < min_E || T.relax_ntran[vac] < 1) {
                edep += Evac;
            } else {
                const int first = T.relax_first[vac], n = T.relax_ntran[vac];
                // sample_alias_histogram (egs_utilities.mortran:1994-2019)
                double r1 = rnd(), r2 = rnd();
                int ibin = 1 + (int)(n * r1);
                if (r2 > T.relax_prob[first + ibin - 1]) ibin = T.relax_atbin[first + ibin - 1];
                const int st = T.relax_state[first + ibin - 1];
                const int base = vac - T.shell_num[vac];   // shell 0 of the element
                double Ef, Ecc;
                int iqf;
                if (st < 0) {
                    Ef = 0; iqf = -1; Ecc = Ec;
                } else if (st <= 64) {                        // fluorescence
                    iqf = 0;
                    const int nw = base + st;
                    Ef = T.shell_be[nw];
                    vacs[nvac++] = nw;
                    Ecc = Pc;
                } else {                                      // Auger, Coster-Kronig
                    iqf = -1;
                    const int n1 = base + st / 64, n2 = base + st % 64;
                    Ef = T.shell_be[n1] + T.shell_be[n2];
                    vacs[nvac++] = n1;
                    vacs[nvac++] = n2;
                    Ecc = Ec;
                }
                if (nvac > 126) throw std::runtime_error("relaxation: too many vacancies");
                relax_depth = std::max(relax_depth, nvac);
                const double Ex = Evac - Ef;
                if (Ex > Ecc) {
                    Particle q = save;
                    q.iq = iqf;
                    const double cost = 2 * rnd() - 1;
                    double sint = 1 - cost * cost;
                    if (sint > 0) {
                        sint = std::sqrt(sint);
                        double cphi, sphi;
                        azimuth(cphi, sphi);
                        q.u = sint * cphi; q.v = sint * sphi; q.w = cost;
                    } else {
                        q.u = 0; q.v = 0; q.w = cost;
                    }
                    q.e = iqf == 0 ? Ex : Ex + RM;
                    stk.push_back(q);
                    c_relax_out++;
                } else {
                    edep += Ex;
                }
            }
            if (nvac == 0) break;
            vac = vacs[--nvac];
            Evac = T.shell_be[vac];
        }
        return edep;
    }

    // $SELECT-PHOTOELECTRON-DIRECTION (egsnrc.macros:1916-1950)
    void photoelectron_direction() {
        Particle &p = top();
        double eelec = p.e;
        if (!(eelec > pb.P.ecut)) return;
        double beta = std::sqrt((eelec - RM) * (eelec + RM)) / eelec;
        double gamma = eelec / RM;
        double alpha = 0.5 * gamma - 0.5 + 1. / gamma;
        double ratio = beta / alpha;
        double costhe, sinth2, xi, rnpht2;
        do {
            double rnpht = rnd();
            rnpht = 2. * rnpht - 1.;
            if (ratio <= 0.2) {
                double fkappa = rnpht + 0.5 * ratio * (1. - rnpht) * (1. + rnpht);
                if (gamma < 100) {
                    costhe = (beta + fkappa) / (1. + beta * fkappa);
                } else {
                    if (fkappa > 0)
                        costhe = 1 - (1 - fkappa) * (gamma - 3) / (2 * (1 + fkappa) * std::pow(gamma - 1, 3));
                    else
                        costhe = (beta + fkappa) / (1. + beta * fkappa);
                }
                xi = (1 + beta * fkappa) * gamma * gamma;
            } else {
                xi = gamma * gamma * (1. + alpha * (std::sqrt(1. + ratio * (2. * rnpht + ratio)) - 1.));
                costhe = (1. - 1. / xi) / beta;
            }
            sinth2 = std::max(0., (1. - costhe) * (1. + costhe));
            rnpht2 = rnd();
        } while (rnpht2 > 0.5 * (1. + gamma) * sinth2 * xi / gamma);
        double sinthe = std::sqrt(sinth2);
        uphi21(costhe, sinthe);
    }

    // -------------------------------------------------------------- PAIR
    // egsnrc.mortran:4518-4731 (Bethe-Heitler, no NRC tables, no triplet)
    // one $SET-PAIR-ANGLE branch: 1 leading term, 2 KM (the 3D-2003
    // distribution by rejection in xi = 1/(1 + (theta E/m)^2)), 3 isotropic,
    // 4 cos = 1 - 2 sqrt(r)
    void pair_angle(int use, const Medium &M, double eig, double ese, double &costhe,
                    double &sinthe) {
        if (use == 1) {
            double pse = std::sqrt(std::max(0.0, (ese - RM) * (ese + RM)));
            costhe = 1.0 - 2.0 * rnd();
            sinthe = RM * std::sqrt((1.0 - costhe) * (1.0 + costhe)) / (pse * costhe + ese);
            costhe = (ese * costhe + pse) / (pse * costhe + ese);
        } else if (use == 2) {
            double ztarg = M.zbrang, tteig = eig / RM, ttese = ese / RM;
            double esedei = ttese / (tteig - ttese), eseder = 1.0 / esedei;
            auto rej = [&](double xi) {   // $SET-PAIR-REJECTION-FUNCTION
                double a = (1.0 + eseder) * (1.0 + esedei) / (2.0 * tteig);
                return 2.0 + 3.0 * (esedei + eseder) -
                       4.0 * (esedei + eseder + 1.0 - 4.0 * (xi - 0.5) * (xi - 0.5)) *
                           (1.0 + 0.25 * std::log(a * a + ztarg * xi * xi));
            };
            double ximin = 1.0 / (1.0 + std::pow(3.141593 * ttese, 2));
            double rejmin = rej(ximin);
            double ya = std::pow(2.0 / tteig, 2);
            double xitry = std::max(0.01, std::max(ximin, std::min(0.5, std::sqrt(ya / ztarg))));
            double galpha = 1.0 + 0.25 * std::log(ya + ztarg * xitry * xitry);
            double gbeta = 0.5 * ztarg * xitry / (ya + ztarg * xitry * xitry);
            galpha = galpha - gbeta * (xitry - 0.5);
            double ximid = galpha / (3.0 * gbeta);
            if (galpha >= 0.0) ximid = 0.5 - ximid + std::sqrt(ximid * ximid + 0.25);
            else ximid = 0.5 - ximid - std::sqrt(ximid * ximid + 0.25);
            ximid = std::max(0.01, std::max(ximin, std::min(0.5, ximid)));
            double rejmid = rej(ximid);
            double rejtop = 1.02 * std::max(rejmin, rejmid);
            double theta;
            for (;;) {
                double xitst = rnd();
                double rejtst = rej(xitst);
                double rtest = rnd();
                theta = std::sqrt(1.0 / xitst - 1.0) / ttese;
                if (rtest <= rejtst / rejtop && theta < 3.14159265358979) break;
            }
            sinthe = std::sin(theta);
            costhe = std::cos(theta);
        } else if (use == 3) {
            costhe = 1.0 - 2.0 * rnd();
            sinthe = (1 - costhe) * (1 + costhe);
            sinthe = sinthe > 0 ? std::sqrt(sinthe) : 0.0;
        } else {
            costhe = 1.0 - 2.0 * std::sqrt(rnd());
            sinthe = (1 - costhe) * (1 + costhe);
            sinthe = sinthe > 0 ? std::sqrt(sinthe) : 0.0;
        }
    }

    void pair(const Medium &M, int medium, double gle) {
        Particle &p = top();
        double peig = p.e, eig = peig;
        double pese1, pese2;
        int iq1, iq2;
        const double k = eig / RM;
        if (pb.P.pair_nrc && k < T.nrcp_emax) {
            // NRC energy sharing (egsnrc.mortran:4596-4620)
            int ibin = 1;
            if (k > T.nrcp_emin) {
                double abin = 1 + std::log((k - 2) / (T.nrcp_emin - 2)) * T.nrcp_dlei;
                ibin = (int)abin;
                abin = abin - ibin;
                double rbin = rnd();
                if (rbin < abin) ibin = ibin + 1;
            }
            if (ibin < 1 || ibin > 84) throw std::runtime_error("NRC pair energy bin out of range");
            double xx = alias_sample1(T.nrcp_x.data(), &M.nrcp_f[(size_t)(ibin - 1) * 65],
                                      &M.nrcp_w[(size_t)(ibin - 1) * 64],
                                      &M.nrcp_i[(size_t)(ibin - 1) * 64], 64);
            if (xx > 0.5) {
                pese1 = PRM * (1 + xx * (k - 2)); iq1 = 1;
                pese2 = peig - pese1; iq2 = -1;
            } else {
                pese2 = PRM * (1 + xx * (k - 2)); iq2 = 1;
                pese1 = peig - pese2; iq1 = -1;
            }
        } else if (eig <= 2.1) {
            double rnno30 = rnd(), rnno34 = rnd();
            pese2 = PRM + 0.5 * rnno30 * (peig - 2 * PRM);
            pese1 = peig - pese2;
            if (rnno34 < 0.5) { iq1 = -1; iq2 = 1; } else { iq1 = 1; iq2 = -1; }
        } else {
            int l, l1;
            double Amax, Bmax, aux1;
            const double(*dl)[8] = M.dl;
            if (eig < 50.) {
                l = 5; l1 = l + 1;
                double delta = 4 * M.delcm / eig;
                if (delta < 1) {
                    Amax = dl[0][l - 1] + delta * (dl[1][l - 1] + delta * dl[2][l - 1]);
                    Bmax = dl[0][l1 - 1] + delta * (dl[1][l1 - 1] + delta * dl[2][l1 - 1]);
                } else {
                    double aux2 = std::log(delta + dl[5][l - 1]);
                    Amax = dl[3][l - 1] + dl[4][l - 1] * aux2;
                    Bmax = dl[3][l1 - 1] + dl[4][l1 - 1] * aux2;
                }
                aux1 = 1 - RMT2 / eig;
                aux1 = aux1 * aux1;
                aux1 = aux1 * Amax / 3;
                aux1 = aux1 / (Bmax + aux1);
            } else {
                l = 7;
                Amax = dl[0][l - 1];
                Bmax = dl[0][l];
                aux1 = M.bpar[1] * (1 - M.bpar[0] * RM / eig);
            }
            double del0 = eig * M.delcm;
            double Eavail = eig - RMT2;
            double Eminus, rejf, rejmax, rnno34;
            do {
                double rnno30 = rnd(), rnno31 = rnd();
                rnno34 = rnd();
                double br;
                if (rnno30 > aux1) {
                    br = 0.5 * rnno31;
                    rejmax = Bmax;
                    l1 = l + 1;
                } else {
                    double rnno32 = rnd(), rnno33 = rnd();
                    br = 0.5 * (1 - std::max(rnno31, std::max(rnno32, rnno33)));
                    rejmax = Amax;
                    l1 = l;
                }
                Eminus = br * Eavail + RM;
                double Eplus = eig - Eminus;
                double delta = del0 / (Eminus * Eplus);
                if (delta < 1)
                    rejf = dl[0][l1 - 1] + delta * (dl[1][l1 - 1] + delta * dl[2][l1 - 1]);
                else
                    rejf = dl[3][l1 - 1] + dl[4][l1 - 1] * std::log(delta + dl[5][l1 - 1]);
            } while (!(rnno34 * rejmax <= rejf));
            pese2 = Eminus;
            pese1 = peig - pese2;
            rnno34 = rnd();
            if (rnno34 < 0.5) { iq1 = -1; iq2 = 1; } else { iq1 = 1; iq2 = -1; }
        }
        // $SET-PAIR-ANGLE (egsnrc.macros:2070-2163)
        top().e = pese1;
        double ese2 = pese2;
        if (pb.P.iprdst > 0) {
            int use = pb.P.iprdst;
            if (use == 4) {
                double rtest = rnd();
                use = rtest < pese1 / (pese1 + 10) ? 1 : 4;
            } else if (use == 2 && eig < 4.14) {
                use = 1;
            }
            for (int ichrg = 1; ichrg <= 2; ichrg++) {
                double ese = ichrg == 1 ? pese1 : ese2;
                if (ichrg == 2 && pb.P.iprdst == 4) {
                    double rtest = rnd();
                    use = rtest < ese / (ese + 10) ? 1 : 4;
                }
                double costhe, sinthe;
                pair_angle(use, M, eig, ese, costhe, sinthe);
                if (ichrg == 1) {
                    uphi21(costhe, sinthe);
                } else {
                    sinthe = -sinthe;
                    push_copy();
                    top().e = ese2;
                    uphi32(costhe, sinthe);
                }
            }
            top().iq = iq2;
            stk[stk.size() - 2].iq = iq1;
        } else {
            // theta = 0: both along the photon direction
            push_copy();
            top().e = ese2;
            top().iq = iq2;
            stk[stk.size() - 2].iq = iq1;
        }
    }

    // ------------------------------------------------------------ ELECTR
    // egsnrc.mortran:985-1823 with PRESTA-II or PRESTA-I steps, exact or
    // PRESTA-I boundary crossing (skin depth in elastic MFPs), spin effects
    // off, EII off, and the egs++ C-interface callbacks.
    int lel(const Medium &M, double elke) const {
        int L = (int)(M.eke1 * elke + M.eke0);
        return L;
    }
    static double ev(const std::vector<double> &a0, const std::vector<double> &a1, int L, double x) {
        return a1[L - 1] * x + a0[L - 1];
    }
    // $COMPUTE-DRANGE (egsnrc.macros:3214-3243)
    double compute_drange(const Medium &M, int lelec, double e1, double e2, int lel1,
                          double el1, double el2) const {
        double fedep = 1 - e2 / e1;
        double elktmp = 0.5 * (el1 + el2 + 0.25 * fedep * fedep * (1 + fedep * (1 + 0.875 * fedep)));
        int lelktmp = lel1;
        double dedxmid, aux;
        if (lelec < 0) {
            dedxmid = ev(M.ededx0, M.ededx1, lelktmp, elktmp);
            dedxmid = 1 / dedxmid;
            aux = M.ededx1[lelktmp - 1] * dedxmid;
        } else {
            dedxmid = ev(M.pdedx0, M.pdedx1, lelktmp, elktmp);
            dedxmid = 1 / dedxmid;
            aux = M.pdedx1[lelktmp - 1] * dedxmid;
        }
        aux = aux * (1 + 2 * aux) * std::pow(fedep / (2 - fedep), 2) / 6;
        return fedep * e1 * dedxmid * (1 + aux);
    }
    // $COMPUTE-ELOSS (egsnrc.macros:3290-3313)
    double compute_eloss(const Medium &M, int lelec, double rhof, double step, double e,
                         double el, int L) const {
        double dedxmid, aux;
        if (lelec < 0) {
            dedxmid = ev(M.ededx0, M.ededx1, L, el);
            aux = M.ededx1[L - 1] / dedxmid;
        } else {
            dedxmid = ev(M.pdedx0, M.pdedx1, L, el);
            aux = M.pdedx1[L - 1] / dedxmid;
        }
        double de = dedxmid * step * rhof;
        double fedep = de / e;
        de = de * (1 - 0.5 * fedep * aux * (1 - 0.333333 * fedep * (aux - 1 - 0.25 * fedep * (2 - aux * (4 - aux)))));
        return de;
    }
    // $COMPUTE-ELOSS-G (egsnrc.macros:3315-3352)
    double compute_eloss_g(const Medium &M, int lelec, int qel, double rhof, double range,
                           double step, double eke, double elke, int lelke) const {
        const std::vector<double> &rep = qel ? M.range_ep1 : M.range_ep0;
        double tuss = range - rep[lelke - 1] / rhof;
        if (tuss >= step) return compute_eloss(M, lelec, rhof, step, eke, elke, lelke);
        int lelktmp = lelke;
        tuss = (range - step) * rhof;
        if (tuss <= 0) return eke - M.te * 0.99;
        while (tuss < rep[lelktmp - 1]) lelktmp--;
        double elktmp = (lelktmp + 1 - M.eke0) / M.eke1;
        double eketmp = M.e_array[lelktmp];
        tuss = (rep[lelktmp] - tuss) / rhof;
        double de = compute_eloss(M, lelec, rhof, tuss, eketmp, elktmp, lelktmp);
        return de + eke - eketmp;
    }

    void electron() {
        const double ecut = pb.P.ecut;
        if (pb.P.kerma) {
            Particle &p = top();
            ecut_discard(p.iq, p.e);
            return;
        }
        Particle *pp = &top();
        int irl = pp->ir;
        int medium = pb.med[irl];
        double rhof = pb.rhor[irl];
    newelectron:
        pp = &top();
        int lelec = pp->iq;
        int qel = (1 + lelec) / 2;
        double peie = pp->e, eie = peie;
        if (eie <= ecut) { ecut_discard(lelec, peie); return; }
        if (pp->wt == 0.0) { user_electron_discard(lelec, 3); return; }
        for (;;) {   // :TSTEP:
            bool compute_tstep = true;
            double eke = eie - RM;
            const Medium *M = &T.med[medium];
            double rnne1 = rnd();
            if (rnne1 == 0.0) rnne1 = 1.e-30;
            double demfp = std::max(-std::log(rnne1), 1.e-8);
            double elke = std::log(eke);
            int lelke = lel(*M, elke);
            double sig0, sigf, dedx0;
            auto eval_sigf = [&]() {
                if (lelec < 0) {
                    sigf = ev(M->esig0, M->esig1, lelke, elke);
                    dedx0 = ev(M->ededx0, M->ededx1, lelke, elke);
                } else {
                    sigf = ev(M->psig0, M->psig1, lelke, elke);
                    dedx0 = ev(M->pdedx0, M->pdedx1, lelke, elke);
                }
                sigf = sigf / dedx0;
            };
            // $EVALUATE-SIG0
            if (M->sig_mono[qel]) { eval_sigf(); sig0 = sigf; }
            else sig0 = lelec < 0 ? M->esig_e : M->psig_e;
            double total_de = 0, total_tstep = 0;
            bool restart_tstep = false;
            do {   // :USTEP:
                Particle &p = top();
                double sig = sig0, tstep;
                if (sig <= 0) {
                    tstep = 1e8;   // vacdst
                    sig0 = 1.E-15;
                } else {
                    // $CALCULATE-TSTEP-FROM-DEMFP
                    if (compute_tstep) {
                        total_de = demfp / sig;
                        double fedep = total_de;
                        double ekef = eke - fedep;
                        if (ekef <= M->e_array[0]) {
                            tstep = 1e8;
                        } else {
                            double elkef = std::log(ekef);
                            int lelkef = lel(*M, elkef);
                            if (lelkef == lelke) {
                                tstep = compute_drange(*M, lelec, eke, ekef, lelke, elke, elkef);
                            } else {
                                double ekei = M->e_array[lelke - 1];
                                double elkei = (lelke - M->eke0) / M->eke1;
                                double tuss = compute_drange(*M, lelec, eke, ekei, lelke, elke, elkei);
                                ekei = M->e_array[lelkef];
                                elkei = (lelkef + 1 - M->eke0) / M->eke1;
                                tstep = compute_drange(*M, lelec, ekei, ekef, lelkef, elkei, elkef);
                                const std::vector<double> &rep = qel ? M->range_ep1 : M->range_ep0;
                                tstep = tstep + tuss + rep[lelke - 1] - rep[lelkef];
                            }
                        }
                        total_tstep = tstep;
                        compute_tstep = false;
                    }
                    tstep = total_tstep / rhof;
                }
                if (lelec < 0) dedx0 = ev(M->ededx0, M->ededx1, lelke, elke);
                else dedx0 = ev(M->pdedx0, M->pdedx1, lelke, elke);
                double dedx = rhof * dedx0;
                double tmxs = ev(M->tmxs0, M->tmxs1, lelke, elke);
                tmxs = tmxs / rhof;
                // $COMPUTE-RANGE
                double range;
                {
                    double ekei = M->e_array[lelke - 1];
                    double elkei = (lelke - M->eke0) / M->eke1;
                    range = compute_drange(*M, lelec, eke, ekei, lelke, elke, elkei);
                    const std::vector<double> &rep = qel ? M->range_ep1 : M->range_ep0;
                    range = (range + rep[lelke - 1]) / rhof;
                }
                tmxs = std::min(tmxs, 1e10);   // smaxir
                double tustep = std::min(tstep, std::min(tmxs, range));
                double tperp = pb.geo.hownear(p.ir, p.x, p.y, p.z);
                p.dnear = tperp;
                // $SET-SKINDEPTH / $CALCULATE-ELASTIC-SCATTERING-MFP
                double blccl = rhof * M->blcc;
                double p2 = eke * (eke + RMT2), beta2 = p2 / (p2 + RMSQ);
                double ssmfp = beta2 / blccl;
                double skindepth = pb.P.skindepth * ssmfp;
                tustep = std::min(tustep, std::max(tperp, skindepth));
                bool callhowfar, dosingle = false, domultiple = false, callmsdist;
                double ustep, de;
                double uscat = 0, vscat = 0, wscat = 0, xtrans = 0, ytrans = 0, ztrans = 0;
                // the approximate step's screening angle, beta^2 and xi
                // ($SET-USTEP / $SET-TVSTEP)
                double ms_chia2 = 0, ms_beta2 = 0, ms_xi = 0;
                const bool exact = pb.P.exact_bca != 0;
                if (tustep <= tperp && (!exact || tustep > skindepth)) {
                    callhowfar = false;
                    dosingle = false;
                    callmsdist = true;
                    de = compute_eloss_g(*M, lelec, qel, rhof, range, tustep, eke, elke, lelke);
                    if (pb.P.presta1)
                        msdist_pI(*M, eke, de, tustep, rhof, p.u, p.v, p.w, p.x, p.y, p.z,
                                  uscat, vscat, wscat, xtrans, ytrans, ztrans, ustep);
                    else
                        msdist_pII(*M, eke, de, tustep, rhof, p.u, p.v, p.w, p.x, p.y, p.z,
                                   uscat, vscat, wscat, xtrans, ytrans, ztrans, ustep);
                } else if (!exact) {
                    // boundary crossing a la EGS4/PRESTA-I with the exact
                    // path-length correction (egsnrc.mortran:1338-1346)
                    callmsdist = false;
                    domultiple = true;
                    ustep = set_ustep(blccl, rhof * M->xcc, eke, tustep, dedx, ms_chia2,
                                      ms_beta2, ms_xi);
                    callhowfar = !(ustep < tperp);
                } else {
                    callmsdist = false;
                    double rnnoss = rnd();
                    if (rnnoss < 1.e-30) rnnoss = 1.e-30;
                    double lambda = -std::log(1 - rnnoss);
                    double lambda_max = 0.5 * blccl * RM / dedx * std::pow(eke / RM + 1, 3);
                    if (lambda >= 0 && lambda_max > 0) {
                        double tuss = lambda < lambda_max ? lambda * ssmfp * (1 - 0.5 * lambda / lambda_max)
                                                          : 0.5 * lambda * ssmfp;
                        if (tuss < tustep) { tustep = tuss; dosingle = true; }
                        else dosingle = false;
                    } else {
                        stk.pop_back();
                        return;
                    }
                    ustep = tustep;
                    callhowfar = !(ustep < tperp);
                }
                int irold = p.ir, irnew = p.ir, idisc = 0;
                const double ustep0 = ustep;
                if (callhowfar || p.wt <= 0) {
                    int inew = pb.geo.howfar(p.ir, p.x, p.y, p.z, p.u, p.v, p.w, ustep);
                    if (inew != p.ir) {
                        irnew = inew;
                        if (inew < 0) idisc = -1;
                    }
                }
                if (ustep <= 0) ustep = 0;   // $CHECK-NEGATIVE-USTEP
                if (ustep == 0) {
                    if (irnew != irold) {
                        p.ir = irnew; irl = irnew;
                        if (irnew >= 0) { medium = pb.med[irnew]; rhof = pb.rhor[irnew]; }
                    }
                    if (eie <= ecut) { ecut_discard(lelec, peie); return; }
                    restart_tstep = true;
                    break;
                }
                double vstep = ustep, tvstep;
                if (callhowfar) {
                    if (exact) {
                        tvstep = vstep;
                        if (tvstep != tustep) dosingle = false;
                    } else {
                        tvstep = set_tvstep(blccl, rhof * M->xcc, eke, tustep, dedx, vstep,
                                            ustep0, ms_chia2, ms_beta2, ms_xi);
                    }
                    de = compute_eloss_g(*M, lelec, qel, rhof, range, tvstep, eke, elke, lelke);
                } else {
                    tvstep = tustep;
                    if (!callmsdist)
                        de = compute_eloss_g(*M, lelec, qel, rhof, range, tvstep, eke, elke, lelke);
                }
                double save_de = de;
                double edep = de;
                double ekef = eke - de;
                double costhe = 1, sinthe = 0;
                if (!callmsdist && domultiple) {
                    // multiple scattering over the approximate step
                    // (egsnrc.mortran:1484-1502)
                    double lambda = blccl * tvstep / ms_beta2 / (1 + ms_chia2);
                    bool fi = true;
                    mscat(lambda, ms_chia2, ms_xi, fi, costhe, sinthe);
                } else if (!callmsdist && dosingle) {
                    double ekems = std::max(ekef, ecut - RM);
                    double p2s = ekems * (ekems + RMT2);
                    double beta2s = p2s / (p2s + RMSQ);
                    double chia2 = M->xcc / (4 * M->blcc * p2s);
                    (void)beta2s;
                    sscat(chia2, costhe, sinthe);
                }
                double xf, yf, zf, uf, vf, wf;
                if (callmsdist) {
                    uf = uscat; vf = vscat; wf = wscat;
                    xf = xtrans; yf = ytrans; zf = ztrans;
                } else {
                    xf = p.x + p.u * vstep;
                    yf = p.y + p.v * vstep;
                    zf = p.z + p.w * vstep;
                    if (domultiple || dosingle) {
                        double ut = p.u, vt = p.v, wt = p.w;
                        uphi21(costhe, sinthe);
                        uf = p.u; vf = p.v; wf = p.w;
                        p.u = ut; p.v = vt; p.w = wt;
                    } else {
                        uf = p.u; vf = p.v; wf = p.w;
                    }
                }
                score(p.ir, edep, p.wt);   // ausgab(TRANAUSB): energy lost along the step
                c_esteps++;
                p.x = xf; p.y = yf; p.z = zf;
                p.u = uf; p.v = vf; p.w = wf;
                p.dnear = p.dnear - vstep;
                irold = p.ir;
                peie = peie - edep;
                eie = peie;
                p.e = peie;
                if (irnew == irl && eie <= ecut) { ecut_discard(lelec, peie); return; }
                int medold = medium;
                eke = eie - RM;
                elke = std::log(eke);
                lelke = lel(*M, elke);
                if (irnew != irold) {   // $electron_region_change
                    p.ir = irnew; irl = irnew;
                    if (irnew >= 0) { medium = pb.med[irnew]; rhof = pb.rhor[irnew]; }
                }
                if (eie <= ecut) { ecut_discard(lelec, peie); return; }
                if (idisc < 0) { user_electron_discard(lelec, 1); return; }
                if (medium != medold) { restart_tstep = true; break; }
                // $UPDATE-DEMFP
                demfp = demfp - save_de * sig;
                total_de = total_de - save_de;
                total_tstep = total_tstep - tvstep * rhof;
                if (total_tstep < 1e-9) demfp = 0;
            } while (!(demfp < 1.e-8));
            if (restart_tstep) continue;
            eval_sigf();
            double sigratio = sigf / sig0;
            double rfict = rnd();
            if (rfict <= sigratio) {
                // ---- discrete interaction
                Particle &p = top();
                if (lelec < 0) {
                    double ebr1 = ev(M->ebr10, M->ebr11, lelke, elke);
                    double rnno24 = rnd();
                    if (rnno24 <= ebr1) {
                        brems(*M, elke);
                        if (top().iq == 0) return;
                        goto newelectron;
                    }
                    if (p.e <= M->thmoll) {
                        if (ebr1 <= 0) goto newelectron;
                        brems(*M, elke);
                        if (top().iq == 0) return;
                        goto newelectron;
                    }
                    moller(*M, elke);
                    if (top().iq == 0) return;
                    goto newelectron;
                }
                double pbr1 = ev(M->pbr10, M->pbr11, lelke, elke);
                double rnno25 = rnd();
                if (rnno25 < pbr1) {
                    brems(*M, elke);
                    if (top().iq == 0) return;
                    goto newelectron;
                }
                double pbr2 = ev(M->pbr20, M->pbr21, lelke, elke);
                if (rnno25 < pbr2) {
                    bhabha(*M);
                    if (top().iq == 0) return;
                    goto newelectron;
                }
                annih();
                return;
            }
        }
    }

    // :USER-ELECTRON-DISCARD: (egsnrc.mortran:1802-1812)
    void user_electron_discard(int lelec, int idisc) {
        Particle &p = top();
        double edep = (lelec < 0 || idisc == 99) ? p.e - PRM : p.e + PRM;
        score(p.ir, edep, p.wt);   // outside the grid (ir < 0) nothing is scored
        stk.pop_back();
    }

    // msdist_pII (egsnrc.mortran:4050-4308), spin effects off
    void msdist_pII(const Medium &M, double e0, double eloss, double tustep, double rhof,
                    double u0, double v0, double w0, double x0, double y0, double z0,
                    double &us, double &vs, double &ws, double &xf, double &yf,
                    double &zf, double &ustep) {
        double blccc = M.blcc, xcccc = M.xcc;
        double e = e0 - 0.5 * eloss;
        double tau = e / PRM, tau2 = tau * tau;
        double epsilon = eloss / e0, epsilonp = eloss / e;
        e = e * (1 - epsilonp * epsilonp * (6 + 10 * tau + 5 * tau2) / (24 * tau2 + 72 * tau + 48));
        double p2 = e * (e + RMT2);
        double beta2 = p2 / (p2 + RMSQ);
        double chia2 = xcccc / (4 * p2 * blccc);
        double lambda = 0.5 * tustep * rhof * blccc / beta2;
        double temp2 = 0.166666 * (4 + tau * (6 + tau * (7 + tau * (4 + tau)))) *
                       std::pow(epsilonp / ((tau + 1) * (tau + 2)), 2);
        lambda = lambda * (1 - temp2);
        double etap = 1, xi_corr = 1, gamma = 1, ms_corr = 1;
        chia2 = chia2 * etap;
        lambda = lambda / (etap * (1 + chia2)) * ms_corr;
        double chilog = std::log(1 + 1 / chia2);
        double q1 = 2 * chia2 * (chilog * (1 + chia2) - 1);
        gamma = 6 * chia2 * (1 + chia2) * (chilog * (1 + 2 * chia2) - 2) / q1 * gamma;
        double xi = q1 * lambda;
        bool find_index = true;
        double w1, sint1, w2, sint2, cphi1, sphi1, cphi2, sphi2;
        mscat(lambda, chia2, xi, find_index, w1, sint1);
        azimuth(cphi1, sphi1);
        mscat(lambda, chia2, xi, find_index, w2, sint2);
        azimuth(cphi2, sphi2);
        double u2 = sint2 * cphi2, v2 = sint2 * sphi2;
        double u2p = w1 * u2 + sint1 * w2;
        us = u2p * cphi1 - v2 * sphi1;
        vs = u2p * sphi1 + v2 * cphi1;
        ws = w1 * w2 - sint1 * u2;
        xi = 2 * xi * xi_corr;
        double eta = rnd();
        eta = std::sqrt(eta);
        double eta1 = 0.5 * (1 - eta);
        double delta = 0.9082483 - (0.1020621 - 0.0263747 * gamma) * xi;
        double temp1 = 2 + tau;
        double temp = (2 + tau * temp1) / ((tau + 1) * temp1);
        temp = temp - (tau + 1) / ((tau + 2) * (chilog * (1 + chia2) - 1));
        temp = temp * epsilonp;
        temp1 = 1 - temp;
        delta = delta + 0.40824829 * (epsilon * (tau + 1) / ((tau + 2) * (chilog * (1 + chia2) - 1) *
                                                             (chilog * (1 + 2 * chia2) - 2)) -
                                      0.25 * temp * temp);
        double b = eta * delta, c = eta * (1 - delta);
        double w1v2 = w1 * v2;
        double ut = b * sint1 * cphi1 + c * (cphi1 * u2 - sphi1 * w1v2) + eta1 * us * temp1;
        double vt = b * sint1 * sphi1 + c * (sphi1 * u2 + cphi1 * w1v2) + eta1 * vs * temp1;
        double wt = eta1 * (1 + temp) + b * w1 + c * w2 + eta1 * ws * temp1;
        ustep = tustep * std::sqrt(ut * ut + vt * vt + wt * wt);
        double sint02 = u0 * u0 + v0 * v0;
        if (sint02 > 1e-20) {
            double sint0 = std::sqrt(sint02), sint0i = 1 / sint0;
            double cphi0 = sint0i * u0, sphi0 = sint0i * v0;
            u2p = w0 * us + sint0 * ws;
            ws = w0 * ws - sint0 * us;
            us = u2p * cphi0 - vs * sphi0;
            vs = u2p * sphi0 + vs * cphi0;
            u2p = w0 * ut + sint0 * wt;
            wt = w0 * wt - sint0 * ut;
            ut = u2p * cphi0 - vt * sphi0;
            vt = u2p * sphi0 + vt * cphi0;
        } else {
            wt = w0 * wt;
            ws = w0 * ws;
        }
        xf = x0 + tustep * ut;
        yf = y0 + tustep * vt;
        zf = z0 + tustep * wt;
    }

    // msdist_pI (egsnrc.mortran:4310-4515), spin effects off: one mscat
    // angle over the whole step, the number of elastic mean free paths
    // corrected for energy loss, the exact Lewis <z> along the initial
    // direction and a lateral offset of half the sine of the angle.
    void msdist_pI(const Medium &M, double e0, double eloss, double tustep, double rhof,
                   double u0, double v0, double w0, double x0, double y0, double z0,
                   double &us, double &vs, double &ws, double &xf, double &yf,
                   double &zf, double &ustep) {
        double blccc = M.blcc, xcccc = M.xcc;
        double e = e0 - 0.5 * eloss;
        double p2 = e * (e + RMT2);
        double p2i = 1 / p2;
        double beta2 = p2 / (p2 + RMSQ);
        double lambda = tustep * rhof * blccc / beta2;
        double factor = 1 / (1 + 0.9784671 * e);   // 1 / (1 + e / 2m)
        double epsilon = eloss / e0;
        epsilon = epsilon / (1 - 0.5 * epsilon);
        double temp = 0.25 * (1 - factor * (1 - 0.333333 * factor)) * epsilon * epsilon;
        lambda = lambda * (1 + temp);
        double chia2 = xcccc * p2i / (4 * blccc);
        lambda = lambda / (1 + chia2);
        double chilog = std::log(1 + 1 / chia2);
        double q1 = 2 * chia2 * (chilog * (1 + chia2) - 1);
        double xi = q1 * lambda;
        bool find_index = true;
        double sint, cphi, sphi;
        mscat(lambda, chia2, xi, find_index, ws, sint);
        azimuth(cphi, sphi);
        us = sint * cphi;
        vs = sint * sphi;
        double z = xi < 0.1 ? 1 - xi * (0.5 - xi * (0.166666667 - 0.041666667 * xi))
                            : (1 - std::exp(-xi)) / xi;
        double r = 0.5 * sint, r2 = r * r, z2 = z * z;
        double r2max = 1 - z2;
        if (r2max < r2) { r2 = r2max; r = std::sqrt(r2); }
        double ut = r * cphi, vt = r * sphi, wt = z;
        ustep = std::sqrt(z2 + r2) * tustep;
        double sint02 = u0 * u0 + v0 * v0;
        if (sint02 > 1e-20) {
            double sint0 = std::sqrt(sint02), sint0i = 1 / sint0;
            double cphi0 = sint0i * u0, sphi0 = sint0i * v0;
            double u2p = w0 * us + sint0 * ws;
            ws = w0 * ws - sint0 * us;
            us = u2p * cphi0 - vs * sphi0;
            vs = u2p * sphi0 + vs * cphi0;
            u2p = w0 * ut + sint0 * wt;
            wt = w0 * wt - sint0 * ut;
            ut = u2p * cphi0 - vt * sphi0;
            vt = u2p * sphi0 + vt * cphi0;
        } else {
            wt = w0 * wt;
            ws = w0 * ws;
        }
        xf = x0 + tustep * ut;
        yf = y0 + tustep * vt;
        zf = z0 + tustep * wt;
    }

    // $CALCULATE-XI (egsnrc.macros:1324-1348), spin effects off: the first
    // GS moment xi of a path `len` at kinetic energy ekems, with the
    // screening angle chia2 and beta^2 left for the multiple-scattering
    // call of an approximate step.  blccl, xccl are rhof * blcc, xcc.
    static double calc_xi(double blccl, double xccl, double ekems, double len,
                          double &chia2, double &beta2) {
        double p2 = ekems * (ekems + RMT2);
        beta2 = p2 / (p2 + RMSQ);
        chia2 = xccl / (4 * blccl * p2);
        double xi = 0.5 * xccl / p2 / beta2 * len;
        return xi * (std::log(1 + 1. / chia2) - 1 / (1 + chia2));
    }

    // $SET-USTEP (egsnrc.macros:1308-1322): the average straight-line
    // distance along the initial direction of a curved path tustep.
    static double set_ustep(double blccl, double xccl, double eke, double tustep, double dedx,
                            double &chia2, double &beta2, double &xi) {
        double ekems = eke - 0.5 * tustep * dedx;
        xi = calc_xi(blccl, xccl, ekems, tustep, chia2, beta2);
        if (xi < 0.1) return tustep * (1 - xi * (0.5 - xi * 0.166667));
        return tustep * (1 - std::exp(-xi)) / xi;
    }

    // $SET-TVSTEP (egsnrc.macros:1350-1384): the curved path of a step that
    // howfar cut from ustep0 to vstep.
    static double set_tvstep(double blccl, double xccl, double eke, double tustep, double dedx,
                             double vstep, double ustep0, double &chia2, double &beta2,
                             double &xi) {
        if (!(vstep < ustep0)) return tustep;
        double ekems = eke - 0.5 * tustep * vstep / ustep0 * dedx;
        xi = calc_xi(blccl, xccl, ekems, vstep, chia2, beta2);
        if (xi < 0.1) return vstep * (1 + xi * (0.5 + xi * 0.333333));
        if (xi < 0.999999) return -vstep * std::log(1 - xi) / xi;
        throw std::runtime_error("SET-TVSTEP: xi > 1");
    }

    // mscat (egsnrc.mortran:2984-3160), spin effects off.  ms_i/ms_j/
    // ms_omega2 persist between the two calls of one msdist_pII step.
    int ms_i = 0, ms_j = 0;
    double ms_omega2 = 0;
    void mscat(double lambda, double chia2, double q1, bool &find_index, double &cost,
               double &sint) {
        if (lambda <= 13.8) {
            double sprob = rnd();
            double explambda = std::exp(-lambda);
            if (sprob < explambda) { cost = 1; sint = 0; return; }
            double wsum = (1 + lambda) * explambda;
            if (sprob < wsum) {
                double xi = rnd();
                xi = 2 * chia2 * xi / (1 - xi + chia2);
                cost = 1 - xi;
                sint = std::sqrt(xi * (2 - xi));
                return;
            }
            if (lambda <= 1) {
                double wprob = explambda;
                wsum = explambda;
                cost = 1; sint = 0;
                int icount = 0;
                do {
                    icount++;
                    if (icount > 20) break;
                    wprob = wprob * lambda / icount;
                    wsum = wsum + wprob;
                    double xi = rnd();
                    xi = 2 * chia2 * xi / (1 - xi + chia2);
                    double cosz = 1 - xi;
                    double sinz = xi * (2 - xi);
                    if (sinz > 1.e-20) {
                        sinz = std::sqrt(sinz);
                        xi = rnd();
                        double phi = xi * 6.2831853;
                        cost = cost * cosz - sint * sinz * std::cos(phi);
                        sint = std::sqrt(std::max(0.0, (1 - cost) * (1 + cost)));
                    }
                } while (!(wsum > sprob));
                return;
            }
        }
        if (lambda <= 1e5) {
            const std::vector<double> &mc = T.ms_const;   // llammin..dqmsi (real*4)
            double dllambi = mc[3], dqmsi = mc[5];
            if (find_index) {
                double llmbda = std::log(lambda);
                double ai = llmbda * dllambi;
                int i = (int)ai;
                ai = ai - i;
                double xi = rnd();
                if (xi < ai) i = i + 1;
                int j;
                if (q1 < 1e-3) j = 0;
                else if (q1 < 0.5) {
                    double aj = q1 * dqmsi;
                    j = (int)aj;
                    aj = aj - j;
                    xi = rnd();
                    if (xi < aj) j = j + 1;
                } else j = 7;
                if (llmbda < 2.2299)
                    ms_omega2 = chia2 * (lambda + 4) *
                                (1.347006 + llmbda * (0.209364 - llmbda * (0.45525 - llmbda * (0.50142 - 0.081234 * llmbda))));
                else
                    ms_omega2 = chia2 * (lambda + 4) *
                                (-2.77164 + llmbda * (2.94874 - llmbda * (0.1535754 - llmbda * 0.00552888)));
                ms_i = i; ms_j = j;
                find_index = false;
            }
            double xi = rnd();
            double ak = xi * 31;
            int k = (int)ak;
            ak = ak - k;
            int base = (ms_i * 8 + ms_j) * 32;
            if (ak > T.ms_w[base + k]) k = T.ms_i[base + k];
            double a = T.ms_f[base + k], u = T.ms_u[base + k];
            double du = T.ms_u[base + k + 1] - u;
            xi = rnd();
            if (std::fabs(a) < 0.2) {
                double x1 = 0.5 * (1 - xi) * a;
                u = u + xi * du * (1 + x1 * (1 - xi * a));
            } else {
                u = u - du / a * (1 - std::sqrt(1 + xi * a * (2 + a)));
            }
            xi = ms_omega2 * u / (1 + 0.5 * ms_omega2 - u);
            if (xi > 1.99999) xi = 1.99999;
            cost = 1 - xi;
            sint = std::sqrt(xi * (2 - xi));
            return;
        }
        throw std::runtime_error("maximum step size in mscat exceeded");
    }

    // sscat (egsnrc.mortran:3226-3262), spin effects off
    void sscat(double chia2, double &cost, double &sint) {
        double xi = rnd();
        xi = 2 * chia2 * xi / (1 - xi + chia2);
        cost = 1 - xi;
        sint = std::sqrt(xi * (2 - xi));
    }

    // MOLLER (egsnrc.mortran:2599-2746), EII off
    // eii_sample (egsnrc.mortran:7760-7830)
    void eii_sample(const Medium &M, int ish, int iZ, double Uj) {
        Particle &p = top();
        double peie = p.e, Tk = peie - RM, tau = Tk / RM, tau1 = tau + 1;
        double tau12 = tau1 * tau1, tau2 = tau * tau, p2 = tau2 + 2 * tau;
        double beta2 = p2 / tau12;
        double Wmax = 0.5 * (Tk + Uj), xmax = Uj / Wmax;
        double c1 = std::pow(Wmax / peie, 2), c2 = (2 * tau + 1) / tau12;
        double fm_s = std::log(RMT2 * p2 / Uj) - beta2 - 0.5;
        double prob_s = 0.66666667 * fm_s * (1 + xmax + xmax * xmax);
        double fm_h = 2 + c1 - c2;
        if (fm_h < 1) fm_h = 1;
        double prob = fm_h + prob_s;
        double wx, frej, r3;
        do {
            double r1 = rnd(), r2 = rnd();
            r3 = rnd();
            if (r1 * prob < fm_h) {
                wx = 1 / (r2 * xmax + 1 - r2);
                double wxx = wx * xmax, aux = wxx / (2 - wxx);
                frej = (1 + aux * (aux - c2) + c1 * wxx * wxx) / fm_h;
            } else {
                wx = 1 / std::pow(r2 * xmax * xmax * xmax + 1 - r2, 0.333333333);
                frej = 1 - std::log(wx) / fm_s;
            }
        } while (!(r3 < frej));
        wx = wx * Uj;
        double h1 = (peie + PRM) / Tk, pese1 = peie - wx;
        p.e = pese1;
        double dcosth = h1 * (pese1 - PRM) / (pese1 + PRM);
        uphi21(std::sqrt(dcosth), std::sqrt(1 - dcosth));
        double pese2 = wx - Uj + PRM;
        double edep = 0;
        if (pese2 > M.ae) {
            push_copy();
            top().e = pese2;
            top().iq = -1;
            dcosth = h1 * (pese2 - PRM) / (pese2 + PRM);
            uphi32(std::sqrt(dcosth), -std::sqrt(1 - dcosth));
        } else {
            edep = wx - Uj;
        }
        edep += eadl_relax(iZ, ish);
        if (edep > 0) score(top().ir, edep, top().wt);
    }

    void moller(const Medium &M, double elke) {
        Particle &p = top();
        double peie = p.e, pekin = peie - PRM, ekin = pekin;
        if (pb.P.eii && M.eii_nsh > 0) {   // egsnrc.mortran:2634-2664
            int L = lel(M, elke);
            double sigm = ev(M.esig0, M.esig1, L, elke) * (1 - ev(M.ebr10, M.ebr11, L, elke));
            double rsh = rnd();
            rsh = sigm * rsh;
            for (int el = 0; el < M.nne; el++) {
                int iZ = (int)(M.z[el] + 0.5);
                int nsh = M.eii_no[el];
                if (nsh <= 0) continue;
                int ifirst = M.eii_first[el];
                for (int ish = 1; ish <= nsh; ish++) {
                    double Uj = T.bind[(size_t)iZ * 30 + ish - 1];
                    if (!(ekin > Uj && (Uj > M.te || Uj > M.ap))) continue;
                    int jj = ifirst + ish - 1;
                    int i = (int)(T.eii_a[jj - 1] * elke + T.eii_b[jj - 1]) + (jj - 1) * 250;
                    if (i < 1 || i > (int)T.eii_xa.size())
                        throw std::runtime_error("EII cross-section index out of range");
                    double sig_j = T.eii_xa[i - 1] * elke + T.eii_xb[i - 1];
                    sig_j = sig_j * M.pz[el] * M.eii_cons;
                    rsh = rsh - sig_j;
                    if (rsh < 0) {
                        eii_sample(M, ish, iZ, Uj);
                        return;
                    }
                }
            }
        }
        if (ekin <= 2 * M.te) return;
        double t0 = ekin / RM, e0 = t0 + 1.0;
        double extrae = peie - M.thmoll;
        double e02 = e0 * e0;
        double g2 = t0 * t0 / e02, g3 = (2. * t0 + 1.) / e02;
        double gmax = (1. + 1.25 * g2);
        double br, rejf4, rnno28;
        do {
            double rnno27 = rnd();
            br = M.te / (ekin - extrae * rnno27);
            double r = br / (1. - br);
            rnno28 = rnd();
            rejf4 = (1. + g2 * br * br + r * (r - g3));
            rnno28 = gmax * rnno28;
        } while (!(rnno28 <= rejf4));
        double pekse2 = br * ekin;
        double pese1 = peie - pekse2, pese2 = pekse2 + PRM;
        p.e = pese1;
        double h1 = (peie + PRM) / pekin;
        double dcosth = h1 * (pese1 - PRM) / (pese1 + PRM);
        double sinthe = std::sqrt(1.0 - dcosth), costhe = std::sqrt(dcosth);
        uphi21(costhe, sinthe);
        push_copy();
        top().iq = -1;
        top().e = pese2;
        dcosth = h1 * (pese2 - PRM) / (pese2 + PRM);
        sinthe = -std::sqrt(1.0 - dcosth);
        costhe = std::sqrt(dcosth);
        uphi32(costhe, sinthe);
    }

    // BHABHA (egsnrc.mortran:229-307)
    void bhabha(const Medium &M) {
        Particle &p = top();
        double peip = p.e, pekin = peip - PRM, ekin = pekin;
        double t0 = ekin / RM, e0 = t0 + 1.;
        double yy = 1. / (t0 + 2.), e02 = e0 * e0;
        double beta2 = (e02 - 1.) / e02;
        double ep0 = M.te / ekin, ep0c = 1. - ep0;
        double y2 = yy * yy, yp = 1. - 2. * yy, yp2 = yp * yp;
        double b4 = yp2 * yp, b3 = b4 + yp2, b2 = yp * (3. + y2), b1 = 2. - y2;
        double br, rejf2, rnno04;
        do {
            double rnno03 = rnd();
            br = ep0 / (1. - ep0c * rnno03);
            rnno04 = rnd();
            rejf2 = (1.0 - beta2 * br * (b1 - br * (b2 - br * (b3 - br * b4))));
        } while (!(rnno04 <= rejf2));
        int iq_np = 1, iq_new = -1;
        if (!(br < 0.5)) { iq_np = -1; iq_new = 1; br = 1. - br; }
        br = std::max(br, 0.0);
        double pekse2 = br * ekin;
        double pese1 = peip - pekse2, pese2 = pekse2 + PRM;
        p.e = pese1;
        p.iq = iq_np;
        double h1 = (peip + PRM) / pekin;
        double dcosth = std::min(1.0, h1 * (pese1 - PRM) / (pese1 + PRM));
        double sinthe = std::sqrt(1.0 - dcosth), costhe = std::sqrt(dcosth);
        uphi21(costhe, sinthe);
        push_copy();
        top().e = pese2;
        top().iq = iq_new;
        dcosth = h1 * (pese2 - PRM) / (pese2 + PRM);
        sinthe = -std::sqrt(1.0 - dcosth);
        costhe = std::sqrt(dcosth);
        uphi32(costhe, sinthe);
    }

    // BREMS (egsnrc.mortran:312-526): Bethe-Heitler (ibr_nist = 0), no
    // splitting, photon angle from the Koch-Motz 2BS (ibrdst = 1) or the
    // leading term (ibrdst = 0)
    // alias_sample1 (egs_utilities.mortran:1901-1931) on the NIST table of
    // grid energy j (1-based)
    double alias_sample1(const Medium &M, int j) {
        if (j < 1 || j > 100) throw std::runtime_error("NIST brems grid index out of range");
        return alias_sample1(&M.nb_x[(size_t)(j - 1) * 51], &M.nb_f[(size_t)(j - 1) * 51],
                             &M.nb_w[(size_t)(j - 1) * 50], &M.nb_i[(size_t)(j - 1) * 50], 50);
    }

    // alias_sample1 over an n-bin table (xs, fs at n+1 points)
    double alias_sample1(const double *xs, const double *fs, const double *ws,
                         const int32_t *ib, int n) {
        double r1 = rnd(), r2 = rnd();
        double aj = 1 + r1 * n;
        int k = (int)aj;
        aj = aj - k;
        if (aj > ws[k - 1]) k = ib[k - 1];
        double x = xs[k - 1], dx = xs[k] - x;
        if (fs[k - 1] > 0) {
            double a = fs[k] / fs[k - 1] - 1;
            if (std::fabs(a) < 0.2) {
                double rnno1 = 0.5 * (1 - r2) * a;
                return x + r2 * dx * (1 + rnno1 * (1 - r2 * a));
            }
            return x - dx / a * (1 - std::sqrt(1 + r2 * a * (2 + a)));
        }
        return x + dx * std::sqrt(r2);
    }

    void brems(const Medium &M, double elke) {
        Particle &p = top();
        double peie = p.e, eie = peie;
        int l = eie < 50.0 ? 1 : 3, l1 = l + 1;
        double ekin = peie - PRM;
        double brmin = M.ap / ekin;
        double waux = elke - M.log_ap;
        double a = p.u, b = p.v, c = p.w;
        double sinpsi = a * a + b * b, sindel = 0, cosdel = 0;
        if (sinpsi > 1e-20) {
            sinpsi = std::sqrt(sinpsi);
            sindel = b / sinpsi;
            cosdel = a / sinpsi;
        }
        double ztarg = M.zbrang, tteie = eie / RM;
        double beta = std::sqrt((tteie - 1) * (tteie + 1)) / tteie;
        double y2max = 2 * beta * (1 + beta) * tteie * tteie, y2maxi = 1 / y2max;
        double z2max = y2max + 1, z2maxi = std::sqrt(z2max);
        double esg, pesg, pese, ese;
        const double(*dl)[8] = M.dl;
        const bool nist = pb.P.ibr_nist >= 1;   // NIST or NRC tables
        int jj = 0;
        double ajj = 0;
        if (nist) {   // egsnrc.mortran:404-408
            ajj = 1 + (waux + M.log_ap - M.nb_lemin) * M.nb_dlei;
            jj = (int)ajj;
            ajj = ajj - jj;
            if (jj > 100) { jj = 100; ajj = -1; }
        }
        for (;;) {
            if (nist) {   // egsnrc.mortran:412-422
                double br;
                if (ekin > M.nb_emin) {
                    double r1 = rnd();
                    int j = r1 < ajj ? jj + 1 : jj;
                    br = alias_sample1(M, j);
                } else {
                    br = rnd();
                }
                esg = M.ap * std::exp(br * waux);
                pesg = esg;
                pese = peie - pesg;
                ese = pese;
                break;
            }
            double rnno06 = rnd(), rnno07 = rnd();
            double br = brmin * std::exp(rnno06 * waux);
            esg = ekin * br;
            pesg = esg;
            pese = peie - pesg;
            ese = pese;
            double delta = esg / eie / ese * M.delcm, aux = ese / eie, phi1, phi2;
            if (delta < 1) {
                phi1 = dl[0][l - 1] + delta * (dl[1][l - 1] + delta * dl[2][l - 1]);
                phi2 = dl[0][l1 - 1] + delta * (dl[1][l1 - 1] + delta * dl[2][l1 - 1]);
            } else {
                phi1 = dl[3][l - 1] + dl[4][l - 1] * std::log(delta + dl[5][l - 1]);
                phi2 = phi1;
            }
            double rejf = (1 + aux * aux) * phi1 - 2 * aux * phi2 / 3;
            if (rnno07 < rejf) break;
        }
        double weight = p.wt;
        push_copy();
        Particle &g = top();
        g.e = pesg;
        g.iq = 0;
        g.wt = weight;
        double y2tst;
        if (pb.P.ibrdst == 1) {
            double ttese = ese / RM, esedei = ttese / tteie;
            double rjarg1 = 1 + esedei * esedei, rjarg2 = rjarg1 + 2 * esedei;
            double aux = 2 * ese * tteie / esg;
            aux = aux * aux;
            double aux1 = aux * ztarg, rjarg3;
            if (aux1 > 10) rjarg3 = M.lzbrang + (1 - aux1) / (aux1 * aux1);
            else rjarg3 = std::log(aux / (1 + aux1));
            double rejmax = rjarg1 * rjarg3 - rjarg2;
            for (;;) {
                y2tst = rnd();
                double rtest = rnd();
                double aux3 = z2maxi / (y2tst + (1 - y2tst) * z2maxi);
                rtest = rtest * aux3 * rejmax;
                y2tst = aux3 * aux3 - 1;
                double y2tst1 = esedei * y2tst / (aux3 * aux3 * aux3 * aux3);
                double aux4 = 16 * y2tst1 - rjarg2, aux5 = rjarg1 - 4 * y2tst1;
                if (rtest < aux4 + aux5 * rjarg3) break;
                double aux2 = std::log(aux / (1 + aux1 / (aux3 * aux3 * aux3 * aux3)));
                double rejtst = aux4 + aux5 * aux2;
                if (rtest < rejtst) break;
            }
        } else {
            y2tst = rnd();
            y2tst = y2tst / (1 - y2tst + y2maxi);
        }
        double costhe = 1 - 2 * y2tst * y2maxi;
        double sinthe = std::sqrt(std::max((1 - costhe) * (1 + costhe), 0.0));
        double cphi, sphi;
        azimuth(cphi, sphi);
        Particle &gg = top();
        if (sinpsi >= 1e-10) {
            double us = sinthe * cphi, vs = sinthe * sphi;
            gg.u = c * cosdel * us - sindel * vs + a * costhe;
            gg.v = c * sindel * us + cosdel * vs + b * costhe;
            gg.w = c * costhe - sinpsi * us;
        } else {
            gg.u = sinthe * cphi;
            gg.v = sinthe * sphi;
            gg.w = c * costhe;
        }
        stk[stk.size() - 2].e = pese;
    }

    // ANNIH (egsnrc.mortran:70-184), two-photon in-flight annihilation
    void annih() {
        Particle &p = top();
        double pavip = p.e + PRM, avip = pavip;
        double a = avip / RM, g = a - 1.0, t = g - 1.0;
        double pp = std::sqrt(a * t), pot = pp / t;
        double ep0 = 1.0 / (a + pp);
        double wsamp = std::log((1.0 - ep0) / ep0);
        double aa = p.u, bb = p.v, cc = p.w;
        double sinpsi = aa * aa + bb * bb, sindel = 0, cosdel = 0;
        if (sinpsi > 1e-20) {
            sinpsi = std::sqrt(sinpsi);
            sindel = bb / sinpsi;
            cosdel = aa / sinpsi;
        }
        double ep, rejf, rnno02;
        do {
            double rnno01 = rnd();
            ep = ep0 * std::exp(rnno01 * wsamp);
            rnno02 = rnd();
            rejf = 1 - (ep * a - 1) * (ep * a - 1) / (ep * (a * a - 2));
        } while (!(rnno02 <= rejf));
        double esg1 = avip * ep;
        p.e = esg1;
        p.iq = 0;
        double costhe = std::max(-1.0, std::min(1.0, (esg1 - RM) * pot / esg1));
        double sinthe = std::sqrt(1.0 - costhe * costhe);
        double cphi, sphi;
        azimuth(cphi, sphi);
        auto set_dir = [&](Particle &q, double ct, double st) {
            if (sinpsi >= 1e-10) {
                double us = st * cphi, vs = st * sphi;
                q.u = cc * cosdel * us - sindel * vs + aa * ct;
                q.v = cc * sindel * us + cosdel * vs + bb * ct;
                q.w = cc * ct - sinpsi * us;
            } else {
                q.u = st * cphi;
                q.v = st * sphi;
                q.w = cc * ct;
            }
        };
        set_dir(p, costhe, sinthe);
        double pesg2 = pavip - esg1, esg2 = pesg2;
        push_copy();
        Particle &q = top();
        q.e = pesg2;
        q.iq = 0;
        costhe = std::max(-1.0, std::min(1.0, (esg2 - RM) * pot / esg2));
        sinthe = -std::sqrt(1.0 - costhe * costhe);
        set_dir(q, costhe, sinthe);
    }

    // :ECUT-DISCARD: / :POSITRON-ANNIHILATION: (egsnrc.mortran:1773-1800)
    // A particle whose last step left the grid is outside: no dose; a
    // positron still annihilates (same random numbers) but its photons
    // start outside and egsHowfar discards them at once
    // (egs_advanced_application.cpp:1355-1358).
    void ecut_discard(int lelec, double peie) {
        Particle &p = top();
        double edep = peie - PRM;   // e-: e - prm; e+: $POSITRON-ECUT-DISCARD
        if (p.ir < 0) {
            const size_t n0 = stk.size();
            if (lelec > 0 && edep < peie) annih_at_rest();
            while (stk.size() >= n0) stk.pop_back();
            return;
        }
        score(p.ir, edep, p.wt);
        if (lelec > 0 && edep < peie) { annih_at_rest(); return; }
        stk.pop_back();
    }

    // ANNIH_AT_REST (egsnrc.mortran:186-224), nbr_split = 1
    void annih_at_rest() {
        double costhe = rnd();
        costhe = 2 * costhe - 1;
        double sinthe = std::sqrt(std::max(0.0, (1 - costhe) * (1 + costhe)));
        double cphi, sphi;
        azimuth(cphi, sphi);
        Particle &p = top();
        p.e = PRM; p.iq = 0;
        p.u = sinthe * cphi; p.v = sinthe * sphi; p.w = costhe;
        push_copy();
        Particle &q = top();
        q.e = PRM; q.iq = 0;
        Particle &pp = stk[stk.size() - 2];
        q.u = -pp.u; q.v = -pp.v; q.w = -pp.w;
    }

    // SHOWER loop (egsnrc.mortran:6646-6717; egs_interface2.c:92-99)
    void shower(const Particle &first) {
        stk.clear();
        stk.push_back(first);
        while (!stk.empty()) {
            if (top().iq == 0) photon();
            else electron();
        }
    }
};

// EGS_AliasTable::make (egs_alias_table.cpp:102-171)
static void alias_make(Source &s, int n, const double *x, const double *f, int type) {
    s.type = type;
    s.np = type == 0 ? n : n - 1;
    int np = s.np;
    s.xi.assign(x, x + n);
    s.fi.assign(f, f + np);
    s.wi.assign(np, 1.0);
    s.bin.assign(np, 0);
    std::vector<double> fcum(np);
    std::vector<bool> not_done(np, true);
    double sum = 0;
    for (int i = 0; i < np; i++) {
        fcum[i] = type == 0 ? s.fi[i] : s.fi[i] * (s.xi[i + 1] - s.xi[i]);
        sum += fcum[i];
    }
    for (int i = 0; i < np; i++) s.fi[i] /= sum;
    sum /= np;
    for (int i = 0; i < np - 1; i++) {
        int jh, jl, high = -1, low = -1;
        for (jh = 0; jh < np; jh++) if (not_done[jh] && fcum[jh] > sum) { high = jh; break; }
        if (high < 0) break;
        for (jl = 0; jl < np; jl++) if (not_done[jl] && fcum[jl] < sum) { low = jl; break; }
        if (low < 0) break;
        double aux = sum - fcum[low];
        fcum[high] -= aux;
        not_done[jl] = false;
        s.wi[low] = fcum[low] / sum;
        s.bin[low] = high;
    }
}

// EGS_AliasTable::sample (egs_alias_table.cpp:268-300), types 0 and 1
static double alias_sample(const Source &s, Shower &sh) {
    double r1 = sh.rnd();
    double aj = r1 * s.np;
    int j = (int)aj;
    aj -= j;
    if (aj > s.wi[j]) j = s.bin[j];
    if (!s.type) return s.xi[j];
    double x = s.xi[j], dx = s.xi[j + 1] - x;
    double r2 = sh.rnd();
    return x + dx * r2;
}

struct Engine {
    Problem pb;
    int nstreams = 1024;
    int ixx = 1802, jxx = 9373;
    std::vector<Ranmar> streams;
    std::vector<double> sum, sum2;
    int64_t nhist = 0;
    int nthreads = 1;

    void seed(int a, int b) {
        ixx = a; jxx = b;
        streams.assign(nstreams, Ranmar());
        for (int s = 0; s < nstreams; s++) {
            int is, js;
            stream_seeds(ixx, jxx, s, is, js);
            streams[s].set_state(is, js);
        }
    }

    // one primary: EGS_Application::simulateSingleShower (egs_application.cpp:919-959)
    void history(Shower &sh) {
        const Source &s = pb.src;
        Particle p{};
        for (int ntry = 0;; ntry++) {
            if (ntry > 100000) throw std::runtime_error("no particle entered the geometry");
            double E = s.spectrum == 0 ? s.energy : alias_sample(s, sh);
            double x = s.x0, y = s.y0, z = s.z;
            if (s.shape == 1) {
                x = s.x0 + (s.x1 - s.x0) * sh.rnd();
                y = s.y0 + (s.y1 - s.y0) * sh.rnd();
            }
            int ireg = pb.geo.is_where(x, y, z);
            if (ireg < 0) {
                double t = 1e30;
                ireg = pb.geo.howfar(-1, x, y, z, s.ux, s.uy, s.uz, t);
                if (ireg >= 0) { x += s.ux * t; y += s.uy * t; z += s.uz * t; }
            }
            if (ireg < 0) continue;
            // EGS_AdvancedApplication::shower (egs_advanced_application.cpp:871-893)
            p.e = s.charge ? E + RM : E;
            p.x = x; p.y = y; p.z = z;
            p.u = s.ux; p.v = s.uy; p.w = s.uz;
            p.wt = 1; p.dnear = 0; p.iq = s.charge; p.ir = ireg;
            break;
        }
        sh.shower(p);
    }

    // Score-function gradient of L = sum_v (D_v - D*_v)^2 with respect to
    // the relative densities, photon kerma mode: forward pass for D, then
    // a replay of the same histories (same streams) with the event log and
    // a reverse sweep per history (downstream deposit weights per depth).
    void grad(int64_t n, const double *target, double *g_out, double *dose, double *loss,
              int mode = 0) {
        std::vector<Ranmar> start = streams;
        std::fill(sum.begin(), sum.end(), 0.0);
        std::fill(sum2.begin(), sum2.end(), 0.0);
        nhist = 0;
        run(n);
        std::vector<double> mass(pb.nvox), w(pb.nvox);
        double L = 0;
        for (int i = 0; i < pb.nvox; i++) {
            int iz = i / pb.geo.nxy, iy = (i / pb.geo.nx) % pb.geo.ny, ix = i % pb.geo.nx;
            double vol = (pb.geo.xp[ix + 1] - pb.geo.xp[ix]) * (pb.geo.yp[iy + 1] - pb.geo.yp[iy]) *
                         (pb.geo.zp[iz + 1] - pb.geo.zp[iz]);
            mass[i] = vol * pb.rhor[i] * pb.T.med[pb.med[i]].rho;
            dose[i] = sum[i] / (double)n / mass[i];
            double r = dose[i] - target[i];
            L += r * r;
            w[i] = mode == 0 ? 2 * r / ((double)n * mass[i]) : 1.0 / (double)n;
        }
        if (mode == 1) {   // linear functional: total energy per history
            L = 0;
            for (int i = 0; i < pb.nvox; i++) L += sum[i] / (double)n;
        }
        *loss = L;
        streams = start;
        int nt = std::max(1, std::min(nthreads, nstreams));
        std::vector<std::vector<double>> gt(nt);
        std::vector<std::thread> th;
        std::vector<std::string> errs(nt);
        for (int t = 0; t < nt; t++) {
            th.emplace_back([&, t]() {
                try {
                    std::vector<double> tmp(pb.nvox, 0.0);
                    std::vector<int> touched;
                    std::vector<Shower::Ev> ev;
                    std::vector<double> acc(256, 0.0);
                    gt[t].assign(pb.nvox, 0.0);
                    Shower sh(pb, tmp, touched);
                    sh.evlog = &ev;
                    for (int s = t; s < nstreams; s += nt) {
                        sh.rng = streams[s];
                        int64_t cnt = n / nstreams + (s < n % nstreams ? 1 : 0);
                        for (int64_t k = 0; k < cnt; k++) {
                            ev.clear();
                            history(sh);
                            for (int ir : touched) tmp[ir] = 0;
                            touched.clear();
                            int dmax = 0;
                            for (auto &e : ev) dmax = std::max(dmax, e.depth);
                            if (dmax >= (int)acc.size()) acc.resize(dmax + 1);
                            std::fill(acc.begin(), acc.begin() + dmax + 1, 0.0);
                            for (size_t q = ev.size(); q-- > 0;) {
                                const auto &e = ev[q];
                                for (int d = e.depth + 1; d <= dmax; d++) acc[d] = 0;
                                if (e.type == 0) {
                                    double wd = w[e.vox] * e.val;
                                    for (int d = 0; d <= e.depth; d++) acc[d] += wd;
                                } else {
                                    gt[t][e.vox] += e.val * acc[e.depth];
                                }
                            }
                        }
                        streams[s] = sh.rng;
                    }
                } catch (const std::exception &ex) {
                    errs[t] = ex.what();
                }
            });
        }
        for (auto &x : th) x.join();
        for (int t = 0; t < nt; t++)
            if (!errs[t].empty()) throw std::runtime_error(errs[t]);
        for (int i = 0; i < pb.nvox; i++) {
            double gsum = 0;
            for (int t = 0; t < nt; t++) gsum += gt[t][i];
            // explicit mass term: D_v = E_v / (rho_v V_v)
            g_out[i] = mode == 0 ? gsum - 2 * (dose[i] - target[i]) * dose[i] / pb.rhor[i] : gsum;
        }
    }

    void run(int64_t n) {
        int nt = std::max(1, std::min(nthreads, nstreams));
        std::vector<std::vector<double>> s1(nt), s2(nt);
        std::vector<std::thread> th;
        std::vector<std::string> errs(nt);
        for (int t = 0; t < nt; t++) {
            th.emplace_back([&, t]() {
                try {
                    std::vector<double> tmp(pb.nvox, 0.0);
                    std::vector<int> touched;
                    s1[t].assign(pb.nvox, 0.0);
                    s2[t].assign(pb.nvox, 0.0);
                    Shower sh(pb, tmp, touched);
                    for (int s = t; s < nstreams; s += nt) {
                        sh.rng = streams[s];
                        // histories h with h % nstreams == s
                        int64_t cnt = n / nstreams + (s < n % nstreams ? 1 : 0);
                        for (int64_t k = 0; k < cnt; k++) {
                            history(sh);
                            for (int ir : touched) {
                                s1[t][ir] += tmp[ir];
                                s2[t][ir] += tmp[ir] * tmp[ir];
                                tmp[ir] = 0;
                            }
                            touched.clear();
                        }
                        streams[s] = sh.rng;
                    }
                } catch (const std::exception &e) {
                    errs[t] = e.what();
                }
            });
        }
        for (auto &x : th) x.join();
        for (int t = 0; t < nt; t++)
            if (!errs[t].empty()) throw std::runtime_error(errs[t]);
        for (int t = 0; t < nt; t++)
            for (int i = 0; i < pb.nvox; i++) { sum[i] += s1[t][i]; sum2[i] += s2[t][i]; }
        nhist += n;
    }
};

}  // namespace oracle

using namespace oracle;

// Same layouts as include/egsgpu.h
struct OGeom {
    int32_t nx, ny, nz;
    const double *xb, *yb, *zb;
    const int32_t *medium;
    const double *rhor;
    double boundary_tolerance;
};
struct OSource {
    int32_t charge, spectrum;
    double energy;
    int32_t nbins;
    const double *spec_e, *spec_w;
    int32_t shape;
    double x0, x1, y0, y1, z, ux, uy, uz;
};
struct OTransport {
    double ecut, pcut;
    int32_t rayleigh, bound_compton, photoelectron_angle, atomic_relaxation,
        pair_angle, brems_angle, spin, exact_bca;
    double estepe, ximax, skin_depth;
    int32_t kerma, fp64, howfarless, n_split, nbr_split, play_rr, ireject;
    double esave;
    int32_t step_algorithm, brems_xsec, pair_xsec;
    char photon_xsec[16], eii_xsec[16];
};

// mscati's default PRESTA-I skin depth (egsnrc.mortran:2790-2817) over the
// media `use` at the lowest ECUT: the blob holds blcc and xcc as mscati
// leaves them (blcc times exp(2 gamma_E - 1), xcc squared), so the PEGS4
// values are recovered first.
static double default_blcmin(const Tables &T, const std::vector<int> &use, double ecutmn,
                             double prm = PRM) {
    double tstbmn = 1e30;
    for (int m : use) {
        const Medium &M = T.med[m];
        double blcc = M.blcc / 1.16699413758864573, xcc = std::sqrt(M.xcc);
        double tstbm = (ecutmn - prm) * (ecutmn + prm) / (ecutmn * ecutmn);
        tstbm = blcc * tstbm * (ecutmn / xcc) * (ecutmn / xcc);
        tstbm = std::log(tstbm / std::log(tstbm));
        tstbmn = std::min(tstbmn, tstbm);
    }
    return tstbmn;
}

static thread_local std::string g_oerr;

extern "C" {

int oracle_last_error(char *buf, size_t n) {
    snprintf(buf, n, "%s", g_oerr.c_str());
    return 0;
}

int oracle_create(const char *blob, const OGeom *g, const OSource *s,
                  const OTransport *tp, void **out) {
    try {
        auto *e = new Engine();
        Problem &pb = e->pb;
        pb.T.load(blob);
        pb.geo.setup(g->nx, g->ny, g->nz, g->xb, g->yb, g->zb, g->boundary_tolerance);
        pb.nvox = g->nx * g->ny * g->nz;
        pb.med.assign(pb.nvox, 0);
        pb.rhor.assign(pb.nvox, 1.0);
        for (int i = 0; i < pb.nvox; i++) {
            if (g->medium) pb.med[i] = g->medium[i];
            if (pb.med[i] < 0 || pb.med[i] >= pb.T.nmed) throw std::runtime_error("medium index out of range");
            if (g->rhor) pb.rhor[i] = g->rhor[i];
        }
        Params &P = pb.P;
        double ecut = 0, pcut = 0;
        for (auto &m : pb.T.med) { ecut = std::max(ecut, m.ae); pcut = std::max(pcut, m.ap); }
        P.ecut = std::max(ecut, tp->ecut);   // $adjust_rhor_ecut_pcut
        P.pcut = std::max(pcut, tp->pcut);
        P.rayleigh = tp->rayleigh; P.ibcmp = tp->bound_compton;
        pb.T.select_compton_rows(P.ibcmp);
        P.iphter = tp->photoelectron_angle; P.iedgfl = tp->atomic_relaxation;
        P.iprdst = tp->pair_angle; P.ibrdst = tp->brems_angle; P.spin = tp->spin;
        P.exact_bca = tp->exact_bca; P.estepe = tp->estepe; P.ximax = tp->ximax;
        P.skindepth = tp->skin_depth; P.kerma = tp->kerma;
        P.presta1 = tp->step_algorithm == 1;
        P.ibr_nist = tp->brems_xsec;
        if (P.ibr_nist != pb.T.ibr_nist)
            throw std::runtime_error("brems_xsec does not match the blob's brems tables");
        P.pair_nrc = tp->pair_xsec;
        if (P.pair_nrc && !pb.T.pair_nrc)
            throw std::runtime_error("pair_xsec = 1 needs the blob's NRC pair tables");
        P.eii = pb.T.eii;   // the blob's electron cross sections include EII or not
        if (P.skindepth <= 1e-4) {   // over the media the voxels use
            std::vector<int> used;
            for (int m = 0; m < pb.T.nmed; m++)
                if (std::find(pb.med.begin(), pb.med.end(), m) != pb.med.end()) used.push_back(m);
            P.skindepth = P.exact_bca ? 3.0 : std::exp(default_blcmin(pb.T, used, P.ecut));
        }
        Source &S = pb.src;
        S.charge = s->charge; S.spectrum = s->spectrum; S.energy = s->energy;
        S.nbins = s->nbins; S.shape = s->shape;
        S.x0 = s->x0; S.x1 = s->x1; S.y0 = s->y0; S.y1 = s->y1; S.z = s->z;
        S.ux = s->ux; S.uy = s->uy; S.uz = s->uz;
        if (S.spectrum == 1) alias_make(S, s->nbins + 1, s->spec_e, s->spec_w, 1);
        else if (S.spectrum == 2) alias_make(S, s->nbins, s->spec_e, s->spec_w, 0);
        e->sum.assign(pb.nvox, 0.0);
        e->sum2.assign(pb.nvox, 0.0);
        e->seed(1802, 9373);
        *out = e;
        return 0;
    } catch (const std::exception &ex) {
        g_oerr = ex.what();
        return -1;
    }
}

int oracle_set_streams(void *h, int32_t nstreams, int32_t nthreads) {
    auto *e = (Engine *)h;
    if (nstreams > 0) e->nstreams = nstreams;
    if (nthreads > 0) e->nthreads = nthreads;
    e->seed(e->ixx, e->jxx);
    return 0;
}

int oracle_seed(void *h, int32_t ixx, int32_t jxx) {
    ((Engine *)h)->seed(ixx, jxx);
    return 0;
}

int oracle_run(void *h, int64_t n) {
    try {
        ((Engine *)h)->run(n);
        return 0;
    } catch (const std::exception &ex) {
        g_oerr = ex.what();
        return -1;
    }
}

int oracle_grad(void *h, int64_t n, const double *target, double *grad,
                double *dose, double *loss, int32_t mode) {
    try {
        auto *e = (Engine *)h;
        if (!e->pb.P.kerma) throw std::runtime_error("the density gradient needs kerma mode");
        e->grad(n, target, grad, dose, loss, mode);
        return 0;
    } catch (const std::exception &ex) {
        g_oerr = ex.what();
        return -1;
    }
}

int oracle_get_tallies(void *h, double *sum, double *sum2, int64_t *nhist) {
    auto *e = (Engine *)h;
    std::copy(e->sum.begin(), e->sum.end(), sum);
    std::copy(e->sum2.begin(), e->sum2.end(), sum2);
    *nhist = e->nhist;
    return 0;
}

int oracle_destroy(void *h) {
    delete (Engine *)h;
    return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------
// Sampler hooks for the oracle's own pinning tests: run one interaction
// routine n times on a particle moving along +z in an infinite medium and
// return per-sample outputs.  kind:
//   0 COMPT (param = IBCMP: 0 Klein-Nishina, 1 bound Compton with rejected
//            interactions returned as -9, 3 bound Compton resampled) -> [k'/k, cos(theta_gamma), T_electron, Uj_deposit]
//   1 MOLLER  -> [T_secondary, T_primary, cos theta_1, cos theta_2]
//   2 BHABHA  -> [T_of_particle_2, T_of_particle_1, q_2, q_1]
//   3 ANNIH   -> [E_gamma1, E_gamma2, cos1, cos2]
//   4 PAIR    -> [E_total_top, E_total_below, q_top, q_below]
//   5 PHOTO   -> [E_electron_total, deposit, cos, 0]
//   6 mscat   -> [cos theta, sin theta, lambda, chia2] at energy E and
//                step of param cm (one call per sample, find_index reset)
//   7 BREMS   -> [E_gamma, E_electron_total, cos theta_gamma, 0]
//   8 RAYLEIGH-> [cos theta, 0, 0, 0]
//   9 SSCAT   -> [cos theta, sin theta, 0, 0] for screening parameter
//                chia2 = param
//  10 MSDIST  -> [x, y, z, w] after one msdist_pII step of param cm (no
//                energy loss, unit density) from the origin along +z, at
//                kinetic energy E
// mscati's default PRESTA-I BLCMIN and skin depth for the media `media`
// (every medium when NULL) of a blob at the lowest ECUT ecutmn, with the
// electron rest mass prm (<= 0: egsnrc.macros:1140's value)
extern "C" int oracle_bca_skindepth(const char *blob, double ecutmn, double prm, int32_t nmed,
                                    const int32_t *media, double *blcmin, double *skindepth) {
    try {
        Tables T;
        T.load(blob);
        std::vector<int> use;
        for (int i = 0; i < (media ? nmed : T.nmed); i++) {
            int m = media ? media[i] : i;
            if (m < 0 || m >= T.nmed) throw std::runtime_error("medium index out of range");
            use.push_back(m);
        }
        *blcmin = default_blcmin(T, use, ecutmn, prm > 0 ? prm : PRM);
        *skindepth = std::exp(*blcmin);
        return 0;
    } catch (const std::exception &ex) {
        g_oerr = ex.what();
        return -1;
    }
}

extern "C" int oracle_sample(const char *blob, int32_t kind, int32_t medium,
                             double energy, double param, int64_t n,
                             int32_t seed, double *out) {
    try {
        static thread_local std::unique_ptr<Problem> pbp;
        static thread_local std::string loaded;
        if (!pbp || loaded != blob) {
            pbp.reset(new Problem());
            pbp->T.load(blob);
            loaded = blob;
        }
        Problem &pb = *pbp;
        double xb[2] = {-1e6, 1e6};
        pb.geo.setup(1, 1, 1, xb, xb, xb, 0);
        pb.nvox = 1;
        pb.med.assign(1, medium);
        pb.rhor.assign(1, 1.0);
        Params &P = pb.P;
        double ecut = 0, pcut = 0;
        for (auto &m : pb.T.med) { ecut = std::max(ecut, m.ae); pcut = std::max(pcut, m.ap); }
        P = Params{ecut, pcut, 1, 3, 1, 1, 1, 1, 0, 1, 0.25, 0.5, 3.0, 0};
        P.ibr_nist = pb.T.ibr_nist;   // the blob's brems cross sections
        P.pair_nrc = pb.T.pair_nrc;   // and pair tables, when it has them
        P.eii = pb.T.eii;
        if (kind == 0) P.ibcmp = (int)param;           // 0 KN, 1 bound with rejection, 3 norej
        std::vector<double> tmp(1, 0.0);
        std::vector<int> touched;
        Shower sh(pb, tmp, touched);
        int is, js;
        stream_seeds(1802, seed, 0, is, js);
        sh.rng.set_state(is, js);
        const Medium &M = pb.T.med[medium];
        for (int64_t k = 0; k < n; k++) {
            Particle p{};
            p.w = 1; p.wt = 1; p.ir = 0;
            double *o = out + 4 * k;
            tmp[0] = 0;
            touched.clear();
            sh.stk.clear();
            if (kind == 0) {
                p.e = energy; p.iq = 0; sh.stk.push_back(p);
                sh.compt(M);
                if (sh.stk.size() == 2) {
                    o[0] = sh.stk[0].e / energy; o[1] = sh.stk[0].w;
                    o[2] = sh.stk[1].e - PRM; o[3] = tmp[0];
                } else { o[0] = o[1] = o[2] = o[3] = -9; }
            } else if (kind == 1 || kind == 2) {
                p.e = energy + PRM; p.iq = kind == 1 ? -1 : 1; sh.stk.push_back(p);
                if (kind == 1) sh.moller(M, std::log(energy)); else sh.bhabha(M);
                if (sh.stk.size() == 2) {
                    o[0] = sh.stk[1].e - PRM; o[1] = sh.stk[0].e - PRM;
                    o[2] = kind == 1 ? sh.stk[0].w : sh.stk[1].iq;
                    o[3] = kind == 1 ? sh.stk[1].w : sh.stk[0].iq;
                } else { o[0] = o[1] = -9; o[2] = o[3] = 0; }
            } else if (kind == 3) {
                p.e = energy + PRM; p.iq = 1; sh.stk.push_back(p);
                sh.annih();
                o[0] = sh.stk[0].e; o[1] = sh.stk[1].e; o[2] = sh.stk[0].w; o[3] = sh.stk[1].w;
            } else if (kind == 4) {
                p.e = energy; p.iq = 0; sh.stk.push_back(p);
                sh.pair(M, medium, std::log(energy));
                o[0] = sh.stk[1].e; o[1] = sh.stk[0].e; o[2] = sh.stk[1].iq; o[3] = sh.stk[0].iq;
            } else if (kind == 5) {
                p.e = energy; p.iq = 0; sh.stk.push_back(p);
                sh.photo(M, std::log(energy));
                o[0] = sh.stk[0].e; o[1] = tmp[0]; o[2] = sh.stk[0].w; o[3] = 0;
            } else if (kind == 6) {
                // msdist-like inputs for one substep of length param
                double e = energy, p2 = e * (e + RMT2), beta2 = p2 / (p2 + RMSQ);
                double chia2 = M.xcc / (4 * p2 * M.blcc);
                double lambda = param * M.blcc / beta2 / (1 + chia2);
                double chilog = std::log(1 + 1 / chia2);
                double q1 = 2 * chia2 * (chilog * (1 + chia2) - 1);
                bool fi = true;
                double c, s;
                sh.mscat(lambda, chia2, q1 * lambda, fi, c, s);
                o[0] = c; o[1] = s; o[2] = lambda; o[3] = chia2;
            } else if (kind == 7) {
                p.e = energy + PRM; p.iq = -1; sh.stk.push_back(p);
                sh.brems(M, std::log(energy));
                o[0] = sh.stk[1].e; o[1] = sh.stk[0].e; o[2] = sh.stk[1].w; o[3] = 0;
            } else if (kind == 8) {
                double gle = std::log(energy);
                int lgle = (int)(M.ge1 * gle + M.ge0);
                double c, s;
                sh.rayleigh_sample(M, energy, gle, lgle, c, s);
                o[0] = c; o[1] = s; o[2] = 0; o[3] = 0;
            } else if (kind == 9) {
                double c, s;
                sh.sscat(param, c, s);
                o[0] = c; o[1] = s; o[2] = 0; o[3] = 0;
            } else if (kind == 10) {
                double us, vs, ws, xf, yf, zf, ustep;
                sh.msdist_pII(M, energy, 0.0, param, 1.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, us, vs, ws,
                              xf, yf, zf, ustep);
                o[0] = xf; o[1] = yf; o[2] = zf; o[3] = ws;
            } else if (kind == 11) {
                // egs_eadl_relax of a vacancy in shell `param` of element
                // Z = energy (the medium sets the cut-offs): the first
                // product (its energy, negative for an electron, 0 if none),
                // photons + 1000 electrons put on the stack, their kinetic
                // energy and the energy absorbed on the spot
                p.e = 0; p.iq = 0; sh.stk.push_back(p);
                sh.relax_depth = 0;
                double edep = sh.eadl_relax((int)(energy + 0.5), (int)param);
                double ekin = 0, first = 0;
                int nph = 0, nel = 0;
                for (size_t i = 1; i < sh.stk.size(); i++) {
                    const Particle &q = sh.stk[i];
                    double ek = q.iq == 0 ? q.e : q.e - PRM;
                    ekin += ek;
                    if (q.iq == 0) nph++; else nel++;
                    if (i == 1) first = q.iq == 0 ? ek : -ek;
                }
                o[0] = first; o[1] = nph + 1000.0 * nel + 1e6 * sh.relax_depth; o[2] = ekin; o[3] = edep;
            } else if (kind == 15) {
                // eii_sample for shell `param` (1 K .. 4 LIII) of the first
                // element of the medium, an electron of kinetic energy
                // `energy`: the energy transfer, the secondary's kinetic
                // energy (0 if absorbed) and the energy absorbed on the spot
                int iZ = (int)(M.z[0] + 0.5), ish = (int)param;
                double Uj = pb.T.bind[(size_t)iZ * 30 + ish - 1];
                p.e = energy + PRM; p.iq = -1; sh.stk.push_back(p);
                tmp[0] = 0;
                sh.eii_sample(M, ish, iZ, Uj);
                double wx = energy + PRM - sh.stk[0].e;
                double sec = sh.stk.size() > 1 && sh.stk[1].iq == -1 && std::fabs(sh.stk[1].e - (wx - Uj + PRM)) < 1e-12
                                 ? sh.stk[1].e - PRM : 0.0;
                o[0] = wx; o[1] = sec; o[2] = tmp[0]; o[3] = Uj;
            } else if (kind == 14) {
                // PAIR with pair angular sampling `param` (0..4) along +z:
                // the two particles' total energies and polar cosines
                P.iprdst = (int)param;
                p.e = energy; p.iq = 0; sh.stk.push_back(p);
                sh.pair(M, medium, std::log(energy));
                o[0] = sh.stk[0].e; o[1] = sh.stk[1].e; o[2] = sh.stk[0].w; o[3] = sh.stk[1].w;
            } else if (kind == 12) {
                // one PRESTA-I step (msdist_pI) of length param without
                // energy loss along +z from the origin
                double us, vs, ws, xf, yf, zf, ustep;
                sh.msdist_pI(M, energy, 0.0, param, 1.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, us, vs, ws,
                             xf, yf, zf, ustep);
                o[0] = xf; o[1] = yf; o[2] = zf; o[3] = ws;
            } else if (kind == 13) {
                // $SET-USTEP for a curved path param at this energy with the
                // electron stopping power of the medium, then $SET-TVSTEP
                // for a boundary just short of the straight-line step
                int lk = sh.lel(M, std::log(energy));
                double dedx = Shower::ev(M.ededx0, M.ededx1, lk, std::log(energy));
                double chia2, beta2, xi;
                double ustep = Shower::set_ustep(M.blcc, M.xcc, energy, param, dedx, chia2, beta2, xi);
                double xi_t = xi;
                double tv = Shower::set_tvstep(M.blcc, M.xcc, energy, param, dedx,
                                               ustep * (1 - 1e-12), ustep, chia2, beta2, xi);
                o[0] = ustep; o[1] = tv; o[2] = xi_t; o[3] = chia2;
            } else {
                throw std::runtime_error("unknown sampler kind");
            }
        }
        return 0;
    } catch (const std::exception &ex) {
        g_oerr = ex.what();
        return -1;
    }
}
