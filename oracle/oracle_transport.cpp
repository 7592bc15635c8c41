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
    int64_t n_overflow = 0;

    Shower(const Problem &p, std::vector<double> &t, std::vector<int> &tl)
        : pb(p), T(p.T), tmp(t), touched(tl) {}

    double rnd() { return rng.uniform(); }

    void score(int ir, double edep, double wt) {
        // EGS_DoseScoring::processEvent: iarg <= 4, ir >= 0, edep > 0
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
                // RELAX / egs_eadl_relax: every relaxation product of the
                // media handled here is below the cut-offs, so the vacancy
                // energy is absorbed on the spot (ausgab iarg = 4).
                double edep = Uj;
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
        int iZ = 0;
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
                    p.iq = -1;
                }
            }
        } else {
            p.e = peig + PRM;
            p.iq = -1;
        }
        if (pb.P.iphter) photoelectron_direction();
        if (do_relax) score(top().ir, e_vac, top().wt);   // relaxation below cut-offs
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
    void pair(const Medium &M, int medium, double gle) {
        Particle &p = top();
        double peig = p.e, eig = peig;
        double pese1, pese2;
        int iq1, iq2;
        if (eig <= 2.1) {
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
        // $SET-PAIR-ANGLE with iprdst = 1 (leading-order angle), the only
        // default-path branch (egsnrc.macros:2060-2120)
        top().e = pese1;
        double ese2 = pese2;
        if (pb.P.iprdst > 0) {
            for (int ichrg = 1; ichrg <= 2; ichrg++) {
                double ese = ichrg == 1 ? pese1 : ese2;
                double pse = std::sqrt(std::max(0.0, (ese - RM) * (ese + RM)));
                double costhe = rnd();
                costhe = 1.0 - 2.0 * costhe;
                double sinthe = RM * std::sqrt((1.0 - costhe) * (1.0 + costhe)) / (pse * costhe + ese);
                costhe = (ese * costhe + pse) / (pse * costhe + ese);
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
    void electron() {
        Particle &p = top();
        int irl = p.ir;
        int lelec = p.iq;
        double peie = p.e, eie = peie;
        if (eie <= pb.P.ecut) { ecut_discard(lelec, peie); return; }
        if (pb.P.kerma) { ecut_discard(lelec, peie); return; }
        throw std::runtime_error("charged-particle transport above ECUT is not restated yet");
    }

    // :ECUT-DISCARD: / :POSITRON-ANNIHILATION: (egsnrc.mortran:1773-1800)
    void ecut_discard(int lelec, double peie) {
        Particle &p = top();
        double edep = peie - PRM;   // e-: e - prm; e+: $POSITRON-ECUT-DISCARD
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
    int32_t kerma, fp64;
};

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
        P.iphter = tp->photoelectron_angle; P.iedgfl = tp->atomic_relaxation;
        P.iprdst = tp->pair_angle; P.ibrdst = tp->brems_angle; P.spin = tp->spin;
        P.exact_bca = tp->exact_bca; P.estepe = tp->estepe; P.ximax = tp->ximax;
        P.skindepth = tp->skin_depth; P.kerma = tp->kerma;
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
