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

// EV = false: the history schedule (a lane owns its histories; per-lane
// tally hash).  EV = true: the event schedule (particles of one history
// may be spread over lanes and kernels; per-slot shared tally tables).
// LUX = false: ranmar streams; true: ranlux streams.
// HFL = true: DOSXYZnrc's HOWFARLESS transport for a homogeneous phantom.
// SPLIT = true: DOSXYZnrc's photon splitting (P.n_split > 1).
// The kernels must inline every member function they call: a call that is
// not inlined passes the shower by address, which moves the whole state
// from registers to scratch memory (about 1.5 KB per lane).  The helpers
// with several call sites, and the large phases the history kernels call
// (photon_trans, photon_interact, msdist_pII, ...), whose size the
// inliner's budget would otherwise refuse, are forced inline; the compiler
// inlines the rest.
template <typename R, bool EV = false, bool LUX = false, bool HFL = false, bool SPLIT = false>
struct Shower {
    const DevTables<R> &T;
    const GeomView<R> *Gp;   // geometry of the current particle's simulation
    const VoxData<R> *__restrict__ vox;
    const DevParams &P;
    const DevLanes<R> &L;
    int lane;
    // RNG: ring of kRing numbers per lane in LDS, read at rpos
    float *rbuf;
    int rpos, ravail;
    // particle stack: top in registers
    Part<R> p;
    int np;
    // UPHI saved state
    R A, B, C, cosphi, sinphi;
    // history tally hash
    int hused;
    int cvox;          // run-length cache: voxel and energy
    double cval;
    const DevTally &tally;
    double *tsum, *tsum2;    // dose sums of the current particle's simulation
    // event schedule: the history slot of the particles on this lane
    const DevSlots *slots = nullptr;
    EvCtl *ev_ctl = nullptr;
    int slot = -1;
    // counters
    // (32-bit per lane and launch; added into 64-bit totals at the end)
    unsigned c_psteps, c_cross, c_dep, c_flush, c_esteps, c_ovf, c_hovf, c_miss;
    unsigned hsteps;   // steps of the current history (runaway guard)
    // density-gradient replay log (null when not replaying): one 8-byte
    // record per event, written with one store
    const DevGrad *grad = nullptr;
    int nlog = 0, gdmax = 0;
    bool log_ovf = false;
    __device__ __forceinline__ void log_event(int type, int vox, R v) {
        if (nlog >= grad->cap) { log_ovf = true; return; }
        size_t i = (size_t)lane * grad->cap + nlog++;
        int depth = np < 63 ? np : 63;
        gdmax = depth > gdmax ? depth : gdmax;
        grad->ev[i] = make_uint2(((unsigned)vox << 7) | ((unsigned)depth << 1) | (unsigned)type,
                                 __float_as_uint((float)v));
    }
    // reverse sweep over the history's events: every log-likelihood term
    // is multiplied by the weighted deposits downstream of it (later events
    // of the same particle and its descendants: stack depth >= its own).
    // acc[d] is the weighted deposit downstream of the last event at depth
    // d; an event at depth `depth` ends the descendants deeper than it.
    // Histories with stacks up to 4 deep (all of them in kerma mode but
    // for pair and relaxation cascades) keep acc in registers.
    __device__ __forceinline__ void grad_sweep() {
        if (log_ovf) { c_hovf++; nlog = 0; gdmax = 0; log_ovf = false; return; }
        const uint2 *ev = grad->ev + (size_t)lane * grad->cap;
        const int dmax = gdmax < kGradDepth - 1 ? gdmax : kGradDepth - 1;
        if (dmax <= 3) {
            double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
            for (int q = nlog - 1; q >= 0; q--) {
                uint2 e = ev[q];
                int depth = (int)((e.x >> 1) & 63u);
                int vox = (int)(e.x >> 7);
                double v = (double)__uint_as_float(e.y);
                if (depth < 1) a1 = 0;
                if (depth < 2) a2 = 0;
                if (depth < 3) a3 = 0;
                if ((e.x & 1u) == 0) {
                    double wd = grad->w[vox] * v;
                    a0 += wd;
                    if (depth >= 1) a1 += wd;
                    if (depth >= 2) a2 += wd;
                    if (depth >= 3) a3 += wd;
                } else {
                    double A = depth == 0 ? a0 : depth == 1 ? a1 : depth == 2 ? a2 : a3;
                    double c = v * A;
                    if (c != 0) atomicAdd(grad->g + vox, c);
                }
            }
        } else {
            double *acc = grad->acc + (size_t)lane * kGradDepth;
            for (int d = 0; d <= dmax; d++) acc[d] = 0;
            for (int q = nlog - 1; q >= 0; q--) {
                uint2 e = ev[q];
                int depth = (int)((e.x >> 1) & 63u);
                depth = depth < dmax ? depth : dmax;
                int vox = (int)(e.x >> 7);
                double v = (double)__uint_as_float(e.y);
                for (int d = depth + 1; d <= dmax; d++) acc[d] = 0;
                if ((e.x & 1u) == 0) {
                    double wd = grad->w[vox] * v;
                    for (int d = 0; d <= depth; d++) acc[d] += wd;
                } else {
                    double c = v * acc[depth];
                    if (c != 0) atomicAdd(grad->g + vox, c);
                }
            }
        }
        nlog = 0;
        gdmax = 0;
    }

    __device__ Shower(const DevTables<R> &t, const GeomView<R> &g,
                      const VoxData<R> *vx, const DevParams &pp,
                      const DevLanes<R> &l, const DevTally &ta, int ln,
                      float *rb)
        : T(t), Gp(&g), vox(vx), P(pp), L(l), lane(ln), rbuf(rb),
          tally(ta), tsum(ta.sum), tsum2(ta.sum2) {
        rpos = 0;
        ravail = 0;
        np = 0;
        A = B = C = 0; cosphi = 1; sinphi = 0;
        hused = 0;
        cvox = -1; cval = 0;
        c_psteps = c_cross = c_dep = c_flush = c_esteps = c_ovf = c_hovf = c_miss = 0;
        hsteps = 0;
    }

    // RNG ring (ring_topup / ring_next): rtopup once per loop turn with the
    // wave converged, rnd for every number
    __device__ __forceinline__ typename RingRec<LUX>::type *rrec() const {
        if constexpr (LUX) return L.rlx + lane;
        else return L.rng + lane;
    }
    __device__ __forceinline__ void rtopup() {
        ring_topup<LUX>(rrec(), rbuf, blockDim.x, rpos, ravail);
    }
    __device__ __forceinline__ R rnd() {
        return (R)ring_next<LUX>(rrec(), rbuf, blockDim.x, rpos, ravail);
    }

    // ---------------------------------------------------------- scoring
    // this lane's history table (history schedule)
    __device__ __forceinline__ SlotEntry *hent() const { return L.hent + (size_t)lane * L.hcap; }
    __device__ __forceinline__ int *hlist() const { return L.hlist + (size_t)lane * L.hcap; }
    __device__ __forceinline__ void hash_add(int vox, double e) {
        SlotEntry *hent = this->hent();
        int *hlist = this->hlist();
        unsigned key = (unsigned)vox + 1u;
        unsigned mask = (unsigned)L.hcap - 1u;
        for (int attempt = 0; attempt < 2; attempt++) {
            unsigned s = (key * 2654435761u) & mask;
            for (int probe = 0; probe < L.hcap; probe++) {
                int k = hent[s].key;
                if (k == (int)key) { hent[s].val += e; return; }
                if (k == 0) {
                    if (hused * 4 >= L.hcap * 3) break;
                    hent[s].key = (int)key;
                    hent[s].val = e;
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
    __device__ __forceinline__ void flush_table() {
        SlotEntry *hent = this->hent();
        int *hlist = this->hlist();
        for (int k = 0; k < hused; k++) {
            int s = hlist[k];
            tally_add(hent[s].key - 1, hent[s].val);
            hent[s].key = 0;
        }
        c_flush += hused;
        hused = 0;
    }
    // The history's sum e of region r into the tallies (sum and sum of
    // squares).  The dose regions are the voxels; the bookkeeping regions
    // past them are spread over kExtraCopies copies (by lane), so that the
    // end of every history does not add into the same two addresses.
    __device__ __forceinline__ void tally_add(int r, double e) {
        const int nv = Gp->nxy * Gp->nz;
        const size_t i = r < nv ? (size_t)r
                                : (size_t)nv + (size_t)(r - nv) * kExtraCopies +
                                      (size_t)(lane & (kExtraCopies - 1));
        atomicAdd(tsum + i, e);
        atomicAdd(tsum2 + i, e * e);
    }
    // History schedule: the first kSmall regions a history scores into are
    // kept in registers (a kerma-mode history touches a handful); later ones
    // go to the lane's table in memory.  A region is in one or the other.
    static constexpr int kSmall = EV ? 1 : 8;
    int sv[kSmall];
    double se[kSmall];
    int sn = 0;
    __device__ __forceinline__ void small_add(int r, double e) {
        bool done = false;
#pragma unroll
        for (int i = 0; i < kSmall; i++)
            if (!done && i < sn && sv[i] == r) { se[i] += e; done = true; }
        if (done) return;
        if (sn < kSmall) {
#pragma unroll
            for (int i = 0; i < kSmall; i++)
                if (i == sn) { sv[i] = r; se[i] = e; }
            sn++;
            return;
        }
        hash_add(r, e);
    }
    __device__ __forceinline__ void flush_small() {
#pragma unroll
        for (int i = 0; i < kSmall; i++)
            if (i < sn) tally_add(sv[i], se[i]);
        c_flush += sn;
        sn = 0;
    }
    // ausgab with iarg <= 4: score edep*wt in region ir (>= 0)
    __device__ __forceinline__ void score(int ir, R edep, R wt) {
        if (grad) {   // replay: log only, the tallies were made by the forward pass
            if (ir >= 0) log_event(0, ir, edep > 0 ? edep * wt : (R)0);
            return;
        }
        if (ir < 0 || !(edep > 0)) return;
        c_dep++;
        score_raw(ir, (double)edep * (double)wt);
    }
    // add e to region ir of the current history (run-length cache in front
    // of the history's table)
    __device__ __forceinline__ void score_raw(int ir, double e) {
        if (ir == cvox) { cval += e; return; }
        if (cvox >= 0) {
            if constexpr (EV) pend_add(cvox, cval);
            else small_add(cvox, cval);
        }
        cvox = ir;
        cval = e;
    }
    // Event schedule: a deposit that leaves the run-length cache waits in a
    // short register queue and is added to the slot table by flush_pend()
    // at the top of the kernel's next loop turn, where the transport
    // temporaries are dead and the wave has re-converged (adding it in the
    // middle of a step raised the charged-particle kernel above its
    // register budget).  One turn scores into at most three regions; a
    // fourth waits too, and a deposit beyond that (none is expected) is
    // added outside the history grouping and counted as a table overflow.
    static constexpr int kPend = EV ? 4 : 1;
    int pv[kPend];
    double pe[kPend];
    int pn = 0;
    __device__ __forceinline__ void pend_add(int v, double e) {
        if (pn < kPend) {
#pragma unroll
            for (int i = 0; i < kPend; i++)
                if (i == pn) { pv[i] = v; pe[i] = e; }
            pn++;
            return;
        }
        c_hovf++;
        tally_add(v, e);
    }
    __device__ __forceinline__ void flush_pend() {
        if constexpr (EV) {
#pragma unroll
            for (int i = 0; i < kPend; i++)
                if (i < pn) slot_add(pv[i], pe[i]);
            pn = 0;
        }
    }
    // Energy bookkeeping of the simulation, scored like deposits into two
    // regions past the voxels so that each has a history-by-history sum:
    // region nvox + kIncident the energy brought in by the primaries that
    // enter the grid, region nvox + kEscaped the energy carried out of it.
    // Without variance reduction the two and the dose sums balance:
    // incident = deposited + escaped.  A particle's energy here is its
    // kinetic energy, plus 2 m c^2 for a positron (its annihilation).
    enum { kEscaped = 0, kIncident = 1 };
    __device__ __forceinline__ void score_extra(int k, const Part<R> &q) {
        if (grad) return;
        const R RM = (R)0.5109989461;
        const R e = q.iq == 0 ? q.e : (q.iq < 0 ? q.e - RM : q.e + RM);
        if (!(e > 0) || !(q.wt > 0)) return;
        score_raw(Gp->nxy * Gp->nz + k, (double)e * (double)q.wt);
    }
    // event schedule: add e to voxel vox of the current slot's table
    __device__ __forceinline__ void slot_add(int vox, double e) {
        const DevSlots &S = *slots;
        int key = vox + 1;
        size_t base = (size_t)slot * S.hcap;
        unsigned mask = (unsigned)S.hcap - 1u;
        unsigned h = ((unsigned)key * 2654435761u) & mask;
        for (int probe = 0; probe < S.hcap; probe++) {
            SlotEntry *en = S.ent + base + h;
            int k = atomicCAS(&en->key, 0, key);
            if (k == 0) {
                int idx = atomicAdd(S.used + slot, 1);
                __hip_atomic_store(S.list + base + idx, (int)h, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
                atomicAdd(&en->val, e);
                return;
            }
            if (k == key) {
                atomicAdd(&en->val, e);
                return;
            }
            h = (h + 1) & mask;
        }
        // table full: score this part outside the history grouping (counted)
        c_hovf++;
        tally_add(vox, e);
    }
    // event schedule: this lane is done with its particles of the slot
    __device__ __forceinline__ void slot_release() {
        flush_pend();
        if (cvox >= 0) slot_add(cvox, cval);
        cvox = -1;
        cval = 0;
        const DevSlots &S = *slots;
        __threadfence();
        int left = atomicSub(S.pending + slot, 1);
        if (left == 1) {   // the history is complete
            __threadfence();
            size_t base = (size_t)slot * S.hcap;
            int n = atomicExch(S.used + slot, 0);
            for (int k = 0; k < n; k++) {
                int h = __hip_atomic_load(S.list + base + k, __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_AGENT);
                int key = atomicExch(&S.ent[base + h].key, 0);
                double e = atomicExch(&S.ent[base + h].val, 0.0);
                tally_add(key - 1, e);
            }
            c_flush += n;
            unsigned long long at = atomicAdd(&ev_ctl->free_tail, 1ull);
            __hip_atomic_store(S.ring + (at % (unsigned long long)S.nslots), slot,
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        slot = -1;
    }
    __device__ __forceinline__ void end_history() {
        hsteps = 0;
        if (cvox >= 0) small_add(cvox, cval);
        cvox = -1;
        cval = 0;
        flush_small();
        if (hused) flush_table();
    }

    // ------------------------------------------------------------ stack
    // element (k, f) of this lane's stack: a 32-bit byte offset from the
    // (uniform) base, so that the access is base register + lane offset
    __device__ __forceinline__ R &sk(int k, int f) const {
        unsigned off = ((unsigned)(k * kStkF + f) * (unsigned)L.nlanes + (unsigned)lane) *
                       (unsigned)sizeof(R);
        return *(R *)((char *)L.stk + off);
    }
    __device__ __forceinline__ static R ibits(int v) {
        if constexpr (sizeof(R) == 4) return __int_as_float(v);
        else return (R)v;
    }
    __device__ __forceinline__ static int bitsi(R v) {
        if constexpr (sizeof(R) == 4) return __float_as_int(v);
        else return (int)v;
    }
    __device__ __forceinline__ void store(int k, const Part<R> &q) {
        sk(k, 0) = q.e; sk(k, 1) = q.x; sk(k, 2) = q.y; sk(k, 3) = q.z;
        sk(k, 4) = q.u; sk(k, 5) = q.v; sk(k, 6) = q.w; sk(k, 7) = q.wt;
        sk(k, 8) = q.dnear; sk(k, 9) = ibits(q.iq); sk(k, 10) = ibits(q.ir);
    }
    __device__ __forceinline__ void load(int k, Part<R> &q) {
        q.e = sk(k, 0); q.x = sk(k, 1); q.y = sk(k, 2); q.z = sk(k, 3);
        q.u = sk(k, 4); q.v = sk(k, 5); q.w = sk(k, 6); q.wt = sk(k, 7);
        q.dnear = sk(k, 8); q.iq = bitsi(sk(k, 9)); q.ir = bitsi(sk(k, 10));
    }
    // push particle q; it is the new top
    __device__ __forceinline__ bool push(const Part<R> &q) {
        if (np >= L.mxstk) { c_ovf++; return false; }
        if (np > 0) store(np - 1, p);
        p = q;
        np++;
        return true;
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
        sk(np - 2, 9) = ibits(iq);
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
    // PHOTON (egsnrc.mortran:6439-6643) cut into phases so that a wave can
    // run one phase for all lanes that need it:
    //   kPEnter  the photon just came to the top of the stack (PCUT test);
    //   kPNewE   :PNEWENERGY: (number of MFPs to the next interaction);
    //   kPTrans  :PTRANS: voxel-by-voxel flight;
    //   kPRayl, kPCompt, kPPhoto, kPPair: the interaction chosen at the end
    //            of the flight (:PNEWMEDIUM: choice of :6560-6600).
    // Random numbers are drawn in exactly the order of the unsplit routine.
    enum { kPEnter = 0, kPNewE = 1, kPTrans = 2, kPRayl = 3, kPCompt = 4, kPPhoto = 5,
           kPPair = 6 };
    int pph = kPEnter;
    int p_med = 0, p_lgle = 1;
    R p_rhof = 1, p_gle = 0, p_dpmfp = 0, p_gmfpr0 = 0, p_cohfac = 1;

    __device__ __forceinline__ void photon_medium() {   // :PNEWMEDIUM:
        const MedScalars<R> &M = T.ms[p_med];
        int lgle = (int)(M.ge1 * p_gle + M.ge0);
        p_lgle = lgle < 1 ? 1 : (lgle > kGE ? kGE : lgle);
        int li = p_med * kGE + p_lgle - 1;
        const R *g = T.grow + (size_t)li * kGRow;
        p_gmfpr0 = g[kGmfp + 1] * p_gle + g[kGmfp];
        if (P.rayleigh) p_cohfac = g[kCohe + 1] * p_gle + g[kCohe];
    }
    // kPEnter -> kPNewE, or the photon is absorbed below PCUT
    __device__ __forceinline__ void photon_enter() {
        { VoxData<R> d = vox[p.ir]; p_med = d.med; p_rhof = d.rho; }
        if (p.e <= (R)P.pcut) { score(p.ir, p.e, p.wt); pop(); return; }
        pph = kPNewE;
    }
    // kPNewE -> kPTrans
    __device__ __forceinline__ void photon_newe() {
        if (p.wt == 0) { pop(); pph = kPEnter; return; }
        p_gle = dlog(p.e);
        R rnno35 = rnd();
        if (rnno35 == 0) rnno35 = (R)1.e-30;
        p_dpmfp = -dlog(rnno35);
        p_cohfac = 1;
        photon_medium();
        pph = kPTrans;
    }
    // kPTrans: up to kmax voxel steps; at the interaction point the type is
    // sampled (kPRayl..kPPair); a photon leaving the grid is discarded
    __device__ __forceinline__ void photon_trans(int kmax) {
        const GeomView<R> &G = *Gp;
        int iz = p.ir / G.nxy, iy = p.ir - iz * G.nxy;
        int ix = iy;
        iy = iy / G.nx;
        ix -= iy * G.nx;
        for (int k = 0; k < kmax; k++) {
            c_psteps++;
            if (++hsteps > kMaxStepsPerHistory) { c_ovf++; np = 0; pph = kPEnter; return; }
            if constexpr (HFL) {
                // HOWFARLESS: one step to the interaction point or out of the
                // grid; the voxel of the point is found for the scoring
                // (dosxyznrc.mortran:3546-3562)
                R gmfp = p_gmfpr0 / p_rhof;
                if (P.rayleigh) gmfp = gmfp * p_cohfac;
                R ustep = gmfp * p_dpmfp;
                R tb = geom_box_exit(G, p.x, p.y, p.z, p.u, p.v, p.w);
                const bool out = tb <= ustep;
                if (out) ustep = tb;
                p.x = p.x + p.u * ustep;
                p.y = p.y + p.v * ustep;
                p.z = p.z + p.w * ustep;
                R d = p_dpmfp - ustep / gmfp;
                p_dpmfp = d > 0 ? d : (R)0;
                if (out) { score_extra(kEscaped, p); pop(); pph = kPEnter; return; }
                p.ir = geom_locate(G, p.x, p.y, p.z);
                if (p_dpmfp <= (R)1.e-8) { photon_choose(); return; }
                continue;
            }
            if (P.woodcock) {
                // Woodcock (delta) tracking: a flight of p_dpmfp mean free
                // paths of the majorant attenuation coefficient ends at a
                // tentative collision, which is a real one with probability
                // mu / mu_maj (mu of the voxel there); flights do not stop
                // at voxel faces.  Same expected dose as the voxel-by-voxel
                // flight below.
                int wl = (int)(T.wm1 * p_gle + T.wm0);
                wl = wl < 1 ? 1 : (wl > T.wmn ? T.wmn : wl);
                const R mumaj = T.wmaj[wl - 1];
                const R dist = p_dpmfp / mumaj;
                const R tb = geom_box_exit(G, p.x, p.y, p.z, p.u, p.v, p.w);
                if (tb <= dist) {   // the flight leaves the grid
                    score_extra(kEscaped, p);
                    pop();
                    pph = kPEnter;
                    return;
                }
                p.x = p.x + p.u * dist;
                p.y = p.y + p.v * dist;
                p.z = p.z + p.w * dist;
                p.dnear = 0;
                p.ir = geom_locate_fast(G, p.x, p.y, p.z);
                c_cross++;
                { VoxData<R> d = vox[p.ir]; p_med = d.med; p_rhof = d.rho; }
                photon_medium();
                const R mu = p_rhof / (p_gmfpr0 * p_cohfac);
                if (mu > mumaj) c_ovf++;   // a majorant below mu (never expected)
                const R r = rnd();
                if (r * mumaj < mu) { photon_choose(); return; }   // real collision
                // virtual collision: d log(1 - mu/mu_maj) / d rho
                if (grad) log_event(1, p.ir, -(mu / p_rhof) / (mumaj - mu));
                R rn = rnd();
                if (rn == 0) rn = (R)1.e-30;
                p_dpmfp = -dlog(rn);
                continue;
            }
            R gmfp = p_gmfpr0 / p_rhof;
            if (P.rayleigh) gmfp = gmfp * p_cohfac;
            R ustep = gmfp * p_dpmfp;
            int irnew = p.ir, idisc = 0;
            if (ustep > p.dnear || p.wt <= 0) {
                int inew = geom_howfar_idx(G, p.ir, ix, iy, iz, p.x, p.y, p.z, p.u, p.v, p.w,
                                           ustep);
                if (inew != p.ir) { irnew = inew; if (inew < 0) idisc = -1; }
            }
            if (grad) log_event(1, p.ir, -ustep / (gmfp * p_rhof));
            p.x = p.x + p.u * ustep;
            p.y = p.y + p.v * ustep;
            p.z = p.z + p.w * ustep;
            p.dnear = p.dnear - ustep;
            R d = p_dpmfp - ustep / gmfp;
            p_dpmfp = d > 0 ? d : (R)0;
            int medold = p_med;
            if (irnew != p.ir) {
                c_cross++;
                p.ir = irnew;
                if (irnew >= 0) { VoxData<R> d = vox[irnew]; p_med = d.med; p_rhof = d.rho; }
                else p_med = -1;
            }
            // left the grid: ausgab(3), not scored as dose
            if (idisc < 0) { score_extra(kEscaped, p); pop(); pph = kPEnter; return; }
            if (p_med != medold) { photon_medium(); continue; }
            if (p_dpmfp <= (R)1.e-8) {   // interaction: choose its type
                photon_choose();
                return;
            }
        }
    }
    // at the interaction point: the interaction type (kPRayl..kPPair)
    __device__ __forceinline__ void photon_choose() {
        int li = p_med * kGE + p_lgle - 1;
        if (grad) log_event(1, p.ir, 1 / p_rhof);
        if (P.rayleigh) {
            R rnno37 = rnd();
            if (rnno37 <= (R)1.0 - p_cohfac) { pph = kPRayl; return; }
        }
        R rnno36 = rnd();
        const R *g = T.grow + (size_t)li * kGRow;
        R gbr1 = g[kGbr1 + 1] * p_gle + g[kGbr1];
        if (rnno36 <= gbr1 && p.e > (R)(2 * 0.5109989461)) { pph = kPPair; return; }
        R gbr2 = g[kGbr2 + 1] * p_gle + g[kGbr2];
        pph = rnno36 < gbr2 ? kPCompt : kPPhoto;
    }
    // kPRayl..kPPair: the interaction; then the photon continues
    // (:PNEWENERGY:) or the stack top changed (kPEnter)
    __device__ __forceinline__ void photon_interact() {
        const int m = p_med;
        const MedScalars<R> &M = T.ms[m];
        if (pph == kPRayl) {
            int li = m * kGE + p_lgle - 1;
            R costhe, sinthe;
            rayleigh_sample(m, p.e, p_gle, li, costhe, sinthe);
            uphi21(costhe, sinthe);
            pph = kPNewE;
            return;
        }
        // with radiative splitting and Russian roulette (play_rr) the charged
        // products are played for
        const int npold = np;
        if (pph == kPPair) {
            if constexpr (SPLIT) {
                if (P.play_rr && !rr_pair()) { pph = kPEnter; return; }
            }
            pair(m, M);
        } else if (pph == kPCompt) {
            compt(m, M);
            if constexpr (SPLIT) {
                if (P.play_rr) rr_electrons(npold);   // from NPold+1 (egsnrc.mortran:798)
            }
        } else {
            photo(m, M, p_gle);
            if constexpr (SPLIT) {   // from NPold (:5192): the photo-electron
                if (P.play_rr && rr_electrons(npold - 1)) { pph = kPEnter; return; }
            }
        }
        if (p.iq != 0) { pph = kPEnter; return; }
        if (p.e < (R)P.pcut) { score(p.ir, p.e, p.wt); pop(); pph = kPEnter; return; }
        pph = kPNewE;
    }
    // $PLAY RUSSIAN ROULETTE WITH ELECTRONS FROM (first+1)
    // (egsnrc.macros:3481-3519): each charged particle of stack entries
    // first..np-1 survives with probability prob_RR = 1/nbr_split and then
    // carries 1/prob_RR times its weight; a killed one is replaced by the
    // top.  Returns true if a particle was killed.
    __device__ bool rr_electrons(int first) {
        const R prr = (R)1 / (R)P.nbr_split;
        bool killed = false;
        int k = first;
        while (k < np) {
            const bool top = k == np - 1;
            const int iq = top ? p.iq : bitsi(sk(k, 9));
            if (iq == 0) { k++; continue; }
            R r = rnd();
            if (r < prr) {
                if (top) p.wt = p.wt / prr; else sk(k, 7) = sk(k, 7) / prr;
                k++;
            } else {
                killed = true;
                if (!top) store(k, p);
                pop();
            }
        }
        return killed;
    }
    // PAIR's own roulette, played before the sampling (egsnrc.mortran
    // :4552-4577): the photon is removed with probability 1 - prob_RR
    // (false), else its weight is raised by 1/prob_RR (true)
    __device__ bool rr_pair() {
        const R prr = (R)1 / (R)P.nbr_split;
        R r = rnd();
        if (r > prr) { pop(); return false; }
        p.wt = p.wt / prr;
        return true;
    }
    // one PHOTON call of the history schedule: runs until the photon is
    // gone or a charged particle is on top
    __device__ __forceinline__ void photon() {
        pph = kPEnter;
        photon_enter();
        while (pph != kPEnter) {
            if (pph == kPNewE) photon_newe();
            else if (pph == kPTrans) photon_trans(1 << 30);
            else photon_interact();
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
    __device__ __forceinline__ void compt(int m, const MedScalars<R> &M) {
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
            // shell vacancy (:778-794): above 1 keV filled by the EADL
            // cascade, whose sub-threshold energy is absorbed here; below,
            // Uj is absorbed
            R edep = Uj > (R)1e-3 ? eadl_relax(T.cs_iz[j], T.cs_shn[j]) : Uj;
            if (edep > 0) score(p.ir, edep, p.wt);
        }
    }

    // ------------------------------------------------------------ PHOTO
    __device__ __forceinline__ void photo(int m, const MedScalars<R> &M, R gle) {
        const R RM = (R)0.5109989461;
        R peig = p.e;
        if (peig < T.ee[1 * 16 + 1]) { p.iq = -1; p.e = peig + RM; return; }
        bool do_relax = false;
        R e_vac = 0;
        int iZ = 0, vac_shell = 0;
        if (P.iedgfl) {
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
                    vac_shell = k;
                    p.iq = -1;
                }
            }
        } else {
            p.e = peig + RM;
            p.iq = -1;
        }
        if (P.iphter) photoelectron_direction();
        if (do_relax) {   // RELAX (:5191-5193)
            R edep = eadl_relax(iZ, vac_shell);
            if (edep > 0) score(p.ir, edep, p.wt);
        }
    }

    // egs_eadl_relax (egsnrc.mortran:6020-6216): fill a vacancy in shell
    // shell_egs of element iZ with fluorescent photons and Auger / Coster-
    // Kronig electrons (EADL transitions sampled from alias histograms),
    // last vacancy first.  Products above PCUT / ECUT go on the stack with
    // the position, region and weight of the particle on top when the
    // cascade starts and isotropic directions; the energy of the others and
    // of vacancies below 1 keV is returned to be absorbed on the spot.  The
    // open vacancies (at most kRelaxVac; 4 are the most a K vacancy of lead
    // needs) are packed 16 bits each into two registers.  A transition from
    // the bin that completes probabilities summing to less than 1 gives its
    // whole energy to an electron.
    static constexpr int kRelaxVac = 8;
    __device__ __forceinline__ R eadl_relax(int iZ, int shell_egs) {
        const R RM = (R)0.5109989461;
        if (!T.relax_zfirst) {   // no relaxation data: absorbed on the spot
            return shell_egs >= 1 && shell_egs <= 30 ? T.bind[iZ * 30 + shell_egs - 1] : (R)0;
        }
        if (iZ < 1 || iZ > 100) return 0;
        const int zf = T.relax_zfirst[iZ];
        if (zf < 0 || shell_egs < 1 || shell_egs > T.relax_nshell[iZ]) return 0;
        int vac = zf + shell_egs - 1;
        R Evac = T.shell_be[vac];
        if (shell_egs > 4) return Evac;
        const R Ec = (R)P.ecut - RM, Pc = (R)P.pcut, min_E = (R)1e-3;
        const Part<R> save = p;
        R edep = 0;
        unsigned long long vs0 = 0, vs1 = 0;   // vacancy stack, 16 bits per entry
        int nvac = 0;
        auto vpush = [&](int v) {
            if (nvac >= kRelaxVac) { c_ovf++; edep += T.shell_be[v]; return; }
            vs1 = (vs1 << 16) | (vs0 >> 48);
            vs0 = (vs0 << 16) | (unsigned long long)(unsigned)v;
            nvac++;
        };
        for (int guard = 0; guard < 256; guard++) {
            if (Evac < min_E || T.relax_ntran[vac] < 1) {
                edep += Evac;
            } else {
                const int first = T.relax_first[vac], n = T.relax_ntran[vac];
                // sample_alias_histogram (egs_utilities.mortran:1994-2019)
                const R r1 = rnd(), r2 = rnd();
                int ibin = 1 + (int)((R)n * r1);
                ibin = ibin > n ? n : ibin;
                if (r2 > T.relax_prob[first + ibin - 1]) ibin = T.relax_atbin[first + ibin - 1];
                const int st = T.relax_state[first + ibin - 1];
                const int base = vac - T.shell_num[vac];
                R Ef, Ecc;
                int iqf;
                if (st < 0) {
                    Ef = 0; iqf = -1; Ecc = Ec;
                } else if (st <= 64) {   // fluorescence
                    iqf = 0;
                    Ef = T.shell_be[base + st];
                    vpush(base + st);
                    Ecc = Pc;
                } else {                  // Auger or Coster-Kronig
                    iqf = -1;
                    const int n1 = base + st / 64, n2 = base + st % 64;
                    Ef = T.shell_be[n1] + T.shell_be[n2];
                    vpush(n1);
                    vpush(n2);
                    Ecc = Ec;
                }
                const R Ex = Evac - Ef;
                if (Ex > Ecc) {
                    Part<R> q = save;
                    q.iq = iqf;
                    const R cost = 2 * rnd() - 1;
                    R sint = 1 - cost * cost;
                    q.u = 0; q.v = 0; q.w = cost;
                    if (sint > 0) {
                        sint = dsqrt(sint);
                        R cphi, sphi;
                        azimuth(cphi, sphi);
                        q.u = sint * cphi;
                        q.v = sint * sphi;
                    }
                    q.e = iqf == 0 ? Ex : Ex + RM;
                    if (!push(q)) edep += Ex;   // stack full: absorbed (counted)
                } else {
                    edep += Ex;
                }
            }
            if (nvac == 0) break;
            vac = (int)(vs0 & 0xffffull);
            vs0 = (vs0 >> 16) | (vs1 << 48);
            vs1 >>= 16;
            nvac--;
            Evac = T.shell_be[vac];
        }
        return edep;
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
    // the polar angle of a pair particle of total energy ese from a photon
    // of energy eig, by the $SET-PAIR-ANGLE branch `use`
    __device__ __forceinline__ void pair_angle(int use, const MedScalars<R> &M, R eig, R ese,
                                               R &costhe, R &sinthe) {
        const R RM = (R)0.5109989461;
        if (use == 1) {
            R pse = (ese - RM) * (ese + RM);
            pse = pse > 0 ? dsqrt(pse) : (R)0;
            costhe = 1 - 2 * rnd();
            sinthe = RM * dsqrt((1 - costhe) * (1 + costhe)) / (pse * costhe + ese);
            costhe = (ese * costhe + pse) / (pse * costhe + ese);
        } else if (use == 2) {
            const R ztarg = M.zbrang, tteig = eig / RM, ttese = ese / RM;
            const R esedei = ttese / (tteig - ttese), eseder = 1 / esedei;
            const R s1 = esedei + eseder;
            const R l0 = (1 + eseder) * (1 + esedei) / (2 * tteig), l02 = l0 * l0;
            // d G(xi) / d xi of the 3D-2003 distribution in xi = 1/(1 + (theta E/m)^2)
            auto rej = [&](R xi) {
                return 2 + 3 * s1 - 4 * (s1 + 1 - 4 * (xi - (R)0.5) * (xi - (R)0.5)) *
                                       (1 + (R)0.25 * dlog(l02 + ztarg * xi * xi));
            };
            const R pt = (R)3.141593 * ttese;
            const R ximin = 1 / (1 + pt * pt);
            const R rejmin = rej(ximin);
            const R ya = (2 / tteig) * (2 / tteig);
            R xitry = dsqrt(ya / ztarg);
            xitry = xitry < (R)0.5 ? xitry : (R)0.5;
            xitry = xitry > ximin ? xitry : ximin;
            xitry = xitry > (R)0.01 ? xitry : (R)0.01;
            R galpha = 1 + (R)0.25 * dlog(ya + ztarg * xitry * xitry);
            const R gbeta = (R)0.5 * ztarg * xitry / (ya + ztarg * xitry * xitry);
            galpha = galpha - gbeta * (xitry - (R)0.5);
            R ximid = galpha / (3 * gbeta);
            ximid = galpha >= 0 ? (R)0.5 - ximid + dsqrt(ximid * ximid + (R)0.25)
                                : (R)0.5 - ximid - dsqrt(ximid * ximid + (R)0.25);
            ximid = ximid < (R)0.5 ? ximid : (R)0.5;
            ximid = ximid > ximin ? ximid : ximin;
            ximid = ximid > (R)0.01 ? ximid : (R)0.01;
            const R rejmid = rej(ximid);
            const R rejtop = (R)1.02 * (rejmin > rejmid ? rejmin : rejmid);
            R theta;
            for (;;) {
                R xitst = rnd();
                R rejtst = rej(xitst);
                R rtest = rnd();
                theta = dsqrt(1 / xitst - 1) / ttese;
                if (rtest <= rejtst / rejtop && theta < (R)3.14159265358979) break;
            }
            sinthe = sin(theta);
            costhe = cos(theta);
        } else if (use == 3) {
            costhe = 1 - 2 * rnd();
            sinthe = (1 - costhe) * (1 + costhe);
            sinthe = sinthe > 0 ? dsqrt(sinthe) : (R)0;
        } else {
            costhe = 1 - 2 * dsqrt(rnd());
            sinthe = (1 - costhe) * (1 + costhe);
            sinthe = sinthe > 0 ? dsqrt(sinthe) : (R)0;
        }
    }

    __device__ __forceinline__ void pair(int m, const MedScalars<R> &M) {
        const R RM = (R)0.5109989461, RMT2 = 2 * RM;
        R peig = p.e, eig = peig;
        R pese1, pese2;
        int iq1, iq2;
        const R *dl = T.dl + m * 48;   // dl[k*8 + l]
        const R kk = eig / RM;
        if (P.pair_nrc && kk < T.nrcp_emax) {
            // NRC energy sharing (egsnrc.mortran:4596-4620): the positron's
            // fraction of the kinetic energy from the alias table of the
            // photon energy bin
            int ibin = 1;
            if (kk > T.nrcp_emin) {
                R abin = 1 + dlog((kk - 2) / (T.nrcp_emin - 2)) * T.nrcp_dlei;
                ibin = (int)abin;
                abin = abin - ibin;
                if (rnd() < abin) ibin = ibin + 1;
                ibin = ibin < 1 ? 1 : (ibin > kNrcpGridE ? kNrcpGridE : ibin);
            }
            const size_t row = (size_t)m * kNrcpGridE + (ibin - 1);
            R xx = alias_sample(T.nrcp_x, T.nrcp_f + row * kNrcpGridX,
                                T.nrcp_w + row * (kNrcpGridX - 1),
                                T.nrcp_i + row * (kNrcpGridX - 1), kNrcpGridX - 1);
            if (xx > (R)0.5) {
                pese1 = RM * (1 + xx * (kk - 2)); iq1 = 1;
                pese2 = peig - pese1; iq2 = -1;
            } else {
                pese2 = RM * (1 + xx * (kk - 2)); iq2 = 1;
                pese1 = peig - pese2; iq1 = -1;
            }
        } else if (eig <= (R)2.1) {
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
            // $SET-PAIR-ANGLE (egsnrc.macros:2070-2163): 1 the leading term,
            // 2 KM (Motz, Olsen and Koch 3D-2003; the leading term below
            // 4.14 MeV), 3 isotropic, 4 a blend of the leading term (weight
            // E/(E+10 MeV) per particle) and 1 - 2 sqrt(r)
            int use = P.iprdst;
            if (use == 4) use = rnd() < pese1 / (pese1 + 10) ? 1 : 4;
            else if (use == 2 && eig < (R)4.14) use = 1;
            for (int ichrg = 1; ichrg <= 2; ichrg++) {
                R ese = ichrg == 1 ? pese1 : pese2;
                if (ichrg == 2 && P.iprdst == 4) use = rnd() < ese / (ese + 10) ? 1 : 4;
                R costhe, sinthe;
                pair_angle(use, M, eig, ese, costhe, sinthe);
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

    // ------------------------------------------------------------ ELECTR
    // egsnrc.mortran:985-1823: PRESTA-II or PRESTA-I condensed history with
    // exact boundary crossing (single elastic scattering within skindepth of
    // a boundary) or PRESTA-I boundary crossing (approximate steps with the
    // path-length correction), spin effects off, EII off.
    __device__ __forceinline__ int lel(const MedScalars<R> &M, R elke) const {
        int L = (int)(M.eke1 * elke + M.eke0);
        return L < 1 ? 1 : (L > M.meke ? M.meke : L);
    }
    // row of the electron tables for medium m, energy bin L (1-based)
    __device__ __forceinline__ const R *erow(int m, int L) const {
        return T.erow + (size_t)(m * kMaxEKE + L - 1) * kERow;
    }
    // PWLF value of field f at bin L
    __device__ __forceinline__ R ev(int f, int m, int L, R x) const {
        const R *r = erow(m, L);
        return r[f + 1] * x + r[f];
    }
    // range_ep(j) and E_array(j) of the Fortran tables (0-based j)
    __device__ __forceinline__ R repv(int m, bool pos, int j) const {
        return erow(m, j + 1)[pos ? kRepP : kRepE];
    }
    __device__ __forceinline__ R earr(int m, int j) const { return erow(m, j + 1)[kEArr]; }
    // $COMPUTE-DRANGE (egsnrc.macros:3214-3243)
    __device__ R drange(int m, bool pos, R e1, R e2, int l1, R el1, R el2) const {
        R fedep = 1 - e2 / e1;
        R elktmp = (R)0.5 * (el1 + el2 + (R)0.25 * fedep * fedep * (1 + fedep * (1 + (R)0.875 * fedep)));
        R dedxmid, aux;
        if (!pos) {
            dedxmid = 1 / ev(kEDedx, m, l1, elktmp);
            aux = erow(m, l1)[kEDedx + 1] * dedxmid;
        } else {
            dedxmid = 1 / ev(kPDedx, m, l1, elktmp);
            aux = erow(m, l1)[kPDedx + 1] * dedxmid;
        }
        R f = fedep / (2 - fedep);
        aux = aux * (1 + 2 * aux) * f * f / 6;
        return fedep * e1 * dedxmid * (1 + aux);
    }
    // $COMPUTE-ELOSS (egsnrc.macros:3290-3313)
    __device__ R eloss(int m, bool pos, R rhof, R step, R e, R el, int L) const {
        R dedxmid, aux;
        if (!pos) {
            dedxmid = ev(kEDedx, m, L, el);
            aux = erow(m, L)[kEDedx + 1] / dedxmid;
        } else {
            dedxmid = ev(kPDedx, m, L, el);
            aux = erow(m, L)[kPDedx + 1] / dedxmid;
        }
        R de = dedxmid * step * rhof;
        R fedep = de / e;
        return de * (1 - (R)0.5 * fedep * aux *
                             (1 - (R)0.333333 * fedep * (aux - 1 - (R)0.25 * fedep * (2 - aux * (4 - aux)))));
    }
    // $COMPUTE-ELOSS-G (egsnrc.macros:3315-3352)
    __device__ __forceinline__ R eloss_g(int m, const MedScalars<R> &M, bool pos, R rhof, R range, R step,
                         R eke, R elke, int lelke) const {
        R tuss = range - repv(m, pos, lelke - 1) / rhof;
        if (tuss >= step) return eloss(m, pos, rhof, step, eke, elke, lelke);
        int lt = lelke;
        tuss = (range - step) * rhof;
        if (tuss <= 0) return eke - M.te * (R)0.99;
        while (lt > 1 && tuss < repv(m, pos, lt - 1)) lt--;
        R elktmp = (R)(lt + 1 - M.eke0) / M.eke1;
        R eketmp = earr(m, lt);
        tuss = (repv(m, pos, lt) - tuss) / rhof;
        return eloss(m, pos, rhof, tuss, eketmp, elktmp, lt) + eke - eketmp;
    }

    // RANGE_ECUTRR of DOSXYZnrc (dosxyznrc.mortran:2514-2545): the
    // electron range from ECUT down to AE in medium m at its default density
    __device__ __forceinline__ R range_ecut(int m, const MedScalars<R> &M) const {
        const R eke = (R)P.ecut - (R)0.5109989461;
        if (!(eke > earr(m, 0))) return 0;
        R elk = dlog(eke);
        int l = lel(M, elk);
        R ekei = earr(m, l - 1);
        R elkei = (R)(l - M.eke0) / M.eke1;
        return drange(m, false, eke, ekei, l, elk, elkei) + repv(m, false, l - 1);
    }
    // :ECUT-DISCARD: (egsnrc.mortran:1773-1800).  A particle whose last
    // step left the grid is outside: nothing is scored as dose (its energy
    // escapes), and a positron still annihilates with the same random
    // numbers, but its photons start outside and are discarded at once, as
    // egsHowfar discards a particle outside the geometry
    // (egs_advanced_application.cpp:1355-1358).
    __device__ __forceinline__ void ecut_discard(int lelec, R peie) {
        const R RM = (R)0.5109989461;
        R edep = peie - RM;
        const bool outside = p.ir < 0;
        if (outside) score_extra(kEscaped, p);
        else score(p.ir, edep, p.wt);
        if (lelec > 0 && edep < peie) {
            const int n0 = np;
            annih_at_rest();
            if (outside) while (np >= n0) pop();
            return;
        }
        pop();
    }

    // ELECTR (egsnrc.mortran:985-1823) cut into single condensed-history
    // steps so that a wave re-converges after every step.  The loop state
    // of ELECTR that lives across steps is kept in the e_* members:
    //   kNew   -> the top of the stack is an electron not yet started
    //             (:NEWELECTRON:);
    //   kTstep -> sample the number of MFPs to the next discrete event
    //             (:TSTEP:);
    //   kUstep -> inside the :USTEP: loop of one TSTEP.
    //   kBrems..kAnnih -> a discrete interaction of that type is due.
    // Each call of electron_step does the work up to and including one
    // USTEP iteration (and the choice of the discrete interaction that
    // ends a TSTEP).  Random numbers are drawn in exactly the order of the
    // unsplit routine.
    enum { kNew = 0, kTstep = 1, kUstep = 2, kBrems = 3, kMoller = 4, kBhabha = 5, kAnnih = 6 };
    int eph = kNew;
    int e_med = 0, e_irl = -1;
    R e_rhof = 1, e_demfp = 0, e_sig0 = 0, e_ttstep = 0, e_elke = 0;
    bool e_ctstep = true;

    // HOWFARLESS deposition of one charged-particle step
    // (dosxyznrc.mortran:3565-3765 and 3830-3860): a step that stays in its
    // voxel deposits there; otherwise the curved path is taken as two
    // straight pieces meeting at a hinge (placed from the initial or, with
    // probability 1/2, the final direction for a condensed-history step;
    // one straight piece for a single-scattering or boundary step) and the
    // energy is shared among the voxels the pieces cross in proportion to
    // the length in each.  Returns the voxel where the path ends, -1 if it
    // leaves the grid.
    __device__ __forceinline__ int hfl_deposit(R edep, R vstep, R tv, bool ch, R xf, R yf, R zf,
                                               R uf, R vf, R wf) {
        const GeomView<R> &G = *Gp;
        const int ir = p.ir;
        int irz = ir / G.nxy, iry = (ir - irz * G.nxy) / G.nx;
        int irx = ir - irz * G.nxy - iry * G.nx;
        R dn = G.xp[irx + 1] - p.x, a;
        a = p.x - G.xp[irx]; dn = a < dn ? a : dn;
        a = G.yp[iry + 1] - p.y; dn = a < dn ? a : dn;
        a = p.y - G.yp[iry]; dn = a < dn ? a : dn;
        a = G.zp[irz + 1] - p.z; dn = a < dn ? a : dn;
        a = p.z - G.zp[irz]; dn = a < dn ? a : dn;
        if (dn >= vstep ||
            (xf > G.xp[irx] && xf < G.xp[irx + 1] && yf > G.yp[iry] && yf < G.yp[iry + 1] &&
             zf > G.zp[irz] && zf < G.zp[irz + 1])) {
            score(ir, edep, p.wt);
            return ir;
        }
        R xi = p.x, yi = p.y, zi = p.z;
        R u1 = p.u, v1 = p.v, w1 = p.w, u2 = 0, v2 = 0, w2 = 0, t1 = tv, t2 = 0;
        const R dx = xf - xi, dy = yf - yi, dz = zf - zi;
        const R d2 = dx * dx + dy * dy + dz * dz;
        auto straight = [&]() {
            t1 = dsqrt(d2);
            if (t1 > 0) { u1 = dx / t1; v1 = dy / t1; w1 = dz / t1; }
            t2 = 0;
        };
        if (ch) {
            if (rnd() < (R)0.5) {   // hinge from the initial direction
                R c = 1 - (dx * u1 + dy * v1 + dz * w1) / tv;
                if (c < (R)1e-3) {
                    straight();
                } else {
                    t1 = (tv * tv - d2) / (2 * tv * c);
                    t2 = tv - t1;
                    if (t2 > 0) {
                        u2 = (dx - t1 * u1) / t2; v2 = (dy - t1 * v1) / t2; w2 = (dz - t1 * w1) / t2;
                    } else {
                        straight();
                    }
                }
            } else {                 // hinge from the final direction
                u2 = uf; v2 = vf; w2 = wf;
                R c = 1 - (dx * u2 + dy * v2 + dz * w2) / tv;
                if (c < (R)1e-3) {
                    straight();
                } else {
                    t2 = (tv * tv - d2) / (2 * tv * c);
                    t1 = tv - t2;
                    if (t1 > 0) {
                        u1 = (dx - t2 * u2) / t1; v1 = (dy - t2 * v2) / t1; w1 = (dz - t2 * w2) / t1;
                    } else {
                        straight();
                    }
                }
            }
        }
        const R ttot = t1 + t2;
        const R efrac = ttot > 0 ? edep / ttot : (R)0;
        int irl = ir;
        bool outside = false;
        for (int seg = 0; seg < 2 && !outside; seg++) {
            const R t = seg ? t2 : t1;
            const R uu = seg ? u2 : u1, vv = seg ? v2 : v1, ww = seg ? w2 : w1;
            if (!(t > 0)) continue;
            if (seg) { xi = xi + u1 * t1; yi = yi + v1 * t1; zi = zi + w1 * t1; }
            int dirx = 0, diry = 0, dirz = 0;
            R uii = 0, vii = 0, wii = 0, distx = (R)1e20, disty = (R)1e20, distz = (R)1e20;
            if (uu > 0) { uii = 1 / uu; dirx = 1; distx = (G.xp[irx + 1] - xi) * uii; }
            else if (uu < 0) { uii = -1 / uu; dirx = -1; distx = (xi - G.xp[irx]) * uii; }
            if (vv > 0) { vii = 1 / vv; diry = 1; disty = (G.yp[iry + 1] - yi) * vii; }
            else if (vv < 0) { vii = -1 / vv; diry = -1; disty = (yi - G.yp[iry]) * vii; }
            if (ww > 0) { wii = 1 / ww; dirz = 1; distz = (G.zp[irz + 1] - zi) * wii; }
            else if (ww < 0) { wii = -1 / ww; dirz = -1; distz = (zi - G.zp[irz]) * wii; }
            R dtot = 0;
            for (int guard = G.nx + G.ny + G.nz + 4; guard > 0; guard--) {
                const R dmax = t - dtot;
                R dl;
                int nxt = irl;
                if (distx < disty && distx < distz && distx < dmax) {
                    dl = distx;
                    irx += dirx;
                    if (irx >= 0 && irx < G.nx) {
                        disty -= distx; distz -= distx; nxt = irl + dirx;
                        distx = (G.xp[irx + 1] - G.xp[irx]) * uii;
                    } else {
                        outside = true;
                    }
                } else if (disty < distz && disty < dmax) {
                    dl = disty;
                    iry += diry;
                    if (iry >= 0 && iry < G.ny) {
                        distx -= disty; distz -= disty; nxt = irl + diry * G.nx;
                        disty = (G.yp[iry + 1] - G.yp[iry]) * vii;
                    } else {
                        outside = true;
                    }
                } else if (distz < dmax) {
                    dl = distz;
                    irz += dirz;
                    if (irz >= 0 && irz < G.nz) {
                        distx -= distz; disty -= distz; nxt = irl + dirz * G.nxy;
                        distz = (G.zp[irz + 1] - G.zp[irz]) * wii;
                    } else {
                        outside = true;
                    }
                } else {
                    dl = dmax;
                }
                score(irl, (dl > 0 ? dl : (R)0) * efrac, p.wt);
                if (outside) break;
                irl = nxt;
                dtot += dl;
                if (!(dtot < t)) break;
            }
        }
        return outside ? -1 : irl;
    }

    __device__ __forceinline__ void electron_step() {
        const R RM = (R)0.5109989461, RMT2 = 2 * RM, RMSQ = RM * RM;
        const R ecut = (R)P.ecut;
        if (P.kerma) { c_esteps++; ecut_discard(p.iq, p.e); eph = kNew; return; }
        const int lelec = p.iq;
        const bool pos = lelec > 0;
        R peie = p.e, eie = peie;
        if (eph == kNew) {   // :NEWELECTRON:
            e_irl = p.ir;
            VoxData<R> d = vox[e_irl];
            e_med = d.med;
            e_rhof = d.rho;
            if (eie <= ecut) { ecut_discard(lelec, peie); return; }
            if (p.wt == 0) { pop(); return; }
            eph = kTstep;
        }
        int m = e_med;
        const MedScalars<R> &M = T.ms[m];
        R eke = eie - RM;
        R elke, rhof = e_rhof;
        int lelke;
        if (eph == kTstep) {   // :TSTEP:
            R rnne1 = rnd();
            if (rnne1 == 0) rnne1 = (R)1.e-30;
            R demfp = -dlog(rnne1);
            e_demfp = demfp > (R)1.e-8 ? demfp : (R)1.e-8;
            elke = dlog(eke);
            lelke = lel(M, elke);
            bool mono = pos ? M.sig_mono_p : M.sig_mono_e;
            if (mono) {
                e_sig0 = ev(pos ? kPSig : kESig, m, lelke, elke) /
                         ev(pos ? kPDedx : kEDedx, m, lelke, elke);
            } else {
                e_sig0 = pos ? M.psig_e : M.esig_e;
            }
            e_ctstep = true;
            e_ttstep = 0;
            eph = kUstep;
        } else {
            elke = dlog(eke);
            lelke = lel(M, elke);
        }
        // ---- one :USTEP: iteration
        if (++hsteps > kMaxStepsPerHistory) { c_ovf++; np = 0; eph = kNew; return; }
        c_esteps++;
        R sig = e_sig0, tstep;
        if (sig <= 0) {
            tstep = (R)1e8;
            e_sig0 = (R)1.E-15;
        } else {
            if (e_ctstep) {
                R total_de = e_demfp / sig;
                R ekef = eke - total_de;
                if (ekef <= earr(m, 0)) {
                    tstep = (R)1e8;
                } else {
                    R elkef = dlog(ekef);
                    int lelkef = lel(M, elkef);
                    if (lelkef == lelke) {
                        tstep = drange(m, pos, eke, ekef, lelke, elke, elkef);
                    } else {
                        R ekei = earr(m, lelke - 1);
                        R elkei = (R)(lelke - M.eke0) / M.eke1;
                        R tuss = drange(m, pos, eke, ekei, lelke, elke, elkei);
                        ekei = earr(m, lelkef);
                        elkei = (R)(lelkef + 1 - M.eke0) / M.eke1;
                        tstep = drange(m, pos, ekei, ekef, lelkef, elkei, elkef);
                        tstep = tstep + tuss + repv(m, pos, lelke - 1) - repv(m, pos, lelkef);
                    }
                }
                e_ttstep = tstep;
                e_ctstep = false;
            }
            tstep = e_ttstep / rhof;
        }
        R dedx0 = ev(pos ? kPDedx : kEDedx, m, lelke, elke);
        R dedx = rhof * dedx0;
        R tmxs = ev(kTmxs, m, lelke, elke) / rhof;
        R range;
        {
            R ekei = earr(m, lelke - 1);
            R elkei = (R)(lelke - M.eke0) / M.eke1;
            range = (drange(m, pos, eke, ekei, lelke, elke, elkei) + repv(m, pos, lelke - 1)) / rhof;
        }
        R tustep = tstep < tmxs ? tstep : tmxs;
        tustep = tustep < range ? tustep : range;
        R tperp;
        if constexpr (HFL) tperp = geom_box_near(*Gp, p.x, p.y, p.z);
        else tperp = geom_hownear(*Gp, p.ir, p.x, p.y, p.z);
        p.dnear = tperp;
        if constexpr (SPLIT && !HFL) {
            // DOSXYZnrc range rejection ($USER-RANGE-DISCARD,
            // dosxyznrc.mortran:1050-1064): below ESAVE_GLOBAL a particle
            // whose range to ECUT is within the distance to the nearest
            // voxel boundary deposits its kinetic energy here (a positron
            // then annihilates at rest)
            if (P.ireject && (double)eie < P.esave &&
                tperp >= range - range_ecut(m, M) / rhof) {
                ecut_discard(lelec, peie);
                eph = kNew;
                return;
            }
        }
        R blccl = rhof * M.blcc;
        R p2 = eke * (eke + RMT2), beta2 = p2 / (p2 + RMSQ);
        R ssmfp = beta2 / blccl;
        R skindepth = (R)P.skindepth * ssmfp;
        R mx = tperp > skindepth ? tperp : skindepth;
        tustep = tustep < mx ? tustep : mx;
        bool callhowfar, dosingle = false, domultiple = false, callmsdist;
        R ustep, de = 0;
        R us = 0, vs = 0, ws = 0, xt = 0, yt = 0, zt = 0;
        // screening angle, beta^2 and xi of a PRESTA-I approximate step
        R ms_chia2 = 0, ms_beta2 = 1, ms_xi = 0;
        const bool exact = P.exact_bca != 0;
        // the CH-versus-boundary decision is taken in FP64
        if ((double)tustep <= (double)tperp && (!exact || (double)tustep > (double)skindepth)) {
            callhowfar = false;
            callmsdist = true;
            de = eloss_g(m, M, pos, rhof, range, tustep, eke, elke, lelke);
            if (P.presta1) msdist_pI(M, eke, de, tustep, rhof, us, vs, ws, xt, yt, zt, ustep);
            else msdist_pII(M, eke, de, tustep, rhof, us, vs, ws, xt, yt, zt, ustep);
        } else if (!exact) {
            // PRESTA-I boundary crossing: an approximate step whose multiple
            // scattering is done at its end (egsnrc.mortran:1338-1346)
            callmsdist = false;
            domultiple = true;
            ustep = set_ustep(blccl, rhof * M.xcc, eke, tustep, dedx, ms_chia2, ms_beta2, ms_xi);
            callhowfar = !(ustep < tperp);
        } else {
            callmsdist = false;
            R rnnoss = rnd();
            if (rnnoss < (R)1.e-30) rnnoss = (R)1.e-30;
            R lambda = -dlog(1 - rnnoss);
            R e1 = eke / RM + 1;
            R lambda_max = (R)0.5 * blccl * RM / dedx * e1 * e1 * e1;
            if (lambda >= 0 && lambda_max > 0) {
                R tuss = lambda < lambda_max ? lambda * ssmfp * (1 - (R)0.5 * lambda / lambda_max)
                                             : (R)0.5 * lambda * ssmfp;
                if (tuss < tustep) { tustep = tuss; dosingle = true; }
            } else {
                pop();
                eph = kNew;
                return;
            }
            ustep = tustep;
            callhowfar = !(ustep < tperp);
        }
        int irold = p.ir, irnew = p.ir, idisc = 0;
        const R ustep0 = ustep;
        if (callhowfar || p.wt <= 0) {
            if constexpr (HFL) {
                R tb = geom_box_exit(*Gp, p.x, p.y, p.z, p.u, p.v, p.w);
                if (tb <= ustep) { ustep = tb; irnew = -1; idisc = -1; }
            } else {
                int inew = geom_howfar_in(*Gp, p.ir, p.x, p.y, p.z, p.u, p.v, p.w, ustep);
                if (inew != p.ir) { irnew = inew; if (inew < 0) idisc = -1; }
            }
        }
        if (ustep <= 0) ustep = 0;
        if (ustep == 0) {
            if (irnew != irold) {
                c_cross++;
                p.ir = irnew; e_irl = irnew;
                if (irnew >= 0) { VoxData<R> d = vox[irnew]; e_med = d.med; e_rhof = d.rho; }
            }
            if (eie <= ecut) { ecut_discard(lelec, peie); eph = kNew; return; }
            eph = kTstep;   // restart the TSTEP
            return;
        }
        R vstep = ustep, tvstep;
        if (callhowfar) {
            if (exact) {
                tvstep = vstep;
                if (tvstep != tustep) dosingle = false;
            } else {
                tvstep = set_tvstep(blccl, rhof * M.xcc, eke, tustep, dedx, vstep, ustep0,
                                    ms_chia2, ms_beta2, ms_xi);
                if (tvstep < 0) { c_ovf++; np = 0; eph = kNew; return; }
            }
        } else {
            tvstep = tustep;
        }
        // the energy loss of every step that did not call msdist: a step
        // cut by howfar and a single-scattering step that stays clear of the
        // boundary alike (egsnrc.mortran:1437-1452)
        if (!callmsdist) de = eloss_g(m, M, pos, rhof, range, tvstep, eke, elke, lelke);
        R save_de = de;
        R edep = de;
        R ekef = eke - de;
        R xf, yf, zf, uf, vf, wf;
        if (callmsdist) {
            uf = us; vf = vs; wf = ws; xf = xt; yf = yt; zf = zt;
        } else {
            xf = p.x + p.u * vstep;
            yf = p.y + p.v * vstep;
            zf = p.z + p.w * vstep;
            if (domultiple) {
                // multiple scattering over the approximate step, then the
                // deflection (egsnrc.mortran:1484-1502, 1564-1570)
                R lambda = blccl * tvstep / ms_beta2 / (1 + ms_chia2);
                bool find_index = true;
                int msi = 0, msj = 0;
                R omega2 = 0, costhe, sinthe;
                mscat(lambda, ms_chia2, ms_xi, find_index, msi, msj, omega2, costhe, sinthe);
                R ut = p.u, vt = p.v, wt = p.w;
                uphi21(costhe, sinthe);
                uf = p.u; vf = p.v; wf = p.w;
                p.u = ut; p.v = vt; p.w = wt;
            } else if (dosingle) {
                R ekems = ekef > ecut - RM ? ekef : ecut - RM;
                R p2s = ekems * (ekems + RMT2);
                R chia2 = M.xcc / (4 * M.blcc * p2s);
                R xi = rnd();
                xi = 2 * chia2 * xi / (1 - xi + chia2);
                R costhe = 1 - xi, sinthe = dsqrt(xi * (2 - xi));
                R ut = p.u, vt = p.v, wt = p.w;
                uphi21(costhe, sinthe);
                uf = p.u; vf = p.v; wf = p.w;
                p.u = ut; p.v = vt; p.w = wt;
            } else {
                uf = p.u; vf = p.v; wf = p.w;
            }
        }
        int hfl_ir = p.ir;
        if constexpr (HFL) hfl_ir = hfl_deposit(edep, vstep, callmsdist ? tvstep : vstep,
                                                callmsdist, xf, yf, zf, uf, vf, wf);
        else score(p.ir, edep, p.wt);
        p.x = xf; p.y = yf; p.z = zf;
        p.u = uf; p.v = vf; p.w = wf;
        p.dnear = p.dnear - vstep;
        if constexpr (HFL) {
            // the voxel at the end of the step (still inside the grid)
            if (irnew >= 0) {
                const int v = hfl_ir >= 0 ? hfl_ir : geom_locate(*Gp, p.x, p.y, p.z);
                p.ir = v; irnew = v; e_irl = v;
            }
        }
        irold = p.ir;
        peie = peie - edep;
        eie = peie;
        p.e = peie;
        if (irnew == e_irl && eie <= ecut) { ecut_discard(lelec, peie); eph = kNew; return; }
        eke = eie - RM;
        elke = dlog(eke);
        lelke = lel(M, elke);
        if (irnew != irold) {
            c_cross++;
            p.ir = irnew; e_irl = irnew;
            if (irnew >= 0) { VoxData<R> d = vox[irnew]; e_med = d.med; e_rhof = d.rho; }
        }
        if (eie <= ecut) { ecut_discard(lelec, peie); eph = kNew; return; }
        if (idisc < 0) { score_extra(kEscaped, p); pop(); eph = kNew; return; }
        if (e_med != m) { eph = kTstep; return; }
        e_demfp = e_demfp - save_de * sig;
        e_ttstep = e_ttstep - tvstep * rhof;
        if (e_ttstep < (R)1e-9) e_demfp = 0;
        if (!(e_demfp < (R)1.e-8)) return;   // next USTEP
        // ---- end of the TSTEP: real or fictitious event
        {
            const R dd = ev(pos ? kPDedx : kEDedx, m, lelke, elke);
            R sigf = ev(pos ? kPSig : kESig, m, lelke, elke) / dd;
            R rfict = rnd();
            if (!(rfict <= sigf / e_sig0)) { eph = kTstep; return; }
        }
        // ---- discrete interaction (:1652-1744): its type is chosen here,
        // the interaction itself is done by electron_interact()
        eph = kNew;
        e_elke = elke;
        if (!pos) {
            R ebr1 = ev(kEBr1, m, lelke, elke);
            R rnno24 = rnd();
            if (rnno24 <= ebr1 || p.e <= M.thmoll) {
                if (!(rnno24 <= ebr1) && ebr1 <= 0) return;
                eph = kBrems;
                return;
            }
            eph = kMoller;
            return;
        }
        R pbr1 = ev(kPBr1, m, lelke, elke);
        R rnno25 = rnd();
        if (rnno25 < pbr1) {
            eph = kBrems;
            return;
        }
        R pbr2 = ev(kPBr2, m, lelke, elke);
        eph = rnno25 < pbr2 ? kBhabha : kAnnih;
    }

    // kBrems..kAnnih: the discrete interaction; afterwards the top of the
    // stack is a new particle (kNew)
    __device__ __forceinline__ void electron_interact() {
        const int m = e_med;
        const MedScalars<R> &M = T.ms[m];
        const int npold = np;
        if (eph == kBrems) brems(m, M, e_elke);
        else if (eph == kMoller) moller(m, M, e_elke);
        else if (eph == kBhabha) bhabha(M);
        else annih();
        if constexpr (SPLIT) {   // ausgab iarg 7 and 13 (dosxyznrc.mortran:3531-3545)
            if (P.n_split > 1 && (eph == kBrems || eph == kAnnih)) rr_photons(npold);
        }
        eph = kNew;
    }

    // ---- DOSXYZnrc photon splitting (n_split) ---------------------------
    // Russian roulette of the photons made by a radiative event of a
    // charged particle (bremsstrahlung, annihilation in flight or at rest:
    // ausgab iarg 7, 13, 14, dosxyznrc.mortran:3531-3545): each survives
    // with probability 1/n_split and then carries n_split times its weight.
    // Entries npold-1 .. np-1 of the stack are the event's products.
    __device__ __forceinline__ void rr_photons(int npold) {
        const int ns = P.n_split;
        for (int k = npold - 1 > 0 ? npold - 1 : 0; k < np; k++) {
            const bool top = k == np - 1;
            const int iq = top ? p.iq : bitsi(sk(k, 9));
            if (iq != 0) continue;
            R xsi = rnd();
            if (xsi * (R)ns > 1) {
                if (top) { p.wt = 0; p.e = 0; } else { sk(k, 7) = 0; sk(k, 0) = 0; }
            } else {
                if (top) p.wt = p.wt * (R)ns; else sk(k, 7) = sk(k, 7) * (R)ns;
            }
        }
    }

    // swap stack entries a > b (entry np-1 is the top, in registers)
    __device__ void swap_down(int a, int b) {
        if (a == b) return;
        Part<R> qa, qb;
        if (a == np - 1) qa = p; else load(a, qa);
        load(b, qb);
        store(b, qa);
        if (a == np - 1) p = qb; else store(a, qb);
    }

    // $SELECT-PHOTON-MFP of DOSXYZnrc with photon splitting
    // (dosxyznrc.mortran:855-1030): the photon on top is replaced by
    // n_split sub-photons of 1/n_split of its weight whose interaction
    // points are stratified in the number of mean free paths along its
    // line of flight (the i-th after -log(1 - (eta+i-1)/n_split) MFPs, one
    // eta for all); each interacts in turn from where the previous one did.
    // Charged products keep the sub-photon's weight; photon products (and a
    // Rayleigh-scattered sub-photon) survive only for the one sub-photon
    // chosen at random, with the full weight; the others are removed.  A
    // photon then on top is split again; otherwise PHOTON returns.
    //
    // The survivor's photons are moved below the charged products of the
    // same split, so that those are transported first: the stack then
    // holds about one split's products instead of growing by the charged
    // particles left under the survivor at every scattering generation.
    // Only the order in which a history's particles are transported
    // changes.  The parent's record waits in the entry past the stack.
    __device__ __forceinline__ void photon_split() {
        const int ns = P.n_split;
        const int isave = L.mxstk;
        const R d_eta = (R)1 / (R)ns;
        for (;;) {
            { VoxData<R> d = vox[p.ir]; p_med = d.med; p_rhof = d.rho; }
            if (p.e <= (R)P.pcut) { score(p.ir, p.e, p.wt); pop(); return; }
            if (p.wt == 0) { pop(); return; }
            p_gle = dlog(p.e);
            const R r35 = rnd() / (R)ns;
            p.wt = p.wt / (R)ns;
            p.dnear = 0;
            store(isave, p);
            pop();
            const int zb = np;   // first stack index of this split's products
            int nz = 0;          // survivor photons moved to zb, zb+1, ...
            const R xxx = rnd();
            const int isurv = 1 + (int)(xxx * (R)ns);
            R dpold = 0, eta = 1 - r35 + d_eta;
            for (int is = 1; is <= ns; is++) {
                rtopup();
                eta = eta - d_eta;
                if (eta <= 0) break;
                const R dp = -dlog(eta) - dpold;
                dpold = dpold + dp;
                {
                    Part<R> q;
                    load(isave, q);
                    if (!push(q)) break;
                }
                { VoxData<R> d = vox[p.ir]; p_med = d.med; p_rhof = d.rho; }
                p_cohfac = 1;
                photon_medium();
                p_dpmfp = dp;
                pph = kPTrans;
                photon_trans(1 << 30);
                if (pph == kPEnter) break;   // it left the grid: so would the rest
                sk(isave, 1) = p.x; sk(isave, 2) = p.y; sk(isave, 3) = p.z;
                sk(isave, 10) = ibits(p.ir);
                const int m = p_med;
                const MedScalars<R> &M = T.ms[m];
                if (pph == kPRayl) {
                    if (is != isurv) { pop(); continue; }
                    p.wt = p.wt * (R)ns;
                    int li = m * kGE + p_lgle - 1;
                    R costhe, sinthe;
                    rayleigh_sample(m, p.e, p_gle, li, costhe, sinthe);
                    uphi21(costhe, sinthe);
                    swap_down(np - 1, zb + nz++);
                    continue;
                }
                const int npold = np;
                if (pph == kPPair) pair(m, M);
                else if (pph == kPCompt) compt(m, M);
                else photo(m, M, p_gle);
                // photon products: removed, or the survivor's at full weight
                int k = npold - 1;
                while (k < np) {
                    const bool top = k == np - 1;
                    const int iq = top ? p.iq : bitsi(sk(k, 9));
                    if (iq == 0 && is != isurv) {
                        if (!top) store(k, p);   // the top takes the place
                        pop();
                    } else {
                        if (iq == 0) {
                            if (top) p.wt = p.wt * (R)ns; else sk(k, 7) = sk(k, 7) * (R)ns;
                            swap_down(k, zb + nz++);
                        }
                        k++;
                    }
                }
            }
            pph = kPEnter;
            if (np <= 0 || p.iq != 0) return;
        }
    }

    // msdist_pII (egsnrc.mortran:4050-4308), spin effects off
    __device__ __forceinline__ void msdist_pII(const MedScalars<R> &M, R e0, R el, R tustep, R rhof,
                               R &us, R &vs, R &ws, R &xf, R &yf, R &zf, R &ustep) {
        const R PRM = (R)0.5109989461, RMT2 = 2 * PRM, RMSQ = PRM * PRM;
        R blccc = M.blcc, xcccc = M.xcc;
        R e = e0 - (R)0.5 * el;
        R tau = e / PRM, tau2 = tau * tau;
        R epsilon = el / e0, epsilonp = el / e;
        e = e * (1 - epsilonp * epsilonp * (6 + 10 * tau + 5 * tau2) / (24 * tau2 + 72 * tau + 48));
        R p2 = e * (e + RMT2);
        R beta2 = p2 / (p2 + RMSQ);
        R chia2 = xcccc / (4 * p2 * blccc);
        R lambda = (R)0.5 * tustep * rhof * blccc / beta2;
        R f = epsilonp / ((tau + 1) * (tau + 2));
        R temp2 = (R)0.166666 * (4 + tau * (6 + tau * (7 + tau * (4 + tau)))) * f * f;
        lambda = lambda * (1 - temp2);
        lambda = lambda / (1 + chia2);
        R chilog = dlog(1 + 1 / chia2);
        R q1 = 2 * chia2 * (chilog * (1 + chia2) - 1);
        R gamma = 6 * chia2 * (1 + chia2) * (chilog * (1 + 2 * chia2) - 2) / q1;
        R xi = q1 * lambda;
        bool find_index = true;
        int msi = 0, msj = 0;
        R omega2 = 0;
        R w1, sint1, w2, sint2, cphi1, sphi1, cphi2, sphi2;
        mscat(lambda, chia2, xi, find_index, msi, msj, omega2, w1, sint1);
        azimuth(cphi1, sphi1);
        mscat(lambda, chia2, xi, find_index, msi, msj, omega2, w2, sint2);
        azimuth(cphi2, sphi2);
        R u2 = sint2 * cphi2, v2 = sint2 * sphi2;
        R u2p = w1 * u2 + sint1 * w2;
        us = u2p * cphi1 - v2 * sphi1;
        vs = u2p * sphi1 + v2 * cphi1;
        ws = w1 * w2 - sint1 * u2;
        xi = 2 * xi;
        R eta = dsqrt(rnd());
        R eta1 = (R)0.5 * (1 - eta);
        R delta = (R)0.9082483 - ((R)0.1020621 - (R)0.0263747 * gamma) * xi;
        R temp1 = 2 + tau;
        R temp = (2 + tau * temp1) / ((tau + 1) * temp1);
        temp = temp - (tau + 1) / ((tau + 2) * (chilog * (1 + chia2) - 1));
        temp = temp * epsilonp;
        temp1 = 1 - temp;
        delta = delta + (R)0.40824829 * (epsilon * (tau + 1) / ((tau + 2) * (chilog * (1 + chia2) - 1) *
                                                                (chilog * (1 + 2 * chia2) - 2)) -
                                         (R)0.25 * temp * temp);
        R b = eta * delta, c = eta * (1 - delta);
        R w1v2 = w1 * v2;
        R ut = b * sint1 * cphi1 + c * (cphi1 * u2 - sphi1 * w1v2) + eta1 * us * temp1;
        R vt = b * sint1 * sphi1 + c * (sphi1 * u2 + cphi1 * w1v2) + eta1 * vs * temp1;
        R wt = eta1 * (1 + temp) + b * w1 + c * w2 + eta1 * ws * temp1;
        ustep = tustep * dsqrt(ut * ut + vt * vt + wt * wt);
        R u0 = p.u, v0 = p.v, w0 = p.w;
        R sint02 = u0 * u0 + v0 * v0;
        if (sint02 > (R)1e-20) {
            R sint0 = dsqrt(sint02), sint0i = 1 / sint0;
            R cphi0 = sint0i * u0, sphi0 = sint0i * v0;
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
        xf = p.x + tustep * ut;
        yf = p.y + tustep * vt;
        zf = p.z + tustep * wt;
    }

    // msdist_pI (egsnrc.mortran:4310-4515), spin effects off: one mscat
    // angle over the step (elastic mean free paths corrected for energy
    // loss), the exact Lewis <z> along the initial direction and a lateral
    // offset of half the sine of the angle
    __device__ __forceinline__ void msdist_pI(const MedScalars<R> &M, R e0, R el, R tustep, R rhof,
                                           R &us, R &vs, R &ws, R &xf, R &yf, R &zf, R &ustep) {
        const R PRM = (R)0.5109989461, RMT2 = 2 * PRM, RMSQ = PRM * PRM;
        R e = e0 - (R)0.5 * el;
        R p2 = e * (e + RMT2);
        R beta2 = p2 / (p2 + RMSQ);
        R lambda = tustep * rhof * M.blcc / beta2;
        R factor = 1 / (1 + (R)0.9784671 * e);   // 1 / (1 + e / 2m)
        R eps = el / e0;
        eps = eps / (1 - (R)0.5 * eps);
        lambda = lambda * (1 + (R)0.25 * (1 - factor * (1 - (R)0.333333 * factor)) * eps * eps);
        R chia2 = M.xcc / (4 * M.blcc * p2);
        lambda = lambda / (1 + chia2);
        R chilog = dlog(1 + 1 / chia2);
        R xi = 2 * chia2 * (chilog * (1 + chia2) - 1) * lambda;
        bool find_index = true;
        int msi = 0, msj = 0;
        R omega2 = 0, sint, cphi, sphi;
        mscat(lambda, chia2, xi, find_index, msi, msj, omega2, ws, sint);
        azimuth(cphi, sphi);
        us = sint * cphi;
        vs = sint * sphi;
        R z = xi < (R)0.1 ? 1 - xi * ((R)0.5 - xi * ((R)0.166666667 - (R)0.041666667 * xi))
                          : (1 - dexp(-xi)) / xi;
        R r = (R)0.5 * sint, r2 = r * r, z2 = z * z;
        if (1 - z2 < r2) { r2 = 1 - z2; r = dsqrt(r2); }
        R ut = r * cphi, vt = r * sphi, wt = z;
        ustep = dsqrt(z2 + r2) * tustep;
        R u0 = p.u, v0 = p.v, w0 = p.w;
        R sint02 = u0 * u0 + v0 * v0;
        if (sint02 > (R)1e-20) {
            R sint0 = dsqrt(sint02), sint0i = 1 / sint0;
            R cphi0 = sint0i * u0, sphi0 = sint0i * v0;
            R u2p = w0 * us + sint0 * ws;
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
        xf = p.x + tustep * ut;
        yf = p.y + tustep * vt;
        zf = p.z + tustep * wt;
    }

    // $CALCULATE-XI (egsnrc.macros:1324-1348), spin effects off: the first
    // GS moment of a path len at kinetic energy ekems; blccl, xccl are
    // rhof * blcc, xcc
    __device__ static R calc_xi(R blccl, R xccl, R ekems, R len, R &chia2, R &beta2) {
        const R PRM = (R)0.5109989461, RMT2 = 2 * PRM, RMSQ = PRM * PRM;
        R p2 = ekems * (ekems + RMT2);
        beta2 = p2 / (p2 + RMSQ);
        chia2 = xccl / (4 * blccl * p2);
        R xi = (R)0.5 * xccl / p2 / beta2 * len;
        return xi * (dlog(1 + 1 / chia2) - 1 / (1 + chia2));
    }

    // $SET-USTEP (egsnrc.macros:1308-1322): mean straight-line distance of
    // the curved path tustep
    __device__ static R set_ustep(R blccl, R xccl, R eke, R tustep, R dedx, R &chia2, R &beta2,
                                  R &xi) {
        R ekems = eke - (R)0.5 * tustep * dedx;
        xi = calc_xi(blccl, xccl, ekems, tustep, chia2, beta2);
        if (xi < (R)0.1) return tustep * (1 - xi * ((R)0.5 - xi * (R)0.166667));
        return tustep * (1 - dexp(-xi)) / xi;
    }

    // $SET-TVSTEP (egsnrc.macros:1350-1384): curved path of a step that
    // howfar cut from ustep0 to vstep; -1 for xi >= 1, where the reference
    // stops with a fatal error
    __device__ static R set_tvstep(R blccl, R xccl, R eke, R tustep, R dedx, R vstep, R ustep0,
                                   R &chia2, R &beta2, R &xi) {
        if (!(vstep < ustep0)) return tustep;
        R ekems = eke - (R)0.5 * tustep * vstep / ustep0 * dedx;
        xi = calc_xi(blccl, xccl, ekems, vstep, chia2, beta2);
        if (xi < (R)0.1) return vstep * (1 + xi * ((R)0.5 + xi * (R)0.333333));
        if (xi < (R)0.999999) return -vstep * dlog(1 - xi) / xi;
        return -1;
    }

    // mscat (egsnrc.mortran:2984-3160), spin effects off
    __device__ __forceinline__ void mscat(R lambda, R chia2, R q1, bool &find_index, int &msi, int &msj,
                          R &omega2, R &cost, R &sint) {
        if (lambda <= (R)13.8) {
            R sprob = rnd();
            R explambda = dexp(-lambda);
            if (sprob < explambda) { cost = 1; sint = 0; return; }
            R wsum = (1 + lambda) * explambda;
            if (sprob < wsum) {
                R xi = rnd();
                xi = 2 * chia2 * xi / (1 - xi + chia2);
                cost = 1 - xi;
                sint = dsqrt(xi * (2 - xi));
                return;
            }
            if (lambda <= 1) {
                R wprob = explambda;
                wsum = explambda;
                cost = 1; sint = 0;
                int icount = 0;
                do {
                    icount++;
                    if (icount > 20) break;
                    wprob = wprob * lambda / icount;
                    wsum = wsum + wprob;
                    R xi = rnd();
                    xi = 2 * chia2 * xi / (1 - xi + chia2);
                    R cosz = 1 - xi;
                    R sinz = xi * (2 - xi);
                    if (sinz > (R)1.e-20) {
                        sinz = dsqrt(sinz);
                        xi = rnd();
                        R phi = xi * (R)6.2831853;
                        cost = cost * cosz - sint * sinz * cos(phi);
                        R s2 = (1 - cost) * (1 + cost);
                        sint = dsqrt(s2 > 0 ? s2 : (R)0);
                    }
                } while (!(wsum > sprob));
                return;
            }
        }
        if (find_index) {
            R llmbda = dlog(lambda);
            R ai = llmbda * (R)T.ms_dllambi;
            int i = (int)ai;
            ai = ai - i;
            R xi = rnd();
            if (xi < ai) i = i + 1;
            int j;
            if (q1 < (R)1e-3) j = 0;
            else if (q1 < (R)0.5) {
                R aj = q1 * (R)T.ms_dqmsi;
                j = (int)aj;
                aj = aj - j;
                xi = rnd();
                if (xi < aj) j = j + 1;
            } else j = 7;
            i = i < 0 ? 0 : (i > 63 ? 63 : i);
            j = j > 7 ? 7 : j;
            if (llmbda < (R)2.2299)
                omega2 = chia2 * (lambda + 4) *
                         ((R)1.347006 + llmbda * ((R)0.209364 - llmbda * ((R)0.45525 - llmbda * ((R)0.50142 - (R)0.081234 * llmbda))));
            else
                omega2 = chia2 * (lambda + 4) *
                         ((R)-2.77164 + llmbda * ((R)2.94874 - llmbda * ((R)0.1535754 - llmbda * (R)0.00552888)));
            msi = i; msj = j;
            find_index = false;
        }
        R xi = rnd();
        R ak = xi * 31;
        int k = (int)ak;
        k = k > 30 ? 30 : k;
        ak = ak - k;
        int base = (msi * 8 + msj) * 32;
        if (ak > (R)T.ms_w[base + k]) k = T.ms_i[base + k];
        R a = (R)T.ms_f[base + k], u = (R)T.ms_u[base + k];
        R du = (R)T.ms_u[base + k + 1] - u;
        xi = rnd();
        if (fabs(a) < (R)0.2) {
            R x1 = (R)0.5 * (1 - xi) * a;
            u = u + xi * du * (1 + x1 * (1 - xi * a));
        } else {
            u = u - du / a * (1 - dsqrt(1 + xi * a * (2 + a)));
        }
        xi = omega2 * u / (1 + (R)0.5 * omega2 - u);
        if (xi > (R)1.99999) xi = (R)1.99999;
        cost = 1 - xi;
        sint = dsqrt(xi * (2 - xi));
    }

    // MOLLER (egsnrc.mortran:2599-2746), EII off
    // eii_sample (egsnrc.mortran:7760-7830): ionisation of shell ish
    // (binding Uj) of element iZ; the energy transfer from the hard (Moller
    // with the binding) or soft (Bethe-like) part by rejection, the
    // secondary on the stack above AE (else absorbed here), then the
    // vacancy's EADL cascade
    __device__ __forceinline__ void eii_sample(const MedScalars<R> &M, int ish, int iZ, R Uj) {
        const R RM = (R)0.5109989461, RMT2 = 2 * RM;
        const R peie = p.e, Tk = peie - RM, tau = Tk / RM, tau1 = tau + 1;
        const R tau12 = tau1 * tau1, p2 = tau * tau + 2 * tau, beta2 = p2 / tau12;
        const R Wmax = (R)0.5 * (Tk + Uj), xmax = Uj / Wmax;
        const R c1 = (Wmax / peie) * (Wmax / peie), c2 = (2 * tau + 1) / tau12;
        const R fm_s = dlog(RMT2 * p2 / Uj) - beta2 - (R)0.5;
        const R prob_s = (R)0.66666667 * fm_s * (1 + xmax + xmax * xmax);
        R fm_h = 2 + c1 - c2;
        if (fm_h < 1) fm_h = 1;
        const R prob = fm_h + prob_s;
        R wx, frej, r3;
        do {
            R r1 = rnd(), r2 = rnd();
            r3 = rnd();
            if (r1 * prob < fm_h) {   // hard collisions
                wx = 1 / (r2 * xmax + 1 - r2);
                R wxx = wx * xmax, aux = wxx / (2 - wxx);
                frej = (1 + aux * (aux - c2) + c1 * wxx * wxx) / fm_h;
            } else {                  // soft collisions
                wx = 1 / pow(r2 * xmax * xmax * xmax + 1 - r2, (R)0.333333333);
                frej = 1 - dlog(wx) / fm_s;
            }
        } while (!(r3 < frej));
        wx = wx * Uj;
        const R h1 = (peie + RM) / Tk, pese1 = peie - wx;
        p.e = pese1;
        R dcosth = h1 * (pese1 - RM) / (pese1 + RM);
        uphi21(dsqrt(dcosth), dsqrt(1 - dcosth));
        const R pese2 = wx - Uj + RM;
        R edep = 0;
        if (pese2 > M.ae) {
            if (!push_copy()) return;
            p.e = pese2;
            p.iq = -1;
            dcosth = h1 * (pese2 - RM) / (pese2 + RM);
            uphi32(dsqrt(dcosth), -dsqrt(1 - dcosth));
        } else {
            edep = wx - Uj;
        }
        edep += eadl_relax(iZ, ish);
        if (edep > 0) score(p.ir, edep, p.wt);
    }

    // MOLLER (egsnrc.mortran:2599-2746), with the EII shell branch first
    // (:2634-2664) when the tables were built with EII
    __device__ void moller(int m, const MedScalars<R> &M, R elke) {
        const R RM = (R)0.5109989461;
        R peie = p.e, pekin = peie - RM, ekin = pekin;
        if (P.eii && M.eii_nsh > 0) {
            const int L = lel(M, elke);
            R rsh = ev(kESig, m, L, elke) * (1 - ev(kEBr1, m, L, elke)) * rnd();
            for (int el = 0; el < M.nne; el++) {
                const int nsh = T.eii_no[m * kMaxNE + el];
                if (nsh <= 0) continue;
                const int iZ = (int)(T.z[m * kMaxNE + el] + (R)0.5);
                const int ifirst = T.eii_first[m * kMaxNE + el];
                for (int ish = 1; ish <= nsh; ish++) {
                    const R Uj = T.bind[iZ * 30 + ish - 1];
                    if (!(ekin > Uj && (Uj > M.te || Uj > M.ap))) continue;
                    const int jj = ifirst + ish - 1;
                    int i = (int)(T.eii_a[jj - 1] * elke + T.eii_b[jj - 1]);
                    i = i < 1 ? 1 : (i > kEiiGridBins ? kEiiGridBins : i);
                    const size_t k = (size_t)(jj - 1) * kEiiGridBins + i - 1;
                    R sig_j = (T.eii_xa[k] * elke + T.eii_xb[k]) * T.pz[m * kMaxNE + el] *
                              M.eii_cons;
                    rsh -= sig_j;
                    if (rsh < 0) {
                        eii_sample(M, ish, iZ, Uj);
                        return;
                    }
                }
            }
        }
        if (ekin <= 2 * M.te) return;
        R t0 = ekin / RM, e0 = t0 + 1;
        R extrae = peie - M.thmoll;
        R e02 = e0 * e0;
        R g2 = t0 * t0 / e02, g3 = (2 * t0 + 1) / e02;
        R gmax = 1 + (R)1.25 * g2;
        R br, rejf4, rnno28;
        do {
            R rnno27 = rnd();
            br = M.te / (ekin - extrae * rnno27);
            R r = br / (1 - br);
            rnno28 = rnd();
            rejf4 = 1 + g2 * br * br + r * (r - g3);
            rnno28 = gmax * rnno28;
        } while (!(rnno28 <= rejf4));
        R pekse2 = br * ekin;
        R pese1 = peie - pekse2, pese2 = pekse2 + RM;
        p.e = pese1;
        R h1 = (peie + RM) / pekin;
        R dcosth = h1 * (pese1 - RM) / (pese1 + RM);
        uphi21(dsqrt(dcosth), dsqrt(1 - dcosth));
        if (!push_copy()) return;
        p.iq = -1;
        p.e = pese2;
        dcosth = h1 * (pese2 - RM) / (pese2 + RM);
        uphi32(dsqrt(dcosth), -dsqrt(1 - dcosth));
    }

    // BHABHA (egsnrc.mortran:229-307)
    __device__ void bhabha(const MedScalars<R> &M) {
        const R RM = (R)0.5109989461;
        R peip = p.e, pekin = peip - RM, ekin = pekin;
        R t0 = ekin / RM, e0 = t0 + 1;
        R yy = 1 / (t0 + 2), e02 = e0 * e0;
        R beta2 = (e02 - 1) / e02;
        R ep0 = M.te / ekin, ep0c = 1 - ep0;
        R y2 = yy * yy, yp = 1 - 2 * yy, yp2 = yp * yp;
        R b4 = yp2 * yp, b3 = b4 + yp2, b2 = yp * (3 + y2), b1 = 2 - y2;
        R br, rejf2, rnno04;
        do {
            R rnno03 = rnd();
            br = ep0 / (1 - ep0c * rnno03);
            rnno04 = rnd();
            rejf2 = 1 - beta2 * br * (b1 - br * (b2 - br * (b3 - br * b4)));
        } while (!(rnno04 <= rejf2));
        int iq_np = 1, iq_new = -1;
        if (!(br < (R)0.5)) { iq_np = -1; iq_new = 1; br = 1 - br; }
        br = br > 0 ? br : (R)0;
        R pekse2 = br * ekin;
        R pese1 = peip - pekse2, pese2 = pekse2 + RM;
        p.e = pese1;
        p.iq = iq_np;
        R h1 = (peip + RM) / pekin;
        R dcosth = h1 * (pese1 - RM) / (pese1 + RM);
        dcosth = dcosth < 1 ? dcosth : (R)1;
        uphi21(dsqrt(dcosth), dsqrt(1 - dcosth));
        if (!push_copy()) return;
        p.e = pese2;
        p.iq = iq_new;
        dcosth = h1 * (pese2 - RM) / (pese2 + RM);
        uphi32(dsqrt(dcosth), -dsqrt(1 - dcosth));
    }

    // BREMS (egsnrc.mortran:312-526): Bethe-Heitler, Koch-Motz 2BS angle
    // alias_sample1 (egs_utilities.mortran:1901-1931) over an n-bin table:
    // a bin by the alias method, then a position inside it from the linear
    // density between its ends
    __device__ R alias_sample(const R *xs, const R *fs, const R *ws, const int *ib, int n) {
        R r1 = rnd(), r2 = rnd();
        R aj = 1 + r1 * n;
        int k = (int)aj;
        if (k > n) k = n;
        aj = aj - k;
        if (aj > ws[k - 1]) k = ib[k - 1];
        R x = xs[k - 1], dx = xs[k] - x, f0 = fs[k - 1];
        if (f0 > 0) {
            R a = fs[k] / f0 - 1;
            if (fabs(a) < (R)0.2) {
                R rnno1 = (R)0.5 * (1 - r2) * a;
                return x + r2 * dx * (1 + rnno1 * (1 - r2 * a));
            }
            return x - dx / a * (1 - dsqrt(1 + r2 * a * (2 + a)));
        }
        return x + dx * dsqrt(r2);
    }

    // the NIST brems table of grid energy j (1-based) of medium m
    __device__ R nist_alias_sample(int m, int j) {
        const size_t row = (size_t)m * kNbGridE + (j - 1);
        return alias_sample(T.nb_x + row * (kNbGridX + 1), T.nb_f + row * (kNbGridX + 1),
                            T.nb_w + row * kNbGridX, T.nb_i + row * kNbGridX, kNbGridX);
    }

    __device__ void brems(int m, const MedScalars<R> &M, R elke) {
        const R RM = (R)0.5109989461;
        const R *dl = T.dl + m * 48;
        R peie = p.e, eie = peie;
        int l = eie < 50 ? 1 : 3, l1 = l + 1;
        R ekin = peie - RM;
        R brmin = M.ap / ekin;
        R waux = elke - M.log_ap;
        R a = p.u, b = p.v, c = p.w;
        R sinpsi = a * a + b * b, sindel = 0, cosdel = 0;
        if (sinpsi > (R)1e-20) {
            sinpsi = dsqrt(sinpsi);
            sindel = b / sinpsi;
            cosdel = a / sinpsi;
        }
        R ztarg = M.zbrang, tteie = eie / RM;
        R beta = dsqrt((tteie - 1) * (tteie + 1)) / tteie;
        R y2max = 2 * beta * (1 + beta) * tteie * tteie, y2maxi = 1 / y2max;
        R z2max = y2max + 1, z2maxi = dsqrt(z2max);
        // radiative splitting: nbr_split photons of 1/nbr_split of the
        // weight, the electron left with the energy of the last sample
        int nbr = 1;
        if constexpr (SPLIT) nbr = P.nbr_split > 1 ? P.nbr_split : 1;
        const int npold = np;
        const R wph = p.wt / (R)nbr;
        R pese = peie;
        // NIST cross sections: the two grid energies around ekin
        // (egsnrc.mortran:404-408)
        const bool nist = P.ibr_nist != 0;
        int jj = 0;
        R ajj = 0;
        if (nist) {
            ajj = 1 + (waux + M.log_ap - M.nb_lemin) * M.nb_dlei;
            jj = (int)ajj;
            ajj = ajj - jj;
            if (jj > kNbGridE) { jj = kNbGridE; ajj = -1; }
        }
        for (int ibr = 0; ibr < nbr; ibr++) {
        R esg, ese;
        for (;;) {
            if (nist) {
                // photon energy from the alias table of grid energy j
                // (egsnrc.mortran:412-422), no rejection
                R br;
                if (ekin > M.nb_emin) {
                    int j = rnd() < ajj ? jj + 1 : jj;
                    j = j < 1 ? 1 : (j > kNbGridE ? kNbGridE : j);
                    br = nist_alias_sample(m, j);
                } else {
                    br = rnd();
                }
                esg = M.ap * dexp(br * waux);
                pese = peie - esg;
                ese = pese;
                break;
            }
            R rnno06 = rnd(), rnno07 = rnd();
            R br = brmin * dexp(rnno06 * waux);
            esg = ekin * br;
            pese = peie - esg;
            ese = pese;
            R delta = esg / eie / ese * M.delcm, aux = ese / eie, phi1, phi2;
            if (delta < 1) {
                phi1 = dl[l - 1] + delta * (dl[8 + l - 1] + delta * dl[16 + l - 1]);
                phi2 = dl[l1 - 1] + delta * (dl[8 + l1 - 1] + delta * dl[16 + l1 - 1]);
            } else {
                phi1 = dl[24 + l - 1] + dl[32 + l - 1] * dlog(delta + dl[40 + l - 1]);
                phi2 = phi1;
            }
            R rejf = (1 + aux * aux) * phi1 - 2 * aux * phi2 / 3;
            if (rnno07 < rejf) break;
        }
        // the electron keeps pese; the photon goes on top
        if (ibr == 0) p.e = pese;
        if (!push_copy()) break;
        p.e = esg;
        p.iq = 0;
        if constexpr (SPLIT) p.wt = wph;
        R y2tst;
        if (P.ibrdst == 1) {
            R ttese = ese / RM, esedei = ttese / tteie;
            R rjarg1 = 1 + esedei * esedei, rjarg2 = rjarg1 + 2 * esedei;
            R aux = 2 * ese * tteie / esg;
            aux = aux * aux;
            R aux1 = aux * ztarg, rjarg3;
            if (aux1 > 10) rjarg3 = M.lzbrang + (1 - aux1) / (aux1 * aux1);
            else rjarg3 = dlog(aux / (1 + aux1));
            R rejmax = rjarg1 * rjarg3 - rjarg2;
            for (;;) {
                y2tst = rnd();
                R rtest = rnd();
                R aux3 = z2maxi / (y2tst + (1 - y2tst) * z2maxi);
                rtest = rtest * aux3 * rejmax;
                y2tst = aux3 * aux3 - 1;
                R a4 = aux3 * aux3 * aux3 * aux3;
                R y2tst1 = esedei * y2tst / a4;
                R aux4 = 16 * y2tst1 - rjarg2, aux5 = rjarg1 - 4 * y2tst1;
                if (rtest < aux4 + aux5 * rjarg3) break;
                R aux2 = dlog(aux / (1 + aux1 / a4));
                if (rtest < aux4 + aux5 * aux2) break;
            }
        } else {
            y2tst = rnd();
            y2tst = y2tst / (1 - y2tst + y2maxi);
        }
        R costhe = 1 - 2 * y2tst * y2maxi;
        R s2 = (1 - costhe) * (1 + costhe);
        R sinthe = dsqrt(s2 > 0 ? s2 : (R)0);
        R cphi, sphi;
        azimuth(cphi, sphi);
        if (sinpsi >= (R)1e-10) {
            R us = sinthe * cphi, vs = sinthe * sphi;
            p.u = c * cosdel * us - sindel * vs + a * costhe;
            p.v = c * sindel * us + cosdel * vs + b * costhe;
            p.w = c * costhe - sinpsi * us;
        } else {
            p.u = sinthe * cphi;
            p.v = sinthe * sphi;
            p.w = c * costhe;
        }
        }
        if constexpr (SPLIT) {
            if (nbr > 1) sk(npold - 1, 0) = pese;
        }
    }

    // ANNIH (egsnrc.mortran:70-184); with radiative splitting nbr_split
    // photon pairs of 1/nbr_split of the positron's weight
    __device__ void annih() {
        const R RM = (R)0.5109989461;
        R pavip = p.e + RM, avip = pavip;
        R a = avip / RM, g = a - 1, t = g - 1;
        R pp = dsqrt(a * t), pot = pp / t;
        R ep0 = 1 / (a + pp);
        R wsamp = dlog((1 - ep0) / ep0);
        R aa = p.u, bb = p.v, cc = p.w;
        R sinpsi = aa * aa + bb * bb, sindel = 0, cosdel = 0;
        if (sinpsi > (R)1e-20) {
            sinpsi = dsqrt(sinpsi);
            sindel = bb / sinpsi;
            cosdel = aa / sinpsi;
        }
        int nbr = 1;
        if constexpr (SPLIT) nbr = P.nbr_split > 1 ? P.nbr_split : 1;
        if (nbr > 1) p.wt = p.wt / (R)nbr;
        for (int ibr = 0; ibr < nbr; ibr++) {
        if (ibr > 0 && !push_copy()) return;
        R ep, rejf, rnno02;
        do {
            R rnno01 = rnd();
            ep = ep0 * dexp(rnno01 * wsamp);
            rnno02 = rnd();
            R q = ep * a - 1;
            rejf = 1 - q * q / (ep * (a * a - 2));
        } while (!(rnno02 <= rejf));
        R esg1 = avip * ep;
        R cphi, sphi;
        R costhe = (esg1 - RM) * pot / esg1;
        costhe = costhe > 1 ? (R)1 : (costhe < -1 ? (R)-1 : costhe);
        R sinthe = dsqrt(1 - costhe * costhe);
        azimuth(cphi, sphi);
        p.e = esg1;
        p.iq = 0;
        auto set_dir = [&](R ct, R st) {
            if (sinpsi >= (R)1e-10) {
                R us = st * cphi, vs = st * sphi;
                p.u = cc * cosdel * us - sindel * vs + aa * ct;
                p.v = cc * sindel * us + cosdel * vs + bb * ct;
                p.w = cc * ct - sinpsi * us;
            } else {
                p.u = st * cphi;
                p.v = st * sphi;
                p.w = cc * ct;
            }
        };
        set_dir(costhe, sinthe);
        R esg2 = pavip - esg1;
        if (!push_copy()) return;
        p.e = esg2;
        costhe = (esg2 - RM) * pot / esg2;
        costhe = costhe > 1 ? (R)1 : (costhe < -1 ? (R)-1 : costhe);
        set_dir(costhe, -dsqrt(1 - costhe * costhe));
        }
    }

    // ANNIH_AT_REST (egsnrc.mortran:186-224); with radiative splitting
    // nbr_split back-to-back pairs of 1/nbr_split of the positron's weight
    __device__ void annih_at_rest() {
        const R RM = (R)0.5109989461;
        const int npold = np;
        int nbr = 1;
        if constexpr (SPLIT) nbr = P.nbr_split > 1 ? P.nbr_split : 1;
        if (nbr > 1) p.wt = p.wt / (R)nbr;
        for (int ibr = 0; ibr < nbr; ibr++) {
            if (ibr > 0 && !push_copy()) return;
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
        if constexpr (SPLIT) {
            if (P.n_split > 1) rr_photons(npold);   // ausgab iarg 14
        }
    }
};

}  // namespace egsgpu
