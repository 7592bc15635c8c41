// egsgpu_device.h -- device-side data layout of the transport engine.
//
// Everything the kernels read lives in HBM in structure-of-arrays form:
//  * cross-section tables (PWLF coefficient pairs on the photon/electron
//    log-energy grids, egsnrc.macros:1226-1248) in precision R; the
//    cumulative-probability tables (Rayleigh CDF, source alias table) in
//    FP64;
//  * the voxel grid: relative density (R) + medium index per voxel in one
//    record, plane positions in R;
//  * per-lane state: ranmar record (400 B), particle stack (SoA, lane
//    fastest), history tally hash (keys int32 + values FP64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "egsgpu_geom.h"
#include "egsgpu_rng.h"

namespace egsgpu {

constexpr int kMaxMed = 16;
constexpr int kMaxNE = 16;        // elements per medium
constexpr int kMaxShMed = 200;    // Compton shells per medium ($MXMDSH)
constexpr int kGE = 2000;         // $MXGE
constexpr int kMaxEKE = 500;      // $MXEKE

// The PWLF tables of one medium and energy bin are interleaved into one
// row, so that the lookups of a step touch one or two cache lines.
// Photon rows (a0, a1 pairs of $EVALUATE, egsnrc.macros:1226-1260):
constexpr int kGRow = 8, kGmfp = 0, kGbr1 = 2, kGbr2 = 4, kCohe = 6;
// Electron rows: the fields a step needs in the first 16 values, the
// discrete-interaction branching ratios after them.
constexpr int kERow = 32, kESig = 0, kPSig = 2, kEDedx = 4, kPDedx = 6, kTmxs = 8,
              kRepE = 10, kRepP = 11, kEArr = 12, kEBr1 = 16, kPBr1 = 18, kPBr2 = 20;

// One voxel: relative density and medium index, read with one load per
// voxel crossing.
template <typename R>
struct alignas(2 * sizeof(R)) VoxData {
    R rho;
    int med;
};

// Per-medium scalars.
template <typename R>
struct MedScalars {
    R ge0, ge1, eke0, eke1;
    R rho, ae, ap, te, thmoll;
    R delcm, bpar0, bpar1, zbrang, lzbrang, log_ap;
    R blcc, xcc, expeke1, esig_e, psig_e;
    R nb_emin, nb_lemin, nb_dlei;     // NIST brems sampling grid (ibr_nist = 1)
    R eii_cons;                       // rho N_A / sum(PZ A) for the EII shells
    int nne, n_shell, meke, sig_mono_e, sig_mono_p;
    int eii_nsh;                      // EII shells of the medium (0: none)
};

// NIST brems sampling grid: energies per medium and bins per alias table
// ($MXBRES, $MXBRXS)
constexpr int kNbGridE = 100, kNbGridX = 50;
// NRC pair grid: photon energies and positron energy-fraction points
// ($NRC-PAIR-NEE, $NRC-PAIR-NXX)
constexpr int kNrcpGridE = 84, kNrcpGridX = 65;
// electron impact ionisation: bins per shell ($N_EII_BINS)
constexpr int kEiiGridBins = 250;

template <typename R>
struct DevTables {
    using value_type = R;
    int nmed;
    const MedScalars<R> *ms;          // [nmed]
    // photon PWLF rows [(m*kGE + L-1)*kGRow + field]
    const R *grow;
    const R *pmax0, *pmax1;
    // Rayleigh sampling [m*100 + i]
    const double *ray_x, *ray_fcum, *ray_b, *ray_c;
    const int *ray_i;
    // Compton shells: per medium [m*kMaxShMed + i]; global [j]
    const R *eno;
    const int *eno_atbin, *shell;
    const R *cs_be, *cs_Jo, *cs_erfJo;
    // NIST brems alias tables [(m*kNbGridE + j-1)*(kNbGridX+1) + k] (x, f)
    // and [(m*kNbGridE + j-1)*kNbGridX + k] (w, bin); null with
    // Bethe-Heitler tables
    const R *nb_x, *nb_f, *nb_w;
    const int *nb_i;
    // NRC pair tables (pair_nrc = 1): positron energy-fraction grid
    // [kNrcpGridX], densities [(m*kNrcpGridE + ie)*kNrcpGridX + ix], alias
    // weights and bins [(m*kNrcpGridE + ie)*(kNrcpGridX-1) + ix]; photon
    // energies k = E/m from nrcp_emin to nrcp_emax, log(k-2) spaced
    const R *nrcp_x, *nrcp_f, *nrcp_w;
    const int *nrcp_i;
    R nrcp_emin, nrcp_emax, nrcp_dlei;
    // electron impact ionisation: per (medium, element) [m*kMaxNE + e] the
    // number of EII shells and the first (1-based) in the global list; per
    // global shell the log-energy grid (a, b) and 250 bins of the cross
    // section [shell*kEiiGridBins + bin] (barn/atom)
    const int *eii_no, *eii_first;
    const R *eii_a, *eii_b, *eii_xa, *eii_xb;
    const int *cs_iz, *cs_shn;   // element and shell number of Compton shell j
    // elements [m*kMaxNE + k]
    const R *z, *pz;
    // EDGE data per element [iz*30 + k], [iz*6 + k], [iz*16 + k]
    const R *bind, *iprob, *ea, *eb, *ec, *ed, *ee;
    const int *enumb;
    // EADL relaxation (egs_init_relax): element Z's shells start at
    // relax_zfirst[Z] (-1: none) in the shell arrays; a shell's transitions
    // are an alias histogram at relax_first (-1: none)
    const int *relax_zfirst, *relax_nshell, *shell_num, *relax_first, *relax_ntran,
        *relax_atbin, *relax_state;
    const R *shell_be, *relax_prob;
    // brems/pair fits [m*48 + k*8 + l]
    const R *dl;
    // electron PWLF rows [(m*kMaxEKE + L-1)*kERow + field]
    const R *erow;
    // screened Rutherford q^(2+) surface (msnew.data), [(i*8 + j)*32 + k]
    const float *ms_u, *ms_f, *ms_w;
    const int *ms_i;
    float ms_dllambi, ms_dqmsi;
    // Woodcock tracking: the majorant attenuation coefficient (1/cm), one
    // value per interval of ln E, interval L = int(wm1 ln E + wm0) in 1..wmn
    const R *wmaj;
    R wm0, wm1;
    int wmn;
};

template <typename R>
struct DevSource {
    int charge, spectrum, shape, np, type;
    R energy, x0, x1, y0, y1, z, ux, uy, uz;
    const double *xi, *wi;   // alias table (FP64 cumulative data); CDFINV columns
    const int *bin;
    int kind;                // EGSGPU_SOURCE_*
    double rot[9];           // DOSXYZnrc isource 1: r_11..r_33
    double iso[3], xcol, ycol, dsource;
    R z1, rho_max;           // egs_fano_source: box z range end, max mass density
    // phase-space particles (kind EGSGPU_SOURCE_PHSP)
    long long nphsp;
    const float *px, *py, *pu, *pv, *pe, *pw;
    const signed char *pq;
};

struct DevParams {
    int rayleigh, ibcmp, iphter, iedgfl, iprdst, ibrdst, exact_bca, kerma;
    int woodcock;             // photon flights by Woodcock (delta) tracking
    int n_split;              // DOSXYZnrc photon splitting number (0 = off)
    int nbr_split;            // radiative splitting number (0 = off)
    int play_rr;              // Russian roulette of photon-made charged particles
    int ireject;              // DOSXYZnrc range rejection below esave
    int presta1;              // transport_algorithm PRESTA-I (msdist_pI)
    int ibr_nist;             // 1: NIST brems photon energies (alias tables)
    int pair_nrc;             // 1: NRC pair energy sharing (alias tables)
    int eii;                  // 1: electron impact ionisation in MOLLER
    double ecut, pcut, skindepth;
    double esave;             // total energy (MeV) below which range rejection applies
};

// Fields of a stacked particle.
constexpr int kStkF = 11;

// Copies of each energy-bookkeeping tally region (a power of two): the
// histories add into copy lane % kExtraCopies.
constexpr int kExtraCopies = 64;

// One entry of a history slot's table: the voxel's key and its energy in
// one 16-byte record, so that a deposit touches one cache line.
struct alignas(16) SlotEntry {
    int key;                  // voxel+1, 0 = empty
    int pad;
    double val;
};

// Per-lane state arrays (lane-fastest SoA for the stack).
template <typename R>
struct DevLanes {
    int nlanes;
    RanmarRec *rng;                 // [nlanes] ranmar streams (ranmar runs)
    RanluxRec *rlx;                 // [nlanes] ranlux streams (ranlux runs)
    int mxstk;
    R *stk;   // [(k*kStkF + field)*nlanes + lane]; iq and ir stored bitwise
    int hcap;                       // hash capacity per lane (power of 2)
    SlotEntry *hent;                // [lane*hcap + slot]: voxel+1 (0 = empty), energy
    int *hlist;                     // touched slots [lane*hcap + k]
};

struct DevTally {
    double *sum, *sum2;
    unsigned long long *counters;   // 8 engine counters
};

// ---- event schedule -----------------------------------------------------
// Particles waiting for a transport kernel (SoA, index fastest).  tag =
// slot*4 + (iq+1): the history slot the particle belongs to and its charge.
// A particle deferred in the middle of its transport also carries the
// loop state of PHOTON / ELECTR (s0..s3, ph); fresh particles have ph = 0.
template <typename R>
struct DevBank {
    int cap;
    R *e, *x, *y, *z, *u, *v, *w, *wt, *dn;
    int *ir, *tag;
    R *s0, *s1, *s2, *s3;
    int *ph;
};

// History slots of the event schedule.  Every history in flight owns one
// slot: an open-addressing table of its per-voxel energy (the
// history-by-history accumulator of egs_scoring.h:63-125, shared by all
// lanes working on particles of that history, updated with atomics), the
// list of claimed entries, and the number of pending particle groups
// (particles in a bank plus lanes holding particles of the history).  The
// lane that takes pending to zero adds the history's sums into the tally
// and returns the slot to the free ring.
struct DevSlots {
    int nslots;
    int hcap;                 // entries per slot (power of 2)
    SlotEntry *ent;           // [slot*hcap + i]
    int *list;                // [slot*hcap + k]: claimed entry indices
    int *used;                // [slot]: number of claimed entries
    int *pending;             // [slot]
    int *ring;                // free slots, ring of nslots entries
};

// Control words of the event schedule (device memory).  Banks 0, 1 hold
// photons, 2, 3 charged particles (two of each: one drained by a kernel
// while particles deferred by it go to the other).
struct EvCtl {
    unsigned long long free_tail;   // ring append position
    int n[4];                       // bank fill counts
    int head[4];                    // bank consume heads
    // graph-driven runs plan the source launches on the device
    unsigned long long free_head;   // first free-ring entry not yet handed out
    long long remaining;            // histories still to start
    long long hist_base;            // index of the next history to start
    unsigned long long plan_head;   // this iteration's source: ring position,
    long long plan_base;            // first history index
    int plan_n;                     // and number of histories
    int stall;                      // no free slot while all banks are empty
    int done;                       // graph-driven runs: the run has ended
    int iters;                      // graph-driven runs: iterations that had work
};

// One simulation of a batch (egsgpu_run_batched): its geometry, voxels,
// dose sums and source.
template <typename R>
struct JobDesc {
    GeomView<R> G;
    const VoxData<R> *vox;
    double *sum, *sum2;
    DevSource<R> S;
    long long hist0;          // batched runs: histories the job had before this run
};

// All active lanes of the wave claim consecutive indices from *counter
// with one atomic (ballot + prefix count); returns this lane's index.
__device__ __forceinline__ int wave_claim(int *counter) {
    unsigned long long mask = __ballot(1);
    int lid = __lane_id();
    int leader = __ffsll((unsigned long long)mask) - 1;
    int rank = __popcll(mask & ((1ull << lid) - 1ull));
    int base = 0;
    if (lid == leader) base = atomicAdd(counter, (int)__popcll(mask));
    base = __shfl(base, leader);
    return base + rank;
}

// Score-function event log of the density-gradient replay: per lane a
// history's events (vox, depth*2 + type, value); type 0 = deposit
// (value = energy * weight), type 1 = d log p / d rho of the voxel.
struct DevGrad {
    int cap;                        // events per lane
    // [lane*cap + i]: x = voxel << 7 | stack depth << 1 | type, y = value bits
    uint2 *ev;
    double *acc;                    // [lane*kGradDepth + d] sweep accumulators (deep stacks)
    const double *w;                // per-voxel loss weights 2(D-D*)/(N m)
    double *g;                      // gradient accumulator [nvox]
};
constexpr int kGradDepth = 32;
constexpr int kGradVoxBits = 25;    // voxels addressable by the event log

}  // namespace egsgpu
