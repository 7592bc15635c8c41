/*
 * egsgpu.h -- C ABI of the MI355X (gfx950) voxel transport engine.
 *
 * The engine replaces the EGSnrc shower loop reached through egsShower()
 * (HEN_HOUSE/interface/egs_interface2.h:761-771, egs_interface2.c:92-99)
 * together with the per-step egs++ callbacks it makes into
 * EGS_XYZGeometry::howfar/hownear/isWhere (egs++/geometry/egs_nd_geometry/
 * egs_nd_geometry.h:791,963,1106), EGS_DoseScoring::processEvent
 * (egs++/ausgab_objects/egs_dose_scoring/egs_dose_scoring.h:203) and the
 * RNG refill egsFillRandomArray (egs++/egs_advanced_application.cpp:1405).
 *
 * Conventions
 *  - every function returns int status: 0 = OK, < 0 = error code;
 *    egsgpu_last_error() gives the text of the last error of the calling
 *    thread.  Nothing in the library calls exit().
 *  - all pointers are host pointers to caller-owned memory unless stated;
 *    the handle owns its device memory; tables/geometry/source are copied
 *    at create time.
 *  - one handle per host thread; calls on one handle are not re-entrant.
 *  - region index = ix + iy*nx + iz*nx*ny (x fastest), as in
 *    EGS_XYZGeometry::isWhere (egs_nd_geometry.h:791-804).
 */
#ifndef EGSGPU_H
#define EGSGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
#define EGSGPU_OK              0
#define EGSGPU_E_ARG          -1   /* invalid argument                   */
#define EGSGPU_E_HIP          -2   /* HIP runtime error                  */
#define EGSGPU_E_IO           -3   /* file could not be read/written     */
#define EGSGPU_E_FORMAT       -4   /* malformed table blob / input       */
#define EGSGPU_E_UNSUPPORTED  -5   /* option not supported by the engine */
#define EGSGPU_E_OVERFLOW     -6   /* particle stack / tally overflow    */
#define EGSGPU_E_NODEVICE     -7   /* no gfx950 device                   */

/* ---- geometry: EGS_XYZGeometry (egs_nd_geometry.h:769-1381) ---------- */
typedef struct EgsXYZGeometry {
    int32_t nx, ny, nz;          /* voxels per axis                        */
    const double *xb;            /* nx+1 increasing plane positions (cm)   */
    const double *yb;            /* ny+1                                   */
    const double *zb;            /* nz+1                                   */
    const int32_t *medium;       /* nx*ny*nz medium indices (0-based, into
                                    the media list of the table blob)      */
    const double *rhor;          /* nx*ny*nz relative density
                                    (EGS_BaseGeometry::setRelativeRho,
                                    egs_base_geometry.cpp:813-860); NULL=1 */
    double boundary_tolerance;   /* EGS_BaseGeometry boundaryTolerance;
                                    <=0 selects distanceEpsilon
                                    (egs_functions.h:80-87)                */
} EgsXYZGeometry;

/* ---- source: EGS_ParallelBeam + spectrum (egs_parallel_beam.h:150,
 *      egs_spectra.cpp:74-100,374-435, egs_alias_table.cpp:268-300) ----- */
#define EGSGPU_SPEC_MONO       0   /* EGS_MonoEnergy                       */
#define EGSGPU_SPEC_HISTOGRAM  1   /* tabulated spectrum type 1 (per MeV)  */
#define EGSGPU_SPEC_LINE       2   /* tabulated spectrum type 0 (lines)    */
#define EGSGPU_SPEC_CDFINV     3   /* DOSXYZnrc inverse-CDF grid (enflag 1):
                                      nbins grid points, spec_e = CDFINV(K,1),
                                      spec_w = CDFINV(K,2) (srcxyznrc.mortran
                                      :1874-1899, sampled as :3440-3446)   */
#define EGSGPU_SOURCE_PARALLEL_BEAM 0 /* egs_parallel_beam                   */
#define EGSGPU_SOURCE_DOSXYZ_FRONT  1 /* DOSXYZnrc isource 0: x0..x1, y0..y1
                                         on the front face (clamped to it),
                                         direction ux, uy, uz               */
#define EGSGPU_SOURCE_DOSXYZ_BEAM   2 /* DOSXYZnrc isource 1: collimator
                                         xcol x ycol at dsource (< 0) from
                                         iso, rotation rot; misses are
                                         resampled and counted             */
#define EGSGPU_SOURCE_PHSP          3 /* phase-space particles given by
                                         egsgpu_set_phsp, translated by
                                         (x0, y0, z)                       */
#define EGSGPU_SOURCE_DOSXYZ_POINT  4 /* DOSXYZnrc isource 3: a point at
                                         distance dsource (SSD) in front of
                                         the grid's first z plane on the z
                                         axis, aimed at x0..x1, y0..y1 of
                                         that plane with the cos^3 weight of
                                         srcxyznrc.mortran:3451-3495       */
#define EGSGPU_SOURCE_FANO          5 /* egs_fano_source (egs_fano_source.h:
                                         128-182): a point uniform in the box
                                         [x0,x1]x[y0,y1]x[z,z1], kept with
                                         probability (mass density of its
                                         voxel)/rho_max, so that particles
                                         start in proportion to mass; the
                                         direction isotropic               */
#define EGSGPU_SHAPE_POINT     0   /* EGS_PointShape                       */
#define EGSGPU_SHAPE_RECTANGLE 1   /* EGS_RectangleShape (z = shape_z)     */
typedef struct EgsSource {
    int32_t charge;              /* 0 photon, -1 electron, +1 positron     */
    int32_t spectrum;            /* EGSGPU_SPEC_*                          */
    double energy;               /* kinetic energy for MONO (MeV)          */
    int32_t nbins;               /* HISTOGRAM: bins; LINE: lines           */
    const double *spec_e;        /* HISTOGRAM: nbins+1 edges; LINE: nbins  */
    const double *spec_w;        /* nbins probabilities                    */
    int32_t shape;               /* EGSGPU_SHAPE_*                         */
    double x0, x1, y0, y1, z;    /* point = (x0,y0,z); rect [x0,x1]x[y0,y1]*/
    double ux, uy, uz;           /* beam direction (normalised)            */
    int32_t kind;                /* EGSGPU_SOURCE_* (0 = egs_parallel_beam) */
    double rot[9];               /* DOSXYZ_BEAM: r_11, r_12, ..., r_33     */
    double iso[3];               /* DOSXYZ_BEAM: isocentre                 */
    double xcol, ycol, dsource;  /* DOSXYZ_BEAM: widths, source distance   */
    double z1;                   /* FANO: far z face of the box            */
    double rho_max;              /* FANO: "max mass density" (g/cm^3), at
                                    least the largest voxel mass density  */
} EgsSource;

/* ---- transport parameters (":start MC transport parameter:",
 *      egs_advanced_application.cpp:482-639) --------------------------- */
typedef struct EgsTransport {
    double ecut;                 /* total energy (MeV); <=0 -> max AE      */
    double pcut;                 /* MeV; <=0 -> max AP                     */
    int32_t rayleigh;            /* IRAYLR (egsnrc.macros:2299)            */
    int32_t bound_compton;       /* IBCMP 0 KN, 1 bound with rejection, 2
                                    simple, 3 norej (default)              */
    int32_t photoelectron_angle; /* IPHTER                                 */
    int32_t atomic_relaxation;   /* IEDGFL                                 */
    int32_t pair_angle;          /* IPRDST: 0 off, 1 leading term (default),
                                    2 KM, 3 isotropic, 4 blend
                                    ($SET-PAIR-ANGLE, egsnrc.macros
                                    :2070-2163)                           */
    int32_t brems_angle;         /* IBRDST 0/1                             */
    int32_t spin;                /* spin effects (needs spinms data)       */
    int32_t exact_bca;           /* 1 exact boundary crossing (default);
                                    0 PRESTA-I boundary crossing (bca_algorithm
                                    1: ELECTR's inexact branches and
                                    $SET-USTEP / $SET-TVSTEP,
                                    egsnrc.mortran:1247-1440)              */
    double estepe;               /* ESTEPE (default 0.25)                  */
    double ximax;                /* XIMAX (default 0.5)                    */
    double skin_depth;           /* skindepth_for_bca (default 3); <= 1e-4
                                    selects mscati's default: 3 with exact
                                    boundary crossing, exp(BLCMIN) at the
                                    lowest ECUT with PRESTA-I
                                    (egsnrc.mortran:2790-2817)             */
    int32_t kerma;               /* 1: deposit every electron locally      */
    int32_t fp64;                /* 1: run the FP64 build of the kernels   */
    int32_t howfarless;          /* 1: DOSXYZnrc HOWFARLESS for a homogeneous
                                    phantom: howfar/hownear see the outer
                                    faces only, a step's energy is shared
                                    along its path (dosxyznrc.mortran
                                    :3546-3765); plain history schedule */
    int32_t n_split;             /* DOSXYZnrc photon splitting (n_split,
                                    dosxyznrc.mortran:855-1030, 3531-3545):
                                    > 1 splits every photon into n_split
                                    sub-photons and plays Russian roulette
                                    with radiative photons; 0 or 1 = off,
                                    at most 256; plain history schedule */
    int32_t nbr_split;           /* EGSnrc radiative splitting (nbr_split of
                                    BREMS, ANNIH, ANNIH_AT_REST,
                                    egsnrc.mortran:70-526; set by the
                                    egs_radiative_splitting ausgab object,
                                    egs_advanced_application.cpp:1224-1239):
                                    > 1 makes nbr_split photons (pairs) of
                                    1/nbr_split of the weight per radiative
                                    event; 0 or 1 = off, at most 256; plain
                                    history schedule; not with n_split */
    int32_t play_rr;             /* 1: Russian roulette of the charged
                                    particles made by photon interactions
                                    with probability 1/nbr_split (i_play_RR,
                                    egsnrc.macros:3481-3519) */
    int32_t ireject;             /* 1: DOSXYZnrc range rejection (IREJECT,
                                    dosxyznrc.mortran:1050-1064): a charged
                                    particle below esave whose range to ECUT
                                    is within the distance to the nearest
                                    voxel boundary deposits its kinetic
                                    energy there; plain history schedule,
                                    not with howfarless */
    double esave;                /* ESAVE_GLOBAL: total energy (MeV) */
    int32_t step_algorithm;      /* transport_algorithm: 0 PRESTA-II
                                    (msdist_pII, default), 1 PRESTA-I
                                    (msdist_pI, egsnrc.mortran:4310-4515) */
    int32_t brems_xsec;          /* ibr_nist: 0 Bethe-Heitler (default), 1 the
                                    NIST cross sections of nist_brems.data,
                                    2 the NRC ones of nrc_brems.data
                                    (init_nist_brems, egsnrc.mortran
                                    :6893-7226); egsgpu_tables_build makes
                                    its tables, egsgpu_create requires a
                                    blob built with the same value      */
    int32_t pair_xsec;           /* pair_nrc: 0 Bethe-Heitler energy sharing
                                    (default), 1 the NRC cross sections of
                                    pair_nrc1.data (init_nrc_pair,
                                    egsnrc.mortran:7256-7352, PAIR :4596
                                    -4620); egsgpu_tables_build with 1 adds
                                    the tables, which egsgpu_create then
                                    requires                              */
    char photon_xsec[16];        /* "Photon cross sections": the data set
                                    prefix xcom, epdl or si (the files
                                    <prefix>_photo/_pair/_triplet/_rayleigh
                                    .data, egsnrc.mortran:8031-8035); empty
                                    = xcom for egsgpu_tables_build and "as
                                    the blob was built" for egsgpu_create,
                                    which refuses a blob of another set  */
    char eii_xsec[16];           /* "Electron impact ionization": empty or
                                    "off", "on" (= ik) or a data set (ik,
                                    casnati, kolbenstvedt, gryzinski,
                                    penelope: eii_<name>.data; eii_init,
                                    egsnrc.mortran:7434-7758, eii_sample
                                    :7760); egsgpu_tables_build makes the
                                    tables (they change the electron cross
                                    sections), egsgpu_create requires a blob
                                    built with the same set; needs atomic
                                    relaxation                           */
} EgsTransport;

/* ---- tables ----------------------------------------------------------- */
typedef struct egsgpu_tables egsgpu_tables;
typedef struct egsgpu_handle egsgpu_handle;

/* Text of the last error on this thread (always NUL terminated). */
int egsgpu_last_error(char *buf, size_t n);

/* Number of visible gfx950 devices (0 when none). */
int egsgpu_device_count(int32_t *n);

/* Load a cross-section table blob written by the table builder
 * (PEGS4 + XCOM + msnew data prepared at initialisation, the state HATCH
 * leaves behind: egsnrc.mortran:1828-2320). */
int egsgpu_tables_load(const char *path, egsgpu_tables **out);

/* Build a table blob from the reference data files: the PEGS4 file
 * pegs_file (media looked up by name, HATCH egsnrc.mortran:2011-2176) and
 * the EGSnrc data directory hen_house/data + hen_house/pegs4/pgs4form.dat.
 * tp supplies ESTEPE/XIMAX (mscati, egsnrc.mortran:2750-2981), IBCMP,
 * IRAYLR, the brems cross sections (brems_xsec 1: NIST tables from
 * hen_house/data/nist_brems.data, 2: NRC tables from nrc_brems.data) and
 * pair_xsec (1: add the NRC pair tables
 * of hen_house/data/pair_nrc1.data).  Writes the blob to out_path. */
int egsgpu_tables_build(const char *hen_house, const char *pegs_file,
                        int32_t nmedia, const char *const *media,
                        const EgsTransport *tp, const char *out_path);
int egsgpu_tables_free(egsgpu_tables *t);
int egsgpu_tables_nmedia(const egsgpu_tables *t, int32_t *n);

/* Name (PEGS4 medium name) and nominal density (g/cm^3) of medium i of a
 * table blob, for mapping the media of a geometry onto the table
 * (egsAddMedium, egs_advanced_application.cpp:460-464). */
int egsgpu_tables_medium_info(const egsgpu_tables *t, int32_t i, char *name,
                              size_t n, double *rho);

/* mscati's default skin depth for PRESTA-I boundary crossing
 * (egsnrc.mortran:2790-2817): BLCMIN = min over media of
 * log(t / log t), t = blcc (E^2 - m^2) / xcc^2 at E = ecut (total energy,
 * MeV) with the PEGS4 blcc and xcc, and skindepth = exp(BLCMIN) elastic
 * mean free paths.  The media are the nmed indices in media, or every
 * medium of the blob when media is NULL (egsgpu_create takes the media its
 * voxels use, the media of the problem). */
int egsgpu_tables_bca_skindepth(const egsgpu_tables *t, double ecut, int32_t nmed,
                                const int32_t *media, double *blcmin, double *skindepth);

/* ---- run control ------------------------------------------------------ */
/* Create a simulation: EGS_AdvancedApplication::initSimulation analogue
 * (egs_application.cpp:753-802).  Nothing runs yet. */
int egsgpu_create(const egsgpu_tables *tab, const EgsXYZGeometry *geom,
                  const EgsSource *src, const EgsTransport *tp,
                  egsgpu_handle **out);
int egsgpu_destroy(egsgpu_handle *h);

/* Launch shape: number of independent RNG streams ("parallel jobs").
 * Stream s is seeded like EGS_RandomGenerator::createRNG(input, s):
 * ranmar (ixx, jxx + s) (egs_rndm.cpp:409-463).  0 = engine default. */
int egsgpu_set_streams(egsgpu_handle *h, int32_t nstreams,
                       int32_t block_size);

/* Scheduling of the transport work.  No counterpart in the reference,
 * whose shower loop runs one history at a time on one CPU core
 * (egsShower, egs_interface2.c:92-99).
 *  EGSGPU_SCHED_HISTORY: stream s simulates its share of the histories one
 *    after the other in one kernel, like egs-parallel job s; the result is
 *    fixed by the seeds and the number of streams.
 *  EGSGPU_SCHED_EVENT: the particles of all histories in flight wait in
 *    two banks (photons, charged particles) that a photon kernel and a
 *    condensed-history kernel drain, each lane taking the next particle of
 *    its type when its own one ends; every history in flight keeps its
 *    per-voxel sums in a slot of its own, so the uncertainty is still
 *    history-by-history.  Same physics and statistics; which stream's
 *    numbers a history uses depends on timing.
 *    In kerma mode the lanes of a wave vote every turn for one action
 *    (voxel steps of the flights, one interaction type, new histories,
 *    end-of-history flushes), so that each lane still runs its own
 *    histories with its own numbers but flights of different lengths do
 *    not serialise the wave.
 *  EGSGPU_SCHED_HISTORY_PLAIN: the history schedule without voting: each
 *    lane runs its histories in one loop.  Same histories, numbers and
 *    counters as EGSGPU_SCHED_HISTORY (the dose sums differ only by the
 *    order of the FP64 additions); kept as the plain baseline.
 * Must be chosen before the first run; the default is EGSGPU_SCHED_EVENT
 * for coupled transport and EGSGPU_SCHED_HISTORY in kerma mode. */
#define EGSGPU_SCHED_HISTORY       0
#define EGSGPU_SCHED_EVENT         1
#define EGSGPU_SCHED_HISTORY_PLAIN 2
int egsgpu_set_schedule(egsgpu_handle *h, int32_t schedule);

/* Launch shape of the event schedule: register budget of the photon and
 * charged-particle kernels as resident waves per SIMD (2, 3 or 4; fewer
 * waves = more registers per lane, 0 = keep), and the number of history
 * slots (histories in flight; 0 = keep, default four times the streams; only
 * before the first run).  No counterpart in the reference. */
int egsgpu_set_event_shape(egsgpu_handle *h, int32_t waves_photon,
                           int32_t waves_electron, int32_t nslots);

/* Work batching inside the event kernels and the kerma-mode history
 * kernel (0 = keep): voxel steps of a photon flight per turn (default 8);
 * an interaction type runs when `batch` lanes of a wave wait for it
 * (default 16) or one lane has waited `max_wait` turns (default 16); once
 * its bank is drained, a wave of the event kernels with fewer than
 * `defer_below` busy lanes (default 32, -1 = never) hands its particles,
 * mid-flight state included, to the next launch.  No counterpart in the
 * reference. */
int egsgpu_set_event_batching(egsgpu_handle *h, int32_t photon_steps,
                              int32_t batch, int32_t max_wait,
                              int32_t defer_below);

/* Event schedule driven by a HIP graph (on = 1, the default): a run (or a
 * batch of runs, egsgpu_run_batched) is one launch of a graph of as many
 * iterations (source planned on the device, photon kernel, charged-particle
 * kernel) as the previous run needed plus a margin; the device decides when
 * the run has ended, after which the kernels still queued return at once,
 * and the host reads the control words once per launch (a run that needs
 * more iterations than the graph holds launches it again).  on = 0
 * launches the kernels one by one from the host, with HIP events around
 * each launch for egsgpu_kernel_time (in graph mode only the whole run is
 * timed, egsgpu_last_kernel_ms).  Same results either way.
 * No counterpart in the reference. */
int egsgpu_set_event_graph(egsgpu_handle *h, int32_t on);

/* The last graph-driven run: graph launches it took, iterations that had
 * work, and iterations the next run's graph will hold.
 * No counterpart in the reference. */
int egsgpu_graph_stats(egsgpu_handle *h, int32_t *launches, int32_t *iterations,
                       int32_t *captured);

/* Autotuner of the event schedule: times short runs of nhist_probe
 * histories while varying, one after the other, the register budget (2 or
 * 3 waves per SIMD), the queue capacity (2, 4 or 8 history slots per
 * stream, the particle banks holding 8 per slot), the workgroup size (128
 * or 256), the deferral threshold (the number of busy lanes below which a
 * wave hands its particles to the next launch), the interaction batch and
 * the photon steps per turn, and keeps a setting when it is more than 1 %
 * faster than the best so far (the probes' own noise is below that when
 * nhist_probe is a production run's size).  The probes
 * advance the RNG streams but leave tallies, history count and counters as
 * they were.  choice (7 ints, may be NULL) receives waves photon, waves
 * electron, defer_below, batch, photon steps, slots per stream, workgroup
 * size.  A no-op for the history schedule.  No counterpart in the
 * reference. */
int egsgpu_autotune(egsgpu_handle *h, int64_t nhist_probe, int32_t *choice);

/* How photons fly between interactions.
 *  EGSGPU_TRACK_VOXEL (default): voxel by voxel, each step ending at the
 *    next voxel face (howfar) or at the interaction point, as PHOTON does
 *    (egsnrc.mortran:6439-6643) with EGS_XYZGeometry::howfar.
 *  EGSGPU_TRACK_WOODCOCK: Woodcock (delta) tracking: a flight of one mean
 *    free path of a majorant attenuation coefficient mu_maj(E) -- factor
 *    times the largest mu any voxel of the simulation has at E (0 = 1.25)
 *    -- ends at a tentative collision that is a real interaction with
 *    probability mu/mu_maj of the voxel there.  The same expected dose
 *    (statistically identical), with a few voxel lookups per photon
 *    instead of one step per voxel crossed; for the density gradient the
 *    score function then sits at the tentative collisions.  The random
 *    numbers of a history differ from the voxel-by-voxel flight.  Not with
 *    HOWFARLESS, photon splitting or batched runs.  No counterpart in the
 *    reference.  Counter 5 also counts tentative collisions where mu
 *    exceeded the majorant (none expected). */
#define EGSGPU_TRACK_VOXEL    0
#define EGSGPU_TRACK_WOODCOCK 1
int egsgpu_set_photon_tracking(egsgpu_handle *h, int32_t mode, double factor);

/* Seed all streams: ranmar initial seeds (ixx, jxx).  Re-seeding resets
 * the streams; tallies are kept.  Same as
 * egsgpu_set_rng(h, EGSGPU_RNG_RANMAR, ixx, jxx). */
int egsgpu_seed(egsgpu_handle *h, int32_t ixx, int32_t jxx);

/* Phase-space source (EGS_PhspSource::getNextParticle/readParticle,
 * egs++/sources/egs_phsp_source/egs_phsp_source.cpp:285-431) for a
 * simulation created with source kind EGSGPU_SOURCE_PHSP.  History i of the
 * simulation (counted over all its runs) takes particle i mod n: position
 * (x + x0, y + y0, z) with x0, y0, z of the EgsSource, direction
 * (u, v, +-sqrt(1 - u^2 - v^2)) with the sign of wt, weight |wt|, kinetic
 * energy e and charge q (-1, 0, 1).  A particle of weight 0 (rejected by a
 * filter) or one that misses the grid ends its history at once.  The
 * arrays are copied; must be called before the first run. */
int egsgpu_set_phsp(egsgpu_handle *h, int64_t n, const float *x, const float *y,
                    const float *u, const float *v, const float *e, const float *wt,
                    const int8_t *q);

/* Choose the generator that drives the transport and seed all streams
 * (any time; tallies are kept):
 *  EGSGPU_RNG_RANMAR (a = ixx, b = jxx): stream s is EGS_Ranmar seeded as
 *    createRNG(input, s) (egs++/egs_rndm.cpp:409-463), the default;
 *  EGSGPU_RNG_RANLUX (a = luxury level 0..4, out of range -> 1, b = seed):
 *    stream s is init_ranlux(luxury, seed + s) (src/ranlux.mortran), the
 *    seeding of parallel job s + 1 of DOSXYZnrc (dosxyznrc.mortran:1916),
 *    consumed in blocks of 24 as ranlux(rng_array) / $RANDOMSET do
 *    (src/ranlux.macros:86-120).
 * Every stream is consumed in order by the histories it runs. */
#define EGSGPU_RNG_RANMAR 0
#define EGSGPU_RNG_RANLUX 1
int egsgpu_set_rng(egsgpu_handle *h, int32_t kind, int32_t a, int32_t b);

/* Simulate nhist further primary histories (runSimulation,
 * egs_application.cpp:839-917).  Histories are dealt round-robin to the
 * streams; tallies accumulate history-by-history (egs_scoring.h:63-125). */
int egsgpu_run(egsgpu_handle *h, int64_t nhist);

/* egsgpu_run without allocation, for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  Runs nhist further
 * histories exactly as egsgpu_run does on the history schedules
 * (EGSGPU_SCHED_HISTORY or EGSGPU_SCHED_HISTORY_PLAIN; same histories, same
 * random numbers, same tallies up to the order of FP64 additions), with the
 * per-stream history tables and particle stacks in a static device array
 * of the library (64 MiB, held by each GPU the library runs on until the
 * process ends) instead of the handle's own.  It calls no hipMalloc,
 * hipFree, event or stream creation.  Requirements: the handle's streams
 * exist (a first egsgpu_run, e.g. of 0 histories, makes them) and fit:
 * streams x (history-table entries x 20 B + stack entries x 44 B in FP32,
 * 88 B in FP64) <= 64 MiB, otherwise EGSGPU_E_ARG; e.g. up to about 10000
 * streams in kerma mode (256 entries, 33 stack entries) and 1500 with
 * coupled transport (2048 entries, 65 stack entries).  Every call on a GPU
 * shares that GPU's array, which keeps the last call's contents; the
 * history tables are cleared at the start of each call.  Calls that
 * overlap on one GPU (two host threads, or two handles' streams) race. */
int egsgpu_run_static(egsgpu_handle *h, int64_t nhist);

/* Run a batch of simulations together (many small jobs): nhist[j] further
 * histories of simulation jobs[j], j < njobs, in one event schedule on the
 * streams and history slots of jobs[0].  The jobs share tables, precision
 * and transport parameters (checked); geometries, media maps, densities and
 * sources differ.  Each job's tallies and history count accumulate as with
 * egsgpu_run.  The reference runs such a set as separate processes
 * (egs-parallel, scripts/bin/egs-parallel); no single-call counterpart. */
int egsgpu_run_batched(egsgpu_handle *const *jobs, int32_t njobs,
                       const int64_t *nhist);

/* Per-voxel FP64 sums of energy deposited (MeV, weight included) and of
 * its per-history squares, plus the number of histories so far.  Arrays
 * of nx*ny*nz, caller-owned. */
int egsgpu_get_tallies(egsgpu_handle *h, double *sum, double *sum2,
                       int64_t *nhist);
int egsgpu_reset_tallies(egsgpu_handle *h);

/* Set the tallies and the history count (arrays of nx*ny*nz): the
 * restart of a run from a state file that holds the per-region sums of a
 * scoring array (EGS_ScoringArray::setState, egs++/egs_scoring.h:340-365,
 * read by EGS_Application::readData, egs_application.cpp:520-556).  The
 * streams are left as they are; seed them apart from the run that made the
 * sums. */
int egsgpu_set_tallies(egsgpu_handle *h, const double *sum, const double *sum2,
                       int64_t nhist);

/* Checkpoint and restart (EGS_Application::outputData / readData,
 * egs_application.cpp:482-560, which store the run control, RNG, source
 * and scoring states): the tallies, the number of histories and the state
 * of every RNG stream, in an engine-private binary layout.  With
 * buf == NULL, egsgpu_checkpoint only returns the size needed.  A restore
 * needs a simulation of the same number of voxels and streams; continuing
 * after a restore gives the same random numbers as continuing the
 * checkpointed run. */
int egsgpu_checkpoint(egsgpu_handle *h, void *buf, size_t *size);
int egsgpu_restore(egsgpu_handle *h, const void *buf, size_t size);

/* Density gradient (photon kerma mode): runs nhist histories forward from
 * the current stream states to get the dose D_v = E_v/(N m_v), then replays
 * the same histories from the saved RNG states and accumulates the
 * score-function (likelihood-ratio) gradient of
 *   L = sum_v (D_v - target_v)^2
 * with respect to the relative voxel densities (FP64 tallies).  Writes
 * dL/drho to grad (nvox), the dose to dose (nvox, may be NULL) and L to
 * loss.  The explicit 1/m_v dependence of the dose is included.  Needs
 * the history schedule and fewer than 2^25 voxels.  No counterpart in the
 * reference (egs_chamber's correlated sampling with
 * RNG save/restore, user_codes/egs_chamber/egs_chamber.cpp:2424-2446, is
 * the precedent for the replay). */
int egsgpu_density_grad(egsgpu_handle *h, int64_t nhist, const double *target,
                        double *grad, double *dose, double *loss);

/* Energy bookkeeping since the last reset (history-by-history FP64 sums,
 * weight included): out[0], out[1] = sum and sum of per-history squares of
 * the energy brought into the grid by the primaries; out[2], out[3] = the
 * same for the energy carried out of the grid by escaping particles.  A
 * particle's energy is its kinetic energy, plus 2 m c^2 for a positron (the
 * energy of its annihilation).  Without variance reduction, incident =
 * deposited (the sum of egsgpu_get_tallies) + escaped, per history, up to
 * the rounding of the transport arithmetic: the energy conservation of
 * the shower (EGSnrc passes every particle's energy to ausgab when it is
 * deposited or discarded, egsnrc.mortran:1800-1816, 6600-6643).  Restored
 * with a checkpoint; restarted by egsgpu_set_tallies. */
int egsgpu_get_energy_balance(egsgpu_handle *h, double out[4]);

/* Engine counters: [0] photon steps, [1] voxel crossings, [2] deposits,
 * [3] history flushes (voxel,history pairs), [4] electron steps,
 * [5] stack overflows, [6] tally-table overflows, [7] primaries that
 * missed the grid and were sampled again (DOSXYZnrc isource 1, nmissm). */
int egsgpu_get_counters(egsgpu_handle *h, int64_t counters[8]);

/* The same counters split by the kernel that counted them (kernel numbers
 * as for egsgpu_kernel_time); egsgpu_get_counters is their sum. */
int egsgpu_kernel_counters(egsgpu_handle *h, int32_t kernel,
                           int64_t counters[8]);

/* Device time (ms) of the last egsgpu_run, from HIP events on the launch
 * stream, and the number of transport kernel launches. */
int egsgpu_last_kernel_ms(egsgpu_handle *h, double *ms, int32_t *launches);

/* Time of one kernel of the last egsgpu_run: kernel 0 = history-schedule
 * transport, 1 = event source, 2 = event photon, 3 = event charged
 * particle.  Sum of its launch durations (ms, HIP events around each
 * launch) and the number of launches. */
int egsgpu_kernel_time(egsgpu_handle *h, int32_t kernel, double *ms,
                       int32_t *launches);

/* ---- test hooks -------------------------------------------------------- */
/* Fill out[s*n + i] with the first n numbers of stream s of ranmar seeded
 * (ixx, jxx + s), s < nstreams, computed on the device
 * (EGS_Ranmar::setState/fillArray, egs_rndm.cpp:303-398), drawn through
 * the per-lane ring the transport uses, with its top-ups and its path for
 * a ring that ran dry. */
int egsgpu_ranmar_fill(int32_t ixx, int32_t jxx, int32_t nstreams,
                       int64_t n, double *out);

/* Same for ranlux (ranlux.mortran) at luxury level lux with stream s seeded
 * with seed + s. */
int egsgpu_ranlux_fill(int32_t lux, int32_t seed, int32_t nstreams,
                       int64_t n, double *out);

/* Trace rays through the XYZ geometry on the device: starting at x, u
 * (3 doubles each per ray) find the region (isWhere, or howfar from
 * outside), then step with howfar(t = 1e30) until the ray leaves.  For ray
 * r, regions[r*max_steps+k] is the k-th region entered (-1 = left) and
 * steps[r*max_steps+k] the step length to it; nsteps[r] the number of
 * entries.  precision is 32 or 64 (the FP32 transport build or the FP64
 * build of the same code). */
int egsgpu_trace_rays(const EgsXYZGeometry *g, int32_t nrays,
                      const double *x, const double *u, int32_t max_steps,
                      int32_t precision, int32_t *regions, double *steps,
                      int32_t *nsteps);

/* Measurement hook, no counterpart in the reference: the rate of FP64
 * atomic adds on this device.  Two launches of about nops adds of 1.0 each
 * into a table of naddr doubles (the second is timed: *ms); pattern 0
 * scatters the adds over the table by a hash (the tallies' pattern),
 * 1 sends the 64 lanes of a wave to one address, 2 gives them 64
 * consecutive addresses.  *check_sum = the table's sum after both
 * launches (2 x the adds issued). */
int egsgpu_atomic_rate(int32_t pattern, int32_t naddr, int64_t nops, double *ms,
                       double *check_sum);

/* hownear for nrays points with known regions (egs_nd_geometry.h:1106). */
int egsgpu_hownear(const EgsXYZGeometry *g, int32_t n, const double *x,
                   const int32_t *ireg, int32_t precision, double *tperp);

#ifdef __cplusplus
}
#endif
#endif /* EGSGPU_H */
