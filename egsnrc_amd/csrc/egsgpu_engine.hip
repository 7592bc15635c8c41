// egsgpu_engine.hip -- the transport engine behind egsgpu_create/run:
// device residency of tables, voxel grid and source, per-lane state, the
// transport kernel launch and the FP64 tallies.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cctype>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/egsgpu.h"
#include "egs_tables.h"
#include "egsgpu_common.h"
#include "egsgpu_device.h"
#include "egsgpu_engine.h"
#include "egsgpu_transport.h"
#include "egsgpu_kernels.h"

namespace egsgpu {

// ---------------------------------------------------------------- kernels
__global__ void k_seed(RanmarRec *recs, int nlanes, int ixx, int jxx) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nlanes) return;
    int is, js;
    stream_seeds(ixx, jxx, s, is, js);
    ranmar_seed(recs + s, is, js);
}

// ranlux stream s: init_ranlux(luxury, seed + s), the seed of parallel job
// s + 1 of DOSXYZnrc (dosxyznrc.mortran:1916, ranlux.macros:96-99)
__global__ void k_seed_lux(RanluxRec *recs, int nlanes, int lux, int seed) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nlanes) return;
    ranlux_seed(recs + s, lux, seed + s);
}

// The transport kernels (templates in egsgpu_kernels.h) are compiled in the
// egsgpu_kern_*.hip files, one group per file, so that the groups build in
// parallel; here they are only declared.
EGSGPU_ALL_KERNELS(EGSGPU_DECLARE, float)
EGSGPU_ALL_KERNELS(EGSGPU_DECLARE, double)
// Bank bookkeeping between kernels: bank b is drained.
__global__ void k_ev_ctl(EvCtl *c, int b) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        c->n[b] = 0;
        c->head[b] = 0;
    }
}

// Plan the next source launch of a graph-driven run: as many new histories
// as there are free slots, histories left and room in the bank they go to.
// Once no history is left to start and every bank is empty the run has
// ended: the plan sets `done`, and the kernels of the iterations still
// queued in the graph return at once.
__global__ void k_ev_plan(EvCtl *c, int bank, int cap) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (c->done) { c->plan_n = 0; return; }
    if (c->remaining == 0 && c->n[0] == 0 && c->n[1] == 0 && c->n[2] == 0 && c->n[3] == 0) {
        c->done = 1;
        c->plan_n = 0;
        return;
    }
    c->iters++;
    unsigned long long fr = c->free_tail - c->free_head;
    long long inj = (long long)fr < c->remaining ? (long long)fr : c->remaining;
    long long room = (long long)cap - c->n[bank];
    inj = inj < room ? inj : room;
    inj = inj > 0 ? inj : 0;
    bool empty = c->n[0] == 0 && c->n[1] == 0 && c->n[2] == 0 && c->n[3] == 0;
    if (inj == 0 && empty && c->remaining > 0) { c->stall = 1; c->done = 1; }
    c->plan_n = (int)inj;
    c->plan_head = c->free_head;
    c->plan_base = c->hist_base;
    c->free_head += (unsigned long long)inj;
    c->remaining -= inj;
    c->hist_base += inj;
}

__global__ void k_ev_setrun(EvCtl *c, long long nhist, unsigned long long free_head) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    c->remaining = nhist;
    c->hist_base = 0;
    c->free_head = free_head;
    c->stall = 0;
    c->done = 0;
    c->iters = 0;
}

__global__ void k_ev_init(DevSlots S, EvCtl *c) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < S.nslots) {
        S.ring[i] = i;
        S.used[i] = 0;
        S.pending[i] = 0;
    }
    if (i == 0) {
        c->free_tail = (unsigned long long)S.nslots;
        for (int k = 0; k < 4; k++) c->n[k] = c->head[k] = 0;
    }
}

// ---- the library's static scratch (egsgpu_run_static) -------------------
// Device memory of a fixed size that the library never allocates: each GPU
// the library runs on holds one copy until the process ends.  Host code
// never uses the array's name as an address (on the host it is only a
// stand-in); it takes the device address once per device with
// hipGetSymbolAddress and passes that.
constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Zero n 16-byte words (the history-table keys must start at zero); a
// kernel rather than a memset so that callers may capture it in a graph.
__global__ void k_zero16(uint4 *p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = make_uint4(0u, 0u, 0u, 0u);
}

// ------------------------------------------------------------ host side
struct DevBuf {
    void *p = nullptr;
    size_t n = 0;
};

template <typename R>
struct Precision {
    DevTables<R> T;
    GeomView<R> G;
    const VoxData<R> *vox = nullptr;
    DevSource<R> S;
    DevLanes<R> L;
    DevBank<R> bank[4]{};    // event-schedule banks: 0, 1 photons; 2, 3 charged
};

}  // namespace egsgpu

using namespace egsgpu;

// tally entries past the voxels: the energy-bookkeeping regions (escaped,
// incident), kExtraCopies entries each
constexpr int kExtraRegions = 2 * kExtraCopies;

constexpr int kMaxGraphIters = 512;   // iterations one run graph may hold

struct egsgpu_handle {
    int fp64 = 0;
    int nvox = 0;
    std::vector<void *> allocs;      // everything to hipFree
    Precision<float> pf;
    Precision<double> pd;
    DevParams P{};
    DevTally tally{};
    RanmarRec *rng = nullptr;
    RanluxRec *rlx = nullptr;
    int rng_kind = EGSGPU_RNG_RANMAR;   // which streams drive the transport
    int lux = 1, lux_seed = 1;          // ranlux luxury level and first seed
    int nlanes = 0, block = 256, mxstk = 48, hcap = 256;
    int ncu = 1;                        // compute units of the device
    int ev_wp = 2, ev_we = 2;           // waves/SIMD variants of the event kernels
    int ev_ksteps = 8, ev_batch = 16, ev_max_wait = 16, ev_defer = 32;
    int pc = 0, ec = 2;                 // banks to drain next (photon, charged)
    int *slot_job = nullptr;            // [nslots] job of each slot (batched runs)
    bool use_graph = true;              // event schedule: replay captured iterations
    int ev_graph_iters = 16;            // iterations captured in the run's graph
    int graph_launches = 0;             // graph launches of the last run
    int graph_iters_used = 0;           // iterations with work in the last run
    void *job_self = nullptr;           // device JobDesc of this simulation alone
    const egsgpu_tables *tab = nullptr; // the tables the simulation was made with
    int ixx = 1802, jxx = 9373;
    bool lanes_ready = false;
    int64_t nhist = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0;
    int last_launches = 0;
    // host copies for dose normalisation and the gradient
    std::vector<double> hx, hy, hz, hrhor, hmass;
    std::vector<double> rho_max;        // largest relative density of each medium (0: absent)
    double woodcock_factor = 0;         // majorant over the largest attenuation (0: off)
    RanmarRec *rng_saved = nullptr;
    RanluxRec *rlx_saved = nullptr;
    DevGrad dgrad{};
    DevGrad *dgrad_dev = nullptr;
    // event schedule
    int schedule = EGSGPU_SCHED_EVENT;
    bool howfarless = false;            // DOSXYZnrc HOWFARLESS (homogeneous phantom)
    bool ev_ready = false;
    size_t ev_alloc_first = 0, ev_alloc_end = 0;   // the event state's entries of allocs
    int ev_block = 0;                   // event kernels' workgroup size (0: block)
    int nslots = 0, slot_hcap = 256;
    DevSlots slots{};
    EvCtl *ctl = nullptr;
    EvCtl *ctl_host = nullptr;          // pinned
    unsigned long long free_head = 0;
    hipEvent_t kev[4][2] = {};          // per-kernel timing events
    double kms[4] = {};
    int klaunch[4] = {};
    ~egsgpu_handle() {
        for (void *p : allocs) (void)hipFree(p);
        if (ctl_host) (void)hipHostFree(ctl_host);
        for (auto &e : kev)
            for (hipEvent_t x : e)
                if (x) (void)hipEventDestroy(x);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

template <typename T>
int upload(egsgpu_handle *h, const T *src, size_t n, const T **dst) {
    void *p = nullptr;
    size_t bytes = std::max<size_t>(n * sizeof(T), 16);
    EGS_HIP(hipMalloc(&p, bytes));
    h->allocs.push_back(p);
    if (n) EGS_HIP(hipMemcpy(p, src, n * sizeof(T), hipMemcpyHostToDevice));
    *dst = (const T *)p;
    return EGSGPU_OK;
}

template <typename T>
int alloc(egsgpu_handle *h, size_t n, T **dst) {
    void *p = nullptr;
    EGS_HIP(hipMalloc(&p, std::max<size_t>(n * sizeof(T), 16)));
    h->allocs.push_back(p);
    *dst = (T *)p;
    return EGSGPU_OK;
}

#define CK(x) do { int rc_ = (x); if (rc_) return rc_; } while (0)

// Pack medium arrays [nmed][stride] in precision R.
template <typename R>
int pack_med(egsgpu_handle *h, const Blob &b, int nmed, const char *field,
             int stride, const R **dst) {
    std::vector<R> v((size_t)nmed * stride, (R)0);
    for (int m = 0; m < nmed; m++) {
        const auto &a = b.F(Blob::mk(m, field));
        if ((int)a.size() > stride) { set_error("table %s too long", field); return EGSGPU_E_FORMAT; }
        for (size_t i = 0; i < a.size(); i++) v[(size_t)m * stride + i] = (R)a[i];
    }
    return upload(h, v.data(), v.size(), dst);
}
// Interleave per-medium tables into rows: dst[((m*nrows + i)*width + off].
template <typename R>
int pack_rows(egsgpu_handle *h, const Blob &b, int nmed, int nrows, int width,
              const std::vector<std::pair<const char *, int>> &fields, const R **dst) {
    std::vector<R> v((size_t)nmed * nrows * width, (R)0);
    for (int m = 0; m < nmed; m++) {
        for (const auto &f : fields) {
            const auto &a = b.F(Blob::mk(m, f.first));
            if ((int)a.size() > nrows) { set_error("table %s too long", f.first); return EGSGPU_E_FORMAT; }
            for (size_t i = 0; i < a.size(); i++)
                v[((size_t)m * nrows + i) * width + f.second] = (R)a[i];
        }
    }
    return upload(h, v.data(), v.size(), dst);
}
int pack_med_i(egsgpu_handle *h, const Blob &b, int nmed, const char *field,
               int stride, const int **dst) {
    std::vector<int> v((size_t)nmed * stride, 0);
    for (int m = 0; m < nmed; m++) {
        const auto &a = b.I(Blob::mk(m, field));
        if ((int)a.size() > stride) { set_error("table %s too long", field); return EGSGPU_E_FORMAT; }
        for (size_t i = 0; i < a.size(); i++) v[(size_t)m * stride + i] = a[i];
    }
    return upload(h, v.data(), v.size(), dst);
}
int pack_med_d(egsgpu_handle *h, const Blob &b, int nmed, const char *field,
               int stride, const double **dst) {
    std::vector<double> v((size_t)nmed * stride, 0.0);
    for (int m = 0; m < nmed; m++) {
        if (!b.has(Blob::mk(m, field))) continue;
        const auto &a = b.F(Blob::mk(m, field));
        for (size_t i = 0; i < a.size() && (int)i < stride; i++) v[(size_t)m * stride + i] = a[i];
    }
    return upload(h, v.data(), v.size(), dst);
}
template <typename R>
int pack_glob(egsgpu_handle *h, const Blob &b, const char *field, const R **dst) {
    const auto &a = b.F(field);
    std::vector<R> v(a.begin(), a.end());
    return upload(h, v.data(), v.size(), dst);
}

template <typename R>
int build_tables(egsgpu_handle *h, const Blob &b, DevTables<R> &T) {
    int nmed = b.i1("nmed");
    T.nmed = nmed;
    std::vector<MedScalars<R>> ms(nmed);
    for (int m = 0; m < nmed; m++) {
        auto f = [&](const char *k) { return (R)b.f1(Blob::mk(m, k)); };
        MedScalars<R> &s = ms[m];
        s.ge0 = f("ge0"); s.ge1 = f("ge1"); s.eke0 = f("eke0"); s.eke1 = f("eke1");
        s.rho = f("rho"); s.ae = f("ae"); s.ap = f("ap"); s.te = f("te"); s.thmoll = f("thmoll");
        s.delcm = f("delcm");
        s.bpar0 = (R)b.F(Blob::mk(m, "bpar"))[0];
        s.bpar1 = (R)b.F(Blob::mk(m, "bpar"))[1];
        s.zbrang = f("zbrang"); s.lzbrang = f("lzbrang"); s.log_ap = f("log_ap");
        s.blcc = f("blcc"); s.xcc = f("xcc"); s.expeke1 = f("expeke1");
        s.esig_e = f("esig_e"); s.psig_e = f("psig_e");
        const bool nist = b.has("ibr_nist") && b.i1("ibr_nist") >= 1;
        s.nb_emin = nist ? f("nb_emin") : (R)0;
        s.nb_lemin = nist ? f("nb_lemin") : (R)0;
        s.nb_dlei = nist ? f("nb_dlei") : (R)0;
        const bool eii = b.has("eii") && b.i1("eii") == 1;
        s.eii_nsh = eii ? b.i1(Blob::mk(m, "eii_nsh")) : 0;
        s.eii_cons = eii ? f("eii_cons") : (R)0;
        s.nne = b.i1(Blob::mk(m, "nne"));
        s.n_shell = b.i1(Blob::mk(m, "n_shell"));
        s.meke = b.i1(Blob::mk(m, "meke"));
        s.sig_mono_e = b.I(Blob::mk(m, "sig_monotone"))[0];
        s.sig_mono_p = b.I(Blob::mk(m, "sig_monotone"))[1];
        if (s.nne > kMaxNE || s.n_shell > kMaxShMed || s.meke > kMaxEKE ||
            b.i1(Blob::mk(m, "mge")) != kGE) {
            set_error("medium %d exceeds the engine's table dimensions", m);
            return EGSGPU_E_UNSUPPORTED;
        }
    }
    CK(upload(h, ms.data(), ms.size(), &T.ms));
    {
        // Klein-Nishina photon rows when the run has bound Compton off and
        // the blob's main rows include the bound-Compton cross section
        const bool kn = h->P.ibcmp <= 1 && b.i1("ibcmp") > 1;
        const std::vector<std::pair<const char *, int>> gf = kn ?
            std::vector<std::pair<const char *, int>>{
                {"gmfp0_kn", kGmfp}, {"gmfp1_kn", kGmfp + 1}, {"gbr10_kn", kGbr1},
                {"gbr11_kn", kGbr1 + 1}, {"gbr20_kn", kGbr2}, {"gbr21_kn", kGbr2 + 1},
                {"cohe0_kn", kCohe}, {"cohe1_kn", kCohe + 1}} :
            std::vector<std::pair<const char *, int>>{
                {"gmfp0", kGmfp}, {"gmfp1", kGmfp + 1}, {"gbr10", kGbr1}, {"gbr11", kGbr1 + 1},
                {"gbr20", kGbr2}, {"gbr21", kGbr2 + 1}, {"cohe0", kCohe}, {"cohe1", kCohe + 1}};
        CK(pack_rows<R>(h, b, nmed, kGE, kGRow, gf, &T.grow));
    }
    bool ray = b.has(Blob::mk(0, "ray_x"));
    if (ray) {
        CK(pack_med<R>(h, b, nmed, "ray_pmax0", kGE, &T.pmax0));
        CK(pack_med<R>(h, b, nmed, "ray_pmax1", kGE, &T.pmax1));
        CK(pack_med_i(h, b, nmed, "ray_i", 100, &T.ray_i));
    } else {
        T.pmax0 = T.pmax1 = nullptr;
        T.ray_i = nullptr;
    }
    CK(pack_med_d(h, b, nmed, "ray_x", 100, &T.ray_x));
    CK(pack_med_d(h, b, nmed, "ray_fcum", 100, &T.ray_fcum));
    CK(pack_med_d(h, b, nmed, "ray_b", 100, &T.ray_b));
    CK(pack_med_d(h, b, nmed, "ray_c", 100, &T.ray_c));
    CK(pack_med<R>(h, b, nmed, "eno", kMaxShMed, &T.eno));
    CK(pack_med_i(h, b, nmed, "eno_atbin", kMaxShMed, &T.eno_atbin));
    CK(pack_med_i(h, b, nmed, "shell", kMaxShMed, &T.shell));
    CK(pack_glob<R>(h, b, "cs_be", &T.cs_be));
    CK(pack_glob<R>(h, b, "cs_Jo", &T.cs_Jo));
    CK(pack_glob<R>(h, b, "cs_erfJo", &T.cs_erfJo));
    {
        const auto &iz = b.I("cs_iz"), &shn = b.I("cs_shn");
        std::vector<int> a(iz.begin(), iz.end()), c(shn.begin(), shn.end());
        CK(upload(h, a.data(), a.size(), &T.cs_iz));
        CK(upload(h, c.data(), c.size(), &T.cs_shn));
    }
    CK(pack_med<R>(h, b, nmed, "z", kMaxNE, &T.z));
    CK(pack_med<R>(h, b, nmed, "pz", kMaxNE, &T.pz));
    CK(pack_glob<R>(h, b, "binding_energies", &T.bind));
    CK(pack_glob<R>(h, b, "interaction_prob", &T.iprob));
    CK(pack_glob<R>(h, b, "edge_a", &T.ea));
    CK(pack_glob<R>(h, b, "edge_b", &T.eb));
    CK(pack_glob<R>(h, b, "edge_c", &T.ec));
    CK(pack_glob<R>(h, b, "edge_d", &T.ed));
    CK(pack_glob<R>(h, b, "edge_energies", &T.ee));
    {
        const auto &en = b.I("edge_number");
        std::vector<int> v(en.begin(), en.end());
        CK(upload(h, v.data(), v.size(), &T.enumb));
    }
    if (b.has("relax_zfirst")) {
        auto up_i = [&](const char *k, const int **dst) {
            const auto &a = b.I(k);
            std::vector<int> v(a.begin(), a.end());
            return upload(h, v.data(), v.size(), dst);
        };
        CK(up_i("relax_zfirst", &T.relax_zfirst));
        CK(up_i("relax_nshell", &T.relax_nshell));
        CK(up_i("shell_num", &T.shell_num));
        CK(up_i("relax_first", &T.relax_first));
        CK(up_i("relax_ntran", &T.relax_ntran));
        CK(up_i("relax_atbin", &T.relax_atbin));
        CK(up_i("relax_state", &T.relax_state));
        CK(pack_glob<R>(h, b, "shell_be", &T.shell_be));
        CK(pack_glob<R>(h, b, "relax_prob", &T.relax_prob));
    } else {   // tables made without relaxation data: vacancies absorbed on the spot
        T.relax_zfirst = T.relax_nshell = T.shell_num = T.relax_first = T.relax_ntran =
            T.relax_atbin = T.relax_state = nullptr;
        T.shell_be = T.relax_prob = nullptr;
    }
    CK(pack_med<R>(h, b, nmed, "dl", 48, &T.dl));
    static_assert(kNbGridE == kNbE && kNbGridX == kNbX, "brems sampling grid");
    if (b.has("ibr_nist") && b.i1("ibr_nist") >= 1) {
        CK(pack_med<R>(h, b, nmed, "nb_xdata", kNbGridE * (kNbGridX + 1), &T.nb_x));
        CK(pack_med<R>(h, b, nmed, "nb_fdata", kNbGridE * (kNbGridX + 1), &T.nb_f));
        CK(pack_med<R>(h, b, nmed, "nb_wdata", kNbGridE * kNbGridX, &T.nb_w));
        CK(pack_med_i(h, b, nmed, "nb_idata", kNbGridE * kNbGridX, &T.nb_i));
    } else {
        T.nb_x = T.nb_f = T.nb_w = nullptr;
        T.nb_i = nullptr;
    }
    static_assert(kEiiGridBins == kEiiBins, "EII grid");
    if (b.has("eii") && b.i1("eii") == 1) {
        CK(pack_med_i(h, b, nmed, "eii_no", kMaxNE, &T.eii_no));
        CK(pack_med_i(h, b, nmed, "eii_first", kMaxNE, &T.eii_first));
        CK(pack_glob<R>(h, b, "eii_a", &T.eii_a));
        CK(pack_glob<R>(h, b, "eii_b", &T.eii_b));
        CK(pack_glob<R>(h, b, "eii_xa", &T.eii_xa));
        CK(pack_glob<R>(h, b, "eii_xb", &T.eii_xb));
    } else {
        T.eii_no = T.eii_first = nullptr;
        T.eii_a = T.eii_b = T.eii_xa = T.eii_xb = nullptr;
    }
    static_assert(kNrcpGridE == kNrcpE && kNrcpGridX == kNrcpX, "NRC pair grid");
    if (b.has("pair_nrc") && b.i1("pair_nrc") == 1) {
        CK(pack_glob<R>(h, b, "nrcp_xdata", &T.nrcp_x));
        CK(pack_med<R>(h, b, nmed, "nrcp_fdata", kNrcpGridE * kNrcpGridX, &T.nrcp_f));
        CK(pack_med<R>(h, b, nmed, "nrcp_wdata", kNrcpGridE * (kNrcpGridX - 1), &T.nrcp_w));
        CK(pack_med_i(h, b, nmed, "nrcp_idata", kNrcpGridE * (kNrcpGridX - 1), &T.nrcp_i));
        T.nrcp_emin = (R)b.F("nrcp_const")[0];
        T.nrcp_emax = (R)b.F("nrcp_const")[1];
        T.nrcp_dlei = (R)b.F("nrcp_const")[2];
    } else {
        T.nrcp_x = T.nrcp_f = T.nrcp_w = nullptr;
        T.nrcp_i = nullptr;
        T.nrcp_emin = T.nrcp_emax = T.nrcp_dlei = 0;
    }
    {
        const std::vector<std::pair<const char *, int>> ef = {
            {"esig0", kESig}, {"esig1", kESig + 1}, {"psig0", kPSig}, {"psig1", kPSig + 1},
            {"ededx0", kEDedx}, {"ededx1", kEDedx + 1}, {"pdedx0", kPDedx},
            {"pdedx1", kPDedx + 1}, {"tmxs0", kTmxs}, {"tmxs1", kTmxs + 1},
            {"range_ep0", kRepE}, {"range_ep1", kRepP}, {"e_array", kEArr},
            {"ebr10", kEBr1}, {"ebr11", kEBr1 + 1}, {"pbr10", kPBr1}, {"pbr11", kPBr1 + 1},
            {"pbr20", kPBr2}, {"pbr21", kPBr2 + 1}};
        CK(pack_rows<R>(h, b, nmed, kMaxEKE, kERow, ef, &T.erow));
    }
    {
        const char *mf[] = {"ms_u", "ms_f", "ms_w"};
        const float **md[] = {&T.ms_u, &T.ms_f, &T.ms_w};
        for (int k = 0; k < 3; k++) {
            const auto &a = b.F(mf[k]);
            std::vector<float> v(a.begin(), a.end());
            CK(upload(h, v.data(), v.size(), md[k]));
        }
        const auto &mi = b.I("ms_i");
        std::vector<int> v(mi.begin(), mi.end());
        CK(upload(h, v.data(), v.size(), &T.ms_i));
        T.ms_dllambi = (float)b.F("ms_const")[3];
        T.ms_dqmsi = (float)b.F("ms_const")[5];
    }
    return EGSGPU_OK;
}

template <typename R>
int build_geom(egsgpu_handle *h, const EgsXYZGeometry *g, const std::vector<uint8_t> &med,
               Precision<R> &P) {
    int n = (g->nx + 1) + (g->ny + 1) + (g->nz + 1);
    std::vector<R> pl(n);
    int k = 0;
    for (int i = 0; i <= g->nx; i++) pl[k++] = (R)g->xb[i];
    for (int i = 0; i <= g->ny; i++) pl[k++] = (R)g->yb[i];
    for (int i = 0; i <= g->nz; i++) pl[k++] = (R)g->zb[i];
    const R *dpl;
    CK(upload(h, pl.data(), pl.size(), &dpl));
    GeomView<R> &v = P.G;
    v.nx = g->nx; v.ny = g->ny; v.nz = g->nz; v.nxy = g->nx * g->ny;
    v.xp = dpl; v.yp = dpl + g->nx + 1; v.zp = dpl + g->nx + 1 + g->ny + 1;
    double tol = g->boundary_tolerance;
    if (tol <= 0) tol = sizeof(R) == 4 ? 1.0 / (1 << 21) : 1.0 / 4294967296.0;
    v.btol = (R)tol;
    v.hbtol = v.btol / (R)2;
    // equally spaced planes: the inverse spacing, for the voxel lookup of
    // Woodcock tracking (0 when the spacing varies by more than 1e-9)
    auto inv_spacing = [](const double *b, int n) -> R {
        const double d = (b[n] - b[0]) / n;
        for (int i = 0; i < n; i++)
            if (std::fabs((b[i + 1] - b[i]) - d) > 1e-9 * std::fabs(d)) return (R)0;
        return (R)(1.0 / d);
    };
    v.ixs = inv_spacing(g->xb, g->nx);
    v.iys = inv_spacing(g->yb, g->ny);
    v.izs = inv_spacing(g->zb, g->nz);
    std::vector<VoxData<R>> vox((size_t)h->nvox);
    for (int i = 0; i < h->nvox; i++) {
        vox[i].rho = g->rhor ? (R)g->rhor[i] : (R)1;
        vox[i].med = med[i];
    }
    CK(upload(h, vox.data(), vox.size(), &P.vox));
    return EGSGPU_OK;
}

template <typename R>
int build_source(egsgpu_handle *h, const EgsSource *s, Precision<R> &P) {
    DevSource<R> &S = P.S;
    S.charge = s->charge; S.spectrum = s->spectrum; S.shape = s->shape;
    S.energy = (R)s->energy;
    S.x0 = (R)s->x0; S.x1 = (R)s->x1; S.y0 = (R)s->y0; S.y1 = (R)s->y1; S.z = (R)s->z;
    S.ux = (R)s->ux; S.uy = (R)s->uy; S.uz = (R)s->uz;
    S.np = 0; S.type = 0;
    S.xi = S.wi = nullptr;
    S.bin = nullptr;
    S.kind = s->kind;
    S.nphsp = 0;
    S.px = S.py = S.pu = S.pv = S.pe = S.pw = nullptr;
    S.pq = nullptr;
    if (s->kind == EGSGPU_SOURCE_PHSP) S.charge = 0;   // primaries enter the photon bank
    for (int k = 0; k < 9; k++) S.rot[k] = s->rot[k];
    for (int k = 0; k < 3; k++) S.iso[k] = s->iso[k];
    S.xcol = s->xcol; S.ycol = s->ycol; S.dsource = s->dsource;
    S.z1 = (R)s->z1; S.rho_max = (R)s->rho_max;
    if (s->spectrum == EGSGPU_SPEC_MONO) return EGSGPU_OK;
    if (s->spectrum == EGSGPU_SPEC_CDFINV) {
        S.np = s->nbins;
        CK(upload(h, s->spec_e, (size_t)s->nbins, &S.xi));
        CK(upload(h, s->spec_w, (size_t)s->nbins, &S.wi));
        return EGSGPU_OK;
    }
    // EGS_AliasTable::make (egs_alias_table.cpp:102-171)
    int type = s->spectrum == EGSGPU_SPEC_HISTOGRAM ? 1 : 0;
    int n = type ? s->nbins + 1 : s->nbins;
    int np = type ? n - 1 : n;
    std::vector<double> xi(s->spec_e, s->spec_e + n), fi(s->spec_w, s->spec_w + np), wi(np, 1.0), fcum(np);
    std::vector<int> bin(np, 0);
    std::vector<bool> not_done(np, true);
    double sum = 0;
    for (int i = 0; i < np; i++) {
        fcum[i] = type == 0 ? fi[i] : fi[i] * (xi[i + 1] - xi[i]);
        sum += fcum[i];
    }
    sum /= np;
    for (int i = 0; i < np - 1; i++) {
        int jh, jl, high = -1, low = -1;
        for (jh = 0; jh < np; jh++) if (not_done[jh] && fcum[jh] > sum) { high = jh; break; }
        if (high < 0) break;
        for (jl = 0; jl < np; jl++) if (not_done[jl] && fcum[jl] < sum) { low = jl; break; }
        if (low < 0) break;
        fcum[high] -= sum - fcum[low];
        not_done[jl] = false;
        wi[low] = fcum[low] / sum;
        bin[low] = high;
    }
    S.np = np; S.type = type;
    CK(upload(h, xi.data(), xi.size(), &S.xi));
    CK(upload(h, wi.data(), wi.size(), &S.wi));
    CK(upload(h, bin.data(), bin.size(), &S.bin));
    return EGSGPU_OK;
}

template <typename R>
int build_lanes(egsgpu_handle *h, Precision<R> &P) {
    DevLanes<R> &L = P.L;
    L.nlanes = h->nlanes;
    L.rng = h->rng;
    L.rlx = h->rlx;
    L.mxstk = h->mxstk;
    // photon splitting keeps the split photon's record past the stack
    const int nstk = h->mxstk + 1;
    if ((double)nstk * kStkF * h->nlanes * sizeof(R) >= 4294967296.0) {
        set_error("particle stacks of %d streams exceed 4 GB", h->nlanes);
        return EGSGPU_E_ARG;
    }
    CK(alloc(h, (size_t)nstk * kStkF * h->nlanes, &L.stk));
    // per-lane history tables: history schedule only
    L.hcap = h->schedule != EGSGPU_SCHED_EVENT ? h->hcap : 0;
    size_t nh = (size_t)L.hcap * h->nlanes;
    CK(alloc(h, nh, &L.hent));
    CK(alloc(h, nh, &L.hlist));
    if (nh) EGS_HIP(hipMemset(L.hent, 0, nh * sizeof(SlotEntry)));
    return EGSGPU_OK;
}

// (re)seed every stream of the kind in use
int seed_streams(egsgpu_handle *h) {
    dim3 grid((h->nlanes + 255) / 256);
    if (h->rng_kind == EGSGPU_RNG_RANLUX)
        hipLaunchKernelGGL(k_seed_lux, grid, dim3(256), 0, h->stream, h->rlx, h->nlanes,
                           h->lux, h->lux_seed);
    else
        hipLaunchKernelGGL(k_seed, grid, dim3(256), 0, h->stream, h->rng, h->nlanes,
                           h->ixx, h->jxx);
    EGS_HIP(hipGetLastError());
    EGS_HIP(hipStreamSynchronize(h->stream));
    return EGSGPU_OK;
}

// the stream records in use and their size
void *rng_records(egsgpu_handle *h, size_t *bytes) {
    bool lux = h->rng_kind == EGSGPU_RNG_RANLUX;
    *bytes = (size_t)h->nlanes * (lux ? sizeof(RanluxRec) : sizeof(RanmarRec));
    return lux ? (void *)h->rlx : (void *)h->rng;
}

int ensure_lanes(egsgpu_handle *h) {
    if (h->lanes_ready) return EGSGPU_OK;
    // default streams: what the kernels can keep resident (event kernels
    // up to 4 waves per SIMD, the history kernel 2)
    if (!h->nlanes) {
        h->nlanes = h->ncu * (h->schedule == EGSGPU_SCHED_EVENT ? 1024
                              : h->P.kerma && h->schedule == EGSGPU_SCHED_HISTORY
                                    ? 256 * kKermaWaves : 512);
        // the deep stacks of photon splitting: fewer streams, stacks < 4 GB
        const size_t per = (size_t)(h->mxstk + 1) * kStkF * (h->fp64 ? 8 : 4);
        const int cap = (int)std::min<size_t>(4294967295ull / per / 256 * 256, 1 << 30);
        if (h->nlanes > cap) h->nlanes = std::max(cap, 256);
    }
    CK(alloc(h, (size_t)h->nlanes, &h->rng));
    CK(alloc(h, (size_t)h->nlanes, &h->rlx));
    if (h->fp64) CK(build_lanes(h, h->pd));
    else CK(build_lanes(h, h->pf));
    CK(seed_streams(h));
    h->lanes_ready = true;
    return EGSGPU_OK;
}

template <typename R>
KArgs<R> kargs(egsgpu_handle *h, Precision<R> &P) {
    KArgs<R> a;
    a.T = P.T; a.G = P.G; a.vox = P.vox; a.P = h->P;
    a.L = P.L; a.tally = h->tally; a.S = P.S;
    a.base = 0;
    a.rem = 0;
    a.grad = nullptr;
    a.hist0 = h->nhist;
    return a;
}

int timing_events(egsgpu_handle *h) {
    for (auto &e : h->kev)
        for (hipEvent_t &x : e)
            if (!x) EGS_HIP(hipEventCreate(&x));
    for (int k = 0; k < 4; k++) { h->kms[k] = 0; h->klaunch[k] = 0; }
    return EGSGPU_OK;
}

// add the duration of the last launch of kernel k (its events completed)
int take_time(egsgpu_handle *h, int k) {
    float ms = 0;
    EGS_HIP(hipEventElapsedTime(&ms, h->kev[k][0], h->kev[k][1]));
    h->kms[k] += ms;
    h->klaunch[k]++;
    return EGSGPU_OK;
}

template <typename R>
int launch(egsgpu_handle *h, Precision<R> &P, int64_t n, const DevGrad *grad = nullptr,
           long long hist0 = -1, bool timed = true) {
    KArgs<R> a = kargs(h, P);
    if (hist0 >= 0) a.hist0 = hist0;   // a replay of histories already counted
    a.base = n / h->nlanes;
    a.rem = (int)(n % h->nlanes);
    a.grad = grad;
    size_t lds = (size_t)kRing * h->block * sizeof(float);
    if (timed) EGS_HIP(hipEventRecord(h->kev[0][0], h->stream));
    dim3 grid((h->nlanes + h->block - 1) / h->block);
    const bool lux = h->rng_kind == EGSGPU_RNG_RANLUX;
    if (h->P.n_split > 1 || h->P.nbr_split > 1 || h->P.ireject) {
        // splitting or range rejection: the plain history loop
        if (h->howfarless && lux)
            hipLaunchKernelGGL((k_transport<R, true, true, true>), grid, dim3(h->block), lds,
                               h->stream, a);
        else if (h->howfarless)
            hipLaunchKernelGGL((k_transport<R, false, true, true>), grid, dim3(h->block), lds,
                               h->stream, a);
        else if (lux)
            hipLaunchKernelGGL((k_transport<R, true, false, true>), grid, dim3(h->block), lds,
                               h->stream, a);
        else
            hipLaunchKernelGGL((k_transport<R, false, false, true>), grid, dim3(h->block), lds,
                               h->stream, a);
    } else if (h->howfarless) {   // HOWFARLESS: the plain history loop, ranmar or ranlux
        if (lux)
            hipLaunchKernelGGL((k_transport<R, true, true>), grid, dim3(h->block), lds, h->stream,
                               a);
        else
            hipLaunchKernelGGL((k_transport<R, false, true>), grid, dim3(h->block), lds, h->stream,
                               a);
    } else if (h->P.kerma && h->schedule == EGSGPU_SCHED_HISTORY) {
        if (lux)
            hipLaunchKernelGGL((k_kerma<R, true>), grid, dim3(h->block), lds, h->stream, a,
                               h->ev_ksteps, h->ev_batch, h->ev_max_wait);
        else
            hipLaunchKernelGGL((k_kerma<R, false>), grid, dim3(h->block), lds, h->stream, a,
                               h->ev_ksteps, h->ev_batch, h->ev_max_wait);
    } else if (lux) {
        hipLaunchKernelGGL((k_transport<R, true>), grid, dim3(h->block), lds, h->stream, a);
    } else {
        hipLaunchKernelGGL((k_transport<R, false>), grid, dim3(h->block), lds, h->stream, a);
    }
    EGS_HIP(hipGetLastError());
    if (timed) EGS_HIP(hipEventRecord(h->kev[0][1], h->stream));
    return EGSGPU_OK;
}

template <typename R>
int alloc_bank(egsgpu_handle *h, int cap, DevBank<R> &b) {
    b.cap = cap;
    R **f[] = {&b.e, &b.x, &b.y, &b.z, &b.u, &b.v, &b.w, &b.wt, &b.dn};
    for (R **x : f) CK(alloc(h, (size_t)cap, x));
    CK(alloc(h, (size_t)cap, &b.ir));
    CK(alloc(h, (size_t)cap, &b.tag));
    R **g[] = {&b.s0, &b.s1, &b.s2, &b.s3};
    for (R **x : g) CK(alloc(h, (size_t)cap, x));
    CK(alloc(h, (size_t)cap, &b.ph));
    return EGSGPU_OK;
}

// Event-schedule state: history slots, banks, control words.
template <typename R>
int ensure_event(egsgpu_handle *h, Precision<R> &P) {
    if (h->ev_ready) return EGSGPU_OK;
    h->ev_alloc_first = h->allocs.size();
    DevSlots &S = h->slots;
    S.nslots = h->nslots > 0 ? h->nslots : 4 * h->nlanes;
    S.hcap = h->slot_hcap;
    size_t ne = (size_t)S.nslots * S.hcap;
    CK(alloc(h, ne, &S.ent));
    CK(alloc(h, ne, &S.list));
    CK(alloc(h, (size_t)S.nslots, &S.used));
    CK(alloc(h, (size_t)S.nslots, &S.pending));
    CK(alloc(h, (size_t)S.nslots, &S.ring));
    EGS_HIP(hipMemset(S.ent, 0, ne * sizeof(SlotEntry)));
    CK(alloc(h, 1, &h->ctl));
    EGS_HIP(hipHostMalloc((void **)&h->ctl_host, sizeof(EvCtl), hipHostMallocDefault));
    CK(alloc(h, (size_t)S.nslots, &h->slot_job));
    {
        JobDesc<R> jd;
        jd.G = P.G;
        jd.vox = P.vox;
        jd.sum = h->tally.sum;
        jd.sum2 = h->tally.sum2;
        jd.S = P.S;
        jd.hist0 = 0;
        const JobDesc<R> *dj = nullptr;
        CK(upload(h, &jd, 1, &dj));
        h->job_self = (void *)dj;
    }
    // a history in flight may hold several photons (bremsstrahlung,
    // annihilation) as well as charged particles: 8 entries per slot for
    // either type
    for (int k = 0; k < 4; k++) CK(alloc_bank(h, 8 * S.nslots, P.bank[k]));
    h->pc = 0;
    h->ec = 2;
    hipLaunchKernelGGL(k_ev_init, dim3((S.nslots + 255) / 256), dim3(256), 0, h->stream,
                       S, h->ctl);
    EGS_HIP(hipGetLastError());
    EGS_HIP(hipStreamSynchronize(h->stream));
    h->free_head = 0;
    h->ev_ready = true;
    h->ev_alloc_end = h->allocs.size();
    return EGSGPU_OK;
}

// Free the event schedule's slots, banks and control words (between runs,
// when every bank is empty), so that the next run makes them again with
// another number of slots.
int release_event(egsgpu_handle *h) {
    if (!h->ev_ready) return EGSGPU_OK;
    EGS_HIP(hipStreamSynchronize(h->stream));
    for (size_t i = h->ev_alloc_first; i < h->ev_alloc_end; i++) (void)hipFree(h->allocs[i]);
    h->allocs.erase(h->allocs.begin() + (long)h->ev_alloc_first,
                    h->allocs.begin() + (long)h->ev_alloc_end);
    if (h->ctl_host) (void)hipHostFree(h->ctl_host);
    h->ctl_host = nullptr;
    h->ctl = nullptr;
    h->ev_ready = false;
    return EGSGPU_OK;
}

template <typename R>
using EvKernel = void (*)(EvArgs<R>);

// ranlux runs use the 2-waves-per-SIMD build only
template <typename R>
EvKernel<R> photon_kernel(int w, bool lux) {
    if (lux) return k_ev_photon<R, 2, true>;
    switch (w) {
    case 3: return k_ev_photon<R, 3, false>;
    case 4: return k_ev_photon<R, 4, false>;
    default: return k_ev_photon<R, 2, false>;
    }
}

template <typename R>
EvKernel<R> electron_kernel(int w, bool lux) {
    if (lux) return k_ev_electron<R, 2, true>;
    switch (w) {
    case 3: return k_ev_electron<R, 3, false>;
    case 4: return k_ev_electron<R, 4, false>;
    default: return k_ev_electron<R, 2, false>;
    }
}

int read_ctl(egsgpu_handle *h, EvCtl &c) {
    EGS_HIP(hipMemcpyAsync(h->ctl_host, h->ctl, sizeof(EvCtl), hipMemcpyDeviceToHost, h->stream));
    EGS_HIP(hipStreamSynchronize(h->stream));
    c = *h->ctl_host;
    return EGSGPU_OK;
}

// Event schedule: inject new histories into free slots while banks drain.
// One iteration = source, photon kernel, charged-particle kernel; the host
// reads the bank counts after each iteration.  Each kernel drains one bank
// of its type (pc, ec) and defers leftovers to the other one, which is
// drained in the next iteration.
// A batch of simulations run together (egsgpu_run_batched).
template <typename R>
struct BatchRun {
    const JobDesc<R> *jobs = nullptr;   // device
    const long long *job_first = nullptr;
    int njobs = 0;
};

template <typename R>
int run_event_graph(egsgpu_handle *h, Precision<R> &P, int64_t nhist, const BatchRun<R> *batch);

template <typename R>
int run_event(egsgpu_handle *h, Precision<R> &P, int64_t nhist,
              const BatchRun<R> *batch = nullptr) {
    if (h->use_graph) return run_event_graph(h, P, nhist, batch);
    CK(ensure_event(h, P));
    EvArgs<R> a;
    a.k = kargs(h, P);
    a.jobs = batch ? batch->jobs : (const JobDesc<R> *)h->job_self;
    a.job_first = batch ? batch->job_first : nullptr;
    a.njobs = batch ? batch->njobs : 1;
    a.slot_job = h->slot_job;
    a.hist_base = 0;
    a.slots = h->slots;
    a.ctl = h->ctl;
    a.ksteps = h->ev_ksteps;
    a.batch = h->ev_batch;
    a.max_wait = h->ev_max_wait;
    a.defer_below = h->ev_defer;
    a.gated = 0;
    const int eb = h->ev_block ? h->ev_block : h->block;   // workgroup size
    const dim3 grid((h->nlanes + eb - 1) / eb), block(eb);
    // persistent kernels: launch what stays resident at the chosen variant
    auto resident = [&](int w) {
        int n = std::min(h->nlanes, h->ncu * 4 * 64 * w);
        return dim3((n + eb - 1) / eb);
    };
    const dim3 grid_p = resident(h->ev_wp), grid_e = resident(h->ev_we);
    const bool lux = h->rng_kind == EGSGPU_RNG_RANLUX;
    auto kp = photon_kernel<R>(h->ev_wp, lux);
    auto ke = electron_kernel<R>(h->ev_we, lux);
    auto ks = lux ? k_ev_source<R, true> : k_ev_source<R, false>;
    const size_t lds = (size_t)kRing * eb * sizeof(float);
    const bool photon_src = P.S.charge == 0;
    auto setb = [&](int ic, int ip, int id) {
        a.ic = ic; a.ip = ip; a.id = id;
        a.cons = P.bank[ic]; a.push = P.bank[ip]; a.defr = P.bank[id];
    };
    auto banks_empty = [](const EvCtl &c) {
        return c.n[0] == 0 && c.n[1] == 0 && c.n[2] == 0 && c.n[3] == 0;
    };
    EvCtl c;
    CK(read_ctl(h, c));
    int64_t remaining = nhist;
    const unsigned long long ns = (unsigned long long)h->slots.nslots;
    for (int64_t iter = 0;; iter++) {
        const int pc = h->pc, ec = h->ec;
        unsigned long long freeslots = c.free_tail - h->free_head;
        int64_t inject = std::min<int64_t>((int64_t)freeslots, remaining);
        int64_t room = photon_src ? (int64_t)P.bank[pc].cap - c.n[pc]
                                  : (int64_t)P.bank[ec].cap - c.n[ec];
        inject = std::max<int64_t>(0, std::min(inject, room));
        if (inject == 0 && banks_empty(c)) {
            set_error("event schedule: no free history slot with empty banks");
            return EGSGPU_E_OVERFLOW;
        }
        bool timed_src = false;
        if (inject > 0) {
            a.free_head = h->free_head;
            a.ninject = (int)inject;
            setb(photon_src ? pc : ec, photon_src ? pc : ec, photon_src ? pc : ec);
            EGS_HIP(hipEventRecord(h->kev[1][0], h->stream));
            hipLaunchKernelGGL(ks, grid, block, lds, h->stream, a);
            EGS_HIP(hipGetLastError());
            EGS_HIP(hipEventRecord(h->kev[1][1], h->stream));
            timed_src = true;
            h->free_head += (unsigned long long)inject;
            remaining -= inject;
            a.hist_base += inject;
        }
        // photons: drain pc, electrons to ec, deferred photons to the other
        setb(pc, ec, 1 - pc);
        EGS_HIP(hipEventRecord(h->kev[2][0], h->stream));
        hipLaunchKernelGGL(kp, grid_p, block, lds, h->stream, a);
        EGS_HIP(hipGetLastError());
        EGS_HIP(hipEventRecord(h->kev[2][1], h->stream));
        hipLaunchKernelGGL(k_ev_ctl, dim3(1), dim3(64), 0, h->stream, h->ctl, pc);
        h->pc = 1 - pc;
        // charged: drain ec, photons to the new pc, deferred to the other
        setb(ec, h->pc, 5 - ec);
        EGS_HIP(hipEventRecord(h->kev[3][0], h->stream));
        hipLaunchKernelGGL(ke, grid_e, block, lds, h->stream, a);
        EGS_HIP(hipGetLastError());
        EGS_HIP(hipEventRecord(h->kev[3][1], h->stream));
        hipLaunchKernelGGL(k_ev_ctl, dim3(1), dim3(64), 0, h->stream, h->ctl, ec);
        EGS_HIP(hipGetLastError());
        h->ec = 5 - ec;
        CK(read_ctl(h, c));
        if (timed_src) CK(take_time(h, 1));
        CK(take_time(h, 2));
        CK(take_time(h, 3));
        if (remaining == 0 && banks_empty(c)) break;
    }
    if (c.free_tail - h->free_head != ns) {
        set_error("event schedule ended with %llu of %llu history slots in use",
                  (unsigned long long)(ns - (c.free_tail - h->free_head)), ns);
        return EGSGPU_E_OVERFLOW;
    }
    return EGSGPU_OK;
}

// The same iteration (plan, source, photon kernel, charged-particle
// kernel, bank bookkeeping) captured twice -- once per bank parity -- into
// a HIP graph that is replayed until every history has been started and
// all banks are empty.  Per-kernel times are not split in this mode (the
// whole run is timed); the host-driven loop gives them.  The source launches are planned on the device
// (k_ev_plan), so a replay needs nothing from the host; the host reads the
// control words once per replay.
template <typename R>
int run_event_graph(egsgpu_handle *h, Precision<R> &P, int64_t nhist,
                    const BatchRun<R> *batch) {
    CK(ensure_event(h, P));
    EvArgs<R> a;
    a.k = kargs(h, P);
    a.jobs = batch ? batch->jobs : (const JobDesc<R> *)h->job_self;
    a.job_first = batch ? batch->job_first : nullptr;
    a.njobs = batch ? batch->njobs : 1;
    a.slot_job = h->slot_job;
    a.hist_base = 0;
    a.free_head = 0;
    a.ninject = -1;   // planned on the device
    a.slots = h->slots;
    a.ctl = h->ctl;
    a.ksteps = h->ev_ksteps;
    a.batch = h->ev_batch;
    a.max_wait = h->ev_max_wait;
    a.defer_below = h->ev_defer;
    a.gated = 1;      // the kernels return at once after the run has ended
    const int eb = h->ev_block ? h->ev_block : h->block;   // workgroup size
    const dim3 grid((h->nlanes + eb - 1) / eb), block(eb);
    auto resident = [&](int w) {
        int n = std::min(h->nlanes, h->ncu * 4 * 64 * w);
        return dim3((n + eb - 1) / eb);
    };
    const dim3 grid_p = resident(h->ev_wp), grid_e = resident(h->ev_we);
    const bool lux = h->rng_kind == EGSGPU_RNG_RANLUX;
    auto kp = photon_kernel<R>(h->ev_wp, lux);
    auto ke = electron_kernel<R>(h->ev_we, lux);
    auto ks = lux ? k_ev_source<R, true> : k_ev_source<R, false>;
    const size_t lds = (size_t)kRing * eb * sizeof(float);
    const bool photon_src = P.S.charge == 0;
    auto setb = [&](int ic, int ip, int id) {
        a.ic = ic; a.ip = ip; a.id = id;
        a.cons = P.bank[ic]; a.push = P.bank[ip]; a.defr = P.bank[id];
    };
    hipLaunchKernelGGL(k_ev_setrun, dim3(1), dim3(64), 0, h->stream, h->ctl, (long long)nhist,
                       h->free_head);
    EGS_HIP(hipGetLastError());
    // One graph holds as many iterations as the run is expected to need
    // (the count of the previous run, with a margin; an even number, so that
    // the bank roles are back where they started); the device plans every
    // iteration and, once the run has ended, the kernels still queued return
    // at once.  A run that needs more iterations launches the graph again.
    const int iters = h->ev_graph_iters;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    EGS_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    int pc = h->pc, ec = h->ec;
    hipError_t cerr = hipSuccess;
    for (int it = 0; it < iters && cerr == hipSuccess; it++) {
        int sb = photon_src ? pc : ec;
        hipLaunchKernelGGL(k_ev_plan, dim3(1), dim3(64), 0, h->stream, h->ctl, sb,
                           P.bank[sb].cap);
        setb(sb, sb, sb);
        hipLaunchKernelGGL(ks, grid, block, lds, h->stream, a);
        setb(pc, ec, 1 - pc);
        hipLaunchKernelGGL(kp, grid_p, block, lds, h->stream, a);
        hipLaunchKernelGGL(k_ev_ctl, dim3(1), dim3(64), 0, h->stream, h->ctl, pc);
        pc = 1 - pc;
        setb(ec, pc, 5 - ec);
        hipLaunchKernelGGL(ke, grid_e, block, lds, h->stream, a);
        hipLaunchKernelGGL(k_ev_ctl, dim3(1), dim3(64), 0, h->stream, h->ctl, ec);
        ec = 5 - ec;
        cerr = hipGetLastError();
    }
    hipError_t eerr = hipStreamEndCapture(h->stream, &graph);
    if (cerr != hipSuccess || eerr != hipSuccess ||
        hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
        // no graph on this runtime: the host-driven loop does the same work
        if (graph) (void)hipGraphDestroy(graph);
        (void)hipGetLastError();
        h->use_graph = false;
        return run_event(h, P, nhist, batch);
    }
    auto cleanup = [&]() {
        (void)hipGraphExecDestroy(exec);
        (void)hipGraphDestroy(graph);
    };
    EvCtl c;
    int rc = EGSGPU_OK;
    const unsigned long long ns = (unsigned long long)h->slots.nslots;
    int launches = 0;
    for (;;) {
        if (hipGraphLaunch(exec, h->stream) != hipSuccess) {
            set_error("graph launch failed");
            rc = EGSGPU_E_HIP;
            break;
        }
        launches++;
        if ((rc = read_ctl(h, c))) break;
        if (c.stall) {
            set_error("event schedule: no free history slot with empty banks");
            rc = EGSGPU_E_OVERFLOW;
            break;
        }
        bool empty = c.n[0] == 0 && c.n[1] == 0 && c.n[2] == 0 && c.n[3] == 0;
        if (c.remaining == 0 && empty) break;
    }
    cleanup();
    if (rc) return rc;
    h->graph_launches = launches;
    h->graph_iters_used = c.iters;
    // size the next run's graph from this run: 25 % more iterations plus
    // two, rounded up to an even number
    int want = c.iters + c.iters / 4 + 2;
    want = std::max(4, std::min(kMaxGraphIters, want + (want & 1)));
    h->ev_graph_iters = want;
    h->free_head = c.free_head;
    if (c.free_tail - h->free_head != ns) {
        set_error("event schedule ended with %llu of %llu history slots in use",
                  (unsigned long long)(ns - (c.free_tail - h->free_head)), ns);
        return EGSGPU_E_OVERFLOW;
    }
    return EGSGPU_OK;
}

}  // namespace

extern "C" {

int egsgpu_create(const egsgpu_tables *tab, const EgsXYZGeometry *geom,
                  const EgsSource *src, const EgsTransport *tp,
                  egsgpu_handle **out) {
    EGS_ARG(tab && geom && src && tp && out, "tables, geometry, source, transport, out");
    EGS_ARG(geom->nx > 0 && geom->ny > 0 && geom->nz > 0 && geom->xb && geom->yb && geom->zb,
            "geometry planes");
    EGS_ARG(src->charge >= -1 && src->charge <= 1, "source charge");
    EGS_ARG(src->spectrum == EGSGPU_SPEC_MONO || ((src->spectrum == EGSGPU_SPEC_HISTOGRAM ||
            src->spectrum == EGSGPU_SPEC_LINE || src->spectrum == EGSGPU_SPEC_CDFINV) &&
            src->nbins > 0 && src->spec_e && src->spec_w),
            "source spectrum");
    EGS_ARG(src->kind >= EGSGPU_SOURCE_PARALLEL_BEAM && src->kind <= EGSGPU_SOURCE_FANO,
            "source kind");
    EGS_ARG(src->kind != EGSGPU_SOURCE_FANO ||
            (src->x0 < src->x1 && src->y0 < src->y1 && src->z < src->z1 && src->rho_max > 0),
            "a Fano source needs a box x0 < x1, y0 < y1, z < z1 and rho_max > 0");
    EGS_ARG(src->kind != EGSGPU_SOURCE_DOSXYZ_POINT ||
            (src->dsource > 0 && src->x0 <= src->x1 && src->y0 <= src->y1),
            "a point source needs SSD (dsource) > 0 and x0 <= x1, y0 <= y1");
    EGS_ARG(src->kind != EGSGPU_SOURCE_DOSXYZ_FRONT || src->uz > 0,
            "a front-face source must point into the grid (uz > 0)");
    {
        const int blob_nist = tab->blob.has("ibr_nist") ? tab->blob.i1("ibr_nist") : 0;
        if (tp->brems_xsec != blob_nist) {
            set_error("brems_xsec = %d but the table blob was built with %s brems cross "
                      "sections; build it with brems_xsec = %d",
                      (int)tp->brems_xsec,
                      blob_nist == 2 ? "NRC" : blob_nist ? "NIST" : "Bethe-Heitler",
                      (int)tp->brems_xsec);
            return EGSGPU_E_ARG;
        }
        // egs_init_user_photon (egsnrc.mortran:7973): the Compton part
        // of the photon cross sections is bound (compton_sigma.data) for
        // IBCMP > 1 and Klein-Nishina otherwise
        const int blob_ibcmp = tab->blob.i1("ibcmp");
        if (tp->bound_compton > 1 && blob_ibcmp <= 1) {
            set_error("bound_compton = %d needs tables built with bound Compton cross "
                      "sections (egsgpu_tables_build with bound_compton > 1)",
                      (int)tp->bound_compton);
            return EGSGPU_E_ARG;
        }
        if (tp->bound_compton <= 1 && blob_ibcmp > 1 && !tab->blob.has(Blob::mk(0, "gmfp0_kn"))) {
            set_error("bound_compton = %d needs the Klein-Nishina photon rows, which this "
                      "table blob lacks; rebuild it", (int)tp->bound_compton);
            return EGSGPU_E_ARG;
        }
        const std::string want(tp->photon_xsec, strnlen(tp->photon_xsec, sizeof tp->photon_xsec));
        std::string have = "xcom";
        if (tab->blob.has("photon_xsec")) {
            have.clear();
            for (int c : tab->blob.I("photon_xsec")) have.push_back((char)c);
        }
        if (!want.empty() && want != have) {
            set_error("photon_xsec = %s but the table blob was built from the %s photon data",
                      want.c_str(), have.c_str());
            return EGSGPU_E_ARG;
        }
        // EII: the blob's electron cross sections include it or not, so the
        // run must ask for what the blob was built with
        std::string eii_want(tp->eii_xsec, strnlen(tp->eii_xsec, sizeof tp->eii_xsec));
        for (auto &c : eii_want) c = (char)std::tolower((unsigned char)c);
        if (eii_want == "on") eii_want = "ik";
        if (eii_want == "off") eii_want.clear();
        std::string eii_have;
        if (tab->blob.has("eii_name"))
            for (int c : tab->blob.I("eii_name")) eii_have.push_back((char)c);
        if (eii_want != eii_have) {
            set_error("eii_xsec = '%s' but the table blob was built with %s%s", eii_want.c_str(),
                      eii_have.empty() ? "EII off" : "EII data ", eii_have.c_str());
            return EGSGPU_E_ARG;
        }
        if (!eii_want.empty() && !tp->atomic_relaxation) {
            set_error("electron impact ionisation needs atomic relaxation (eii_init)");
            return EGSGPU_E_ARG;
        }
        const int blob_nrcp = tab->blob.has("pair_nrc") ? tab->blob.i1("pair_nrc") : 0;
        if (tp->pair_xsec != 0 && (tp->pair_xsec != 1 || !blob_nrcp)) {
            set_error("pair_xsec = %d needs a table blob built with the NRC pair tables "
                      "(pair_xsec = 1)", (int)tp->pair_xsec);
            return EGSGPU_E_ARG;
        }
    }
    int rc = require_device();
    if (rc) return rc;
    if (tp->spin) {
        set_error("spin effects need the spinms data, which the engine does not have");
        return EGSGPU_E_UNSUPPORTED;
    }
    if (tp->bound_compton < 0 || tp->bound_compton > 3 || tp->rayleigh < 0 ||
        tp->rayleigh > 1 || tp->atomic_relaxation < 0 || tp->atomic_relaxation > 1 ||
        tp->photoelectron_angle < 0 || tp->photoelectron_angle > 1) {
        set_error("bound_compton %d (0 Off, 1 On, 2 Simple, 3 norej), rayleigh %d, "
                  "atomic_relaxation %d (0 Off, 1 EADL), photoelectron_angle %d (0/1)",
                  (int)tp->bound_compton, (int)tp->rayleigh, (int)tp->atomic_relaxation,
                  (int)tp->photoelectron_angle);
        return EGSGPU_E_ARG;
    }
    if (tp->pair_angle < 0 || tp->pair_angle > 4 || tp->brems_angle < 0 || tp->brems_angle > 1) {
        set_error("pair_angle %d / brems_angle %d: pair angles take 0 Off, 1 Simple, 2 KM, "
                  "3 Uniform, 4 Blend; brems angles 0 Simple, 1 KM",
                  (int)tp->pair_angle, (int)tp->brems_angle);
        return EGSGPU_E_ARG;
    }
    if (tp->step_algorithm != 0 && tp->step_algorithm != 1) {
        set_error("step_algorithm = %d: 0 is PRESTA-II, 1 is PRESTA-I", (int)tp->step_algorithm);
        return EGSGPU_E_ARG;
    }
    if (tp->n_split < 0 || tp->n_split > 256) {
        set_error("n_split = %d: photon splitting takes 0..256", (int)tp->n_split);
        return EGSGPU_E_ARG;
    }
    if (tp->nbr_split < 0 || tp->nbr_split > 256) {
        set_error("nbr_split = %d: radiative splitting takes 0..256", (int)tp->nbr_split);
        return EGSGPU_E_ARG;
    }
    if (tp->n_split > 1 && tp->nbr_split > 1) {
        set_error("photon splitting (n_split) and radiative splitting (nbr_split) are "
                  "separate variance-reduction schemes; choose one");
        return EGSGPU_E_ARG;
    }
    if (tp->ireject && tp->howfarless) {
        set_error("range rejection is not implemented with HOWFARLESS");
        return EGSGPU_E_UNSUPPORTED;
    }
    if (tp->ireject && tp->kerma) {
        set_error("range rejection needs charged-particle transport (kerma = 0)");
        return EGSGPU_E_ARG;
    }
    if (tp->play_rr && tp->nbr_split <= 1) {
        set_error("Russian roulette (play_rr) plays with probability 1/nbr_split: it needs "
                  "nbr_split > 1");
        return EGSGPU_E_ARG;
    }
    if (tp->howfarless) {   // one medium and one density everywhere
        const int64_t nv = (int64_t)geom->nx * geom->ny * geom->nz;
        for (int64_t i = 1; i < nv; i++)
            if ((geom->medium && geom->medium[i] != geom->medium[0]) ||
                (geom->rhor && geom->rhor[i] != geom->rhor[0])) {
                set_error("HOWFARLESS needs a homogeneous phantom (voxel %lld differs)",
                          (long long)i);
                return EGSGPU_E_ARG;
            }
    }
    const Blob &b = tab->blob;
    auto *h = new egsgpu_handle();
    auto fail = [&](int code) { delete h; return code; };
    h->fp64 = tp->fp64 ? 1 : 0;
    h->tab = tab;
    h->nvox = geom->nx * geom->ny * geom->nz;
    int nmed = b.i1("nmed");
    if (nmed > kMaxMed) { set_error("too many media"); return fail(EGSGPU_E_UNSUPPORTED); }
    if (tp->rayleigh && !b.has(Blob::mk(0, "ray_x"))) {
        set_error("table blob has no Rayleigh data"); return fail(EGSGPU_E_FORMAT);
    }
    // media map
    std::vector<uint8_t> med(h->nvox, 0);
    for (int i = 0; i < h->nvox; i++) {
        int m = geom->medium ? geom->medium[i] : 0;
        if (m < 0 || m >= nmed) { set_error("voxel %d has medium %d outside the table", i, m); return fail(EGSGPU_E_ARG); }
        med[i] = (uint8_t)m;
    }
    if (hipStreamCreate(&h->stream) != hipSuccess || hipEventCreate(&h->ev0) != hipSuccess ||
        hipEventCreate(&h->ev1) != hipSuccess) {
        set_error("stream/event creation failed");
        return fail(EGSGPU_E_HIP);
    }
    h->hx.assign(geom->xb, geom->xb + geom->nx + 1);
    h->hy.assign(geom->yb, geom->yb + geom->ny + 1);
    h->hz.assign(geom->zb, geom->zb + geom->nz + 1);
    h->hrhor.assign(h->nvox, 1.0);
    h->hmass.assign(h->nvox, 0.0);
    h->rho_max.assign(nmed, 0.0);
    for (int i = 0; i < h->nvox; i++) {
        if (geom->rhor) h->hrhor[i] = geom->rhor[i];
        h->rho_max[med[i]] = std::max(h->rho_max[med[i]], h->hrhor[i]);
        int iz = i / (geom->nx * geom->ny), iy = (i / geom->nx) % geom->ny, ix = i % geom->nx;
        double vol = (geom->xb[ix + 1] - geom->xb[ix]) * (geom->yb[iy + 1] - geom->yb[iy]) *
                     (geom->zb[iz + 1] - geom->zb[iz]);
        h->hmass[i] = vol * h->hrhor[i] * b.f1(Blob::mk(med[i], "rho"));
        // egs_fano_source: the acceptance ratio must stay <= 1 in every voxel
        if (src->kind == EGSGPU_SOURCE_FANO &&
            h->hrhor[i] * b.f1(Blob::mk(med[i], "rho")) > src->rho_max * (1 + 1e-9)) {
            set_error("Fano source: voxel %d has mass density %g g/cm^3 above rho_max %g", i,
                      h->hrhor[i] * b.f1(Blob::mk(med[i], "rho")), src->rho_max);
            return fail(EGSGPU_E_ARG);
        }
    }
    // cut-offs: $adjust_rhor_ecut_pcut
    double ecut = 0, pcut = 0;
    for (int m = 0; m < nmed; m++) {
        ecut = std::max(ecut, b.f1(Blob::mk(m, "ae")));
        pcut = std::max(pcut, b.f1(Blob::mk(m, "ap")));
    }
    h->P.ecut = std::max(ecut, tp->ecut);
    h->P.pcut = std::max(pcut, tp->pcut);
    h->P.rayleigh = tp->rayleigh; h->P.ibcmp = tp->bound_compton;
    h->P.iphter = tp->photoelectron_angle; h->P.iedgfl = tp->atomic_relaxation;
    h->P.iprdst = tp->pair_angle; h->P.ibrdst = tp->brems_angle;
    h->P.exact_bca = tp->exact_bca; h->P.kerma = tp->kerma;
    h->P.n_split = tp->n_split > 1 ? tp->n_split : 0;
    h->P.nbr_split = tp->nbr_split > 1 ? tp->nbr_split : 0;
    h->P.play_rr = tp->play_rr ? 1 : 0;
    // DOSXYZnrc switches range rejection off when ESAVE_GLOBAL <= ECUT
    h->P.esave = tp->esave;
    h->P.ireject = tp->ireject && tp->esave > h->P.ecut ? 1 : 0;
    h->P.presta1 = tp->step_algorithm == 1;
    h->P.ibr_nist = tp->brems_xsec;
    h->P.pair_nrc = tp->pair_xsec;
    h->P.eii = b.has("eii") && b.i1("eii") == 1 ? 1 : 0;
    // mscati (egsnrc.mortran:2790-2817): skindepth_for_bca <= 1e-4 selects
    // 3 elastic mean free paths with exact boundary crossing and
    // exp(BLCMIN) at the lowest ECUT with PRESTA-I, the minimum taken over
    // the media of the problem (those the voxels use)
    if (tp->skin_depth > 1e-4) {
        h->P.skindepth = tp->skin_depth;
    } else if (tp->exact_bca) {
        h->P.skindepth = 3.0;
    } else {
        std::vector<char> seen(nmed, 0);
        for (int i = 0; i < h->nvox; i++) seen[med[i]] = 1;
        std::vector<int> used;
        for (int m = 0; m < nmed; m++)
            if (seen[m]) used.push_back(m);
        h->P.skindepth = std::exp(bca_blcmin(b, used, h->P.ecut));
    }
    try {
        if (h->fp64) {
            if ((rc = build_tables(h, b, h->pd.T))) return fail(rc);
            if ((rc = build_geom(h, geom, med, h->pd))) return fail(rc);
            if ((rc = build_source(h, src, h->pd))) return fail(rc);
        } else {
            if ((rc = build_tables(h, b, h->pf.T))) return fail(rc);
            if ((rc = build_geom(h, geom, med, h->pf))) return fail(rc);
            if ((rc = build_source(h, src, h->pf))) return fail(rc);
        }
    } catch (const std::exception &e) {
        set_error("%s", e.what());
        return fail(EGSGPU_E_FORMAT);
    }
    // the dose regions, then the energy-bookkeeping regions (escaped,
    // incident) of Shower::score_extra
    if ((rc = alloc(h, (size_t)h->nvox + kExtraRegions, &h->tally.sum))) return fail(rc);
    if ((rc = alloc(h, (size_t)h->nvox + kExtraRegions, &h->tally.sum2))) return fail(rc);
    if ((rc = alloc(h, 32, &h->tally.counters))) return fail(rc);
    if (hipMemset(h->tally.sum, 0, sizeof(double) * (h->nvox + kExtraRegions)) != hipSuccess ||
        hipMemset(h->tally.sum2, 0, sizeof(double) * (h->nvox + kExtraRegions)) != hipSuccess ||
        hipMemset(h->tally.counters, 0, 256) != hipSuccess) {
        set_error("tally clear failed");
        return fail(EGSGPU_E_HIP);
    }
    hipDeviceProp_t prop;
    int dev = 0;
    (void)hipGetDevice(&dev);
    (void)hipGetDeviceProperties(&prop, dev);
    h->ncu = std::max(1, prop.multiProcessorCount);
    h->nlanes = 0;   // chosen at the first run
    // per-history tally hash: a kerma history touches a few voxels, a
    // coupled electron shower up to several hundred
    h->hcap = tp->kerma ? 256 : 2048;
    h->mxstk = tp->kerma ? 32 : 64;
    // the event schedule pays off where charged particles are transported
    h->schedule = tp->kerma ? EGSGPU_SCHED_HISTORY : EGSGPU_SCHED_EVENT;
    h->howfarless = tp->howfarless != 0;
    if (h->howfarless) h->schedule = EGSGPU_SCHED_HISTORY_PLAIN;
    if (h->P.ireject) h->schedule = EGSGPU_SCHED_HISTORY_PLAIN;
    if (h->P.nbr_split) {
        // a radiative event puts nbr_split photons (pairs) on the stack
        h->schedule = EGSGPU_SCHED_HISTORY_PLAIN;
        h->mxstk = std::max(h->mxstk, 4 * h->P.nbr_split + 64);
    }
    if (h->P.n_split) {
        // a split leaves about one charged particle per sub-photon on the
        // stack; a photon from a radiative event of one of them is split
        // on top of those
        h->schedule = EGSGPU_SCHED_HISTORY_PLAIN;
        h->mxstk = std::max(h->mxstk, 4 * h->P.n_split + 64);
    }
    // electron beams touch more voxels per history than photon beams
    h->slot_hcap = tp->kerma ? 64 : (src->charge || src->kind == EGSGPU_SOURCE_PHSP ? 512 : 256);
    *out = h;
    return EGSGPU_OK;
}

int egsgpu_destroy(egsgpu_handle *h) {
    delete h;
    return EGSGPU_OK;
}

int egsgpu_set_streams(egsgpu_handle *h, int32_t nstreams, int32_t block_size) {
    EGS_ARG(h, "handle");
    EGS_ARG(nstreams >= 0 && (block_size == 0 || (block_size % 64 == 0 && block_size <= 256)),
            "nstreams >= 0, block size a multiple of 64 up to 256");
    if (h->lanes_ready) {
        set_error("launch shape must be set before the first run");
        return EGSGPU_E_ARG;
    }
    if (nstreams > 0) h->nlanes = nstreams;
    if (block_size > 0) h->block = block_size;
    return EGSGPU_OK;
}

int egsgpu_set_phsp(egsgpu_handle *h, int64_t n, const float *x, const float *y,
                    const float *u, const float *v, const float *e, const float *wt,
                    const int8_t *q) {
    EGS_ARG(h && n > 0 && x && y && u && v && e && wt && q, "handle, n > 0, particle arrays");
    DevSource<float> &Sf = h->pf.S;
    DevSource<double> &Sd = h->pd.S;
    if (Sf.kind != EGSGPU_SOURCE_PHSP) {
        set_error("the simulation was not created with a phase-space source");
        return EGSGPU_E_ARG;
    }
    if (h->lanes_ready) {
        set_error("phase-space particles must be given before the first run");
        return EGSGPU_E_ARG;
    }
    for (int64_t k = 0; k < n; k++)
        if (q[k] < -1 || q[k] > 1 || !(e[k] >= 0)) {
            set_error("phase-space particle %lld: charge %d, energy %g", (long long)k, (int)q[k],
                      (double)e[k]);
            return EGSGPU_E_ARG;
        }
    const float *src[6] = {x, y, u, v, e, wt};
    const float *dst[6] = {};
    for (int f = 0; f < 6; f++) CK(upload(h, src[f], (size_t)n, &dst[f]));
    const signed char *dq = nullptr;
    CK(upload(h, (const signed char *)q, (size_t)n, &dq));
    auto set = [&](auto &S) {
        S.nphsp = n;
        S.px = dst[0]; S.py = dst[1]; S.pu = dst[2]; S.pv = dst[3]; S.pe = dst[4]; S.pw = dst[5];
        S.pq = dq;
    };
    set(Sf);
    set(Sd);
    return EGSGPU_OK;
}

int egsgpu_seed(egsgpu_handle *h, int32_t ixx, int32_t jxx) {
    return egsgpu_set_rng(h, EGSGPU_RNG_RANMAR, ixx, jxx);
}

int egsgpu_set_rng(egsgpu_handle *h, int32_t kind, int32_t a, int32_t b) {
    EGS_ARG(h, "handle");
    EGS_ARG(kind == EGSGPU_RNG_RANMAR || kind == EGSGPU_RNG_RANLUX, "rng kind");
    if (kind == EGSGPU_RNG_RANMAR) {
        h->ixx = a;
        h->jxx = b;
    } else {
        h->lux = a >= 0 && a <= 4 ? a : 1;   // init_ranlux: out of range -> 1
        h->lux_seed = b;
    }
    h->rng_kind = kind;
    if (h->lanes_ready) CK(seed_streams(h));
    return EGSGPU_OK;
}

int egsgpu_run_batched(egsgpu_handle *const *jobs, int32_t njobs, const int64_t *nhist) {
    EGS_ARG(jobs && njobs > 0 && nhist, "jobs, njobs > 0, nhist");
    egsgpu_handle *h = jobs[0];
    EGS_ARG(h, "jobs[0]");
    int64_t total = 0;
    for (int j = 0; j < njobs; j++) {
        egsgpu_handle *g = jobs[j];
        EGS_ARG(g && nhist[j] >= 0, "every job handle, nhist >= 0");
        if (g->tab != h->tab || g->fp64 != h->fp64 ||
            memcmp(&g->P, &h->P, sizeof(DevParams)) != 0) {
            set_error("job %d differs from job 0 in tables, precision or transport "
                      "parameters; a batch shares them", j);
            return EGSGPU_E_ARG;
        }
        for (int k = 0; k < j; k++)
            if (jobs[k] == g) { set_error("job %d appears twice in the batch", j); return EGSGPU_E_ARG; }
        total += nhist[j];
    }
    if (h->schedule != EGSGPU_SCHED_EVENT) {
        set_error("a batch runs on the event schedule of job 0");
        return EGSGPU_E_UNSUPPORTED;
    }
    if (h->P.woodcock) {
        set_error("Woodcock tracking needs a majorant for each phantom; a batch shares job "
                  "0's: run batches with voxel-by-voxel tracking");
        return EGSGPU_E_UNSUPPORTED;
    }
    // the history slots (job 0's) must hold the voxels a history of any job
    // touches: the largest table any job asks for (charged-particle sources
    // ask for more than photon beams)
    int hcap = h->slot_hcap;
    for (int j = 1; j < njobs; j++) hcap = std::max(hcap, jobs[j]->slot_hcap);
    if (hcap > h->slot_hcap) {
        if (h->ev_ready) {
            set_error("job 0 already ran with history tables of %d entries; this batch needs "
                      "%d (a job with a charged-particle source): run the batch before job 0 "
                      "runs alone", h->slot_hcap, hcap);
            return EGSGPU_E_ARG;
        }
        h->slot_hcap = hcap;
    }
    int rc = ensure_lanes(h);
    if (rc) return rc;
    CK(timing_events(h));
    std::vector<long long> first(njobs + 1, 0);
    for (int j = 0; j < njobs; j++) first[j + 1] = first[j] + nhist[j];
    auto go = [&](auto &P0, auto pick) -> int {
        using R = typename std::remove_reference<decltype(P0.T)>::type::value_type;
        std::vector<JobDesc<R>> jd(njobs);
        for (int j = 0; j < njobs; j++) {
            auto &Pj = pick(jobs[j]);
            jd[j].G = Pj.G;
            jd[j].vox = Pj.vox;
            jd[j].sum = jobs[j]->tally.sum;
            jd[j].sum2 = jobs[j]->tally.sum2;
            jd[j].S = Pj.S;
            jd[j].hist0 = jobs[j]->nhist;
        }
        JobDesc<R> *djobs = nullptr;
        long long *dfirst = nullptr;
        EGS_HIP(hipMalloc(&djobs, sizeof(JobDesc<R>) * njobs));
        if (hipMalloc(&dfirst, sizeof(long long) * (njobs + 1)) != hipSuccess) {
            (void)hipFree(djobs);
            set_error("hipMalloc failed");
            return EGSGPU_E_HIP;
        }
        int r = EGSGPU_OK;
        if (hipMemcpy(djobs, jd.data(), sizeof(JobDesc<R>) * njobs, hipMemcpyHostToDevice) !=
                hipSuccess ||
            hipMemcpy(dfirst, first.data(), sizeof(long long) * (njobs + 1),
                      hipMemcpyHostToDevice) != hipSuccess) {
            set_error("job table upload failed");
            r = EGSGPU_E_HIP;
        }
        if (!r) {
            BatchRun<R> b;
            b.jobs = djobs;
            b.job_first = dfirst;
            b.njobs = njobs;
            EGS_HIP(hipEventRecord(h->ev0, h->stream));
            r = run_event(h, P0, total, &b);
            if (!r) {
                EGS_HIP(hipEventRecord(h->ev1, h->stream));
                EGS_HIP(hipStreamSynchronize(h->stream));
            }
        }
        (void)hipFree(djobs);
        (void)hipFree(dfirst);
        return r;
    };
    rc = h->fp64 ? go(h->pd, [](egsgpu_handle *g) -> Precision<double> & { return g->pd; })
                 : go(h->pf, [](egsgpu_handle *g) -> Precision<float> & { return g->pf; });
    if (rc) return rc;
    float ms = 0;
    EGS_HIP(hipEventElapsedTime(&ms, h->ev0, h->ev1));
    h->last_ms = ms;
    h->last_launches = h->klaunch[1] + h->klaunch[2] + h->klaunch[3];
    for (int j = 0; j < njobs; j++) jobs[j]->nhist += nhist[j];
    return EGSGPU_OK;
}

int egsgpu_set_schedule(egsgpu_handle *h, int32_t schedule) {
    EGS_ARG(h, "handle");
    EGS_ARG(schedule == EGSGPU_SCHED_HISTORY || schedule == EGSGPU_SCHED_EVENT ||
            schedule == EGSGPU_SCHED_HISTORY_PLAIN, "schedule");
    if ((h->howfarless || h->P.n_split || h->P.nbr_split || h->P.ireject) &&
        schedule != EGSGPU_SCHED_HISTORY_PLAIN) {
        set_error("HOWFARLESS, photon or radiative splitting and range rejection run on the "
                  "plain history schedule");
        return EGSGPU_E_UNSUPPORTED;
    }
    if (h->lanes_ready) {
        set_error("the schedule must be chosen before the first run");
        return EGSGPU_E_ARG;
    }
    h->schedule = schedule;
    return EGSGPU_OK;
}

int egsgpu_set_event_shape(egsgpu_handle *h, int32_t waves_photon, int32_t waves_electron,
                           int32_t nslots) {
    EGS_ARG(h, "handle");
    EGS_ARG((waves_photon == 0 || (waves_photon >= 2 && waves_photon <= 4)) &&
            (waves_electron == 0 || (waves_electron >= 2 && waves_electron <= 4)) && nslots >= 0,
            "waves per SIMD 2-4 (0 = keep), nslots >= 0");
    if (nslots && h->ev_ready) {
        set_error("the number of history slots must be set before the first run");
        return EGSGPU_E_ARG;
    }
    if (waves_photon) h->ev_wp = waves_photon;
    if (waves_electron) h->ev_we = waves_electron;
    if (nslots) h->nslots = nslots;
    return EGSGPU_OK;
}

int egsgpu_set_event_graph(egsgpu_handle *h, int32_t on) {
    EGS_ARG(h, "handle");
    h->use_graph = on != 0;
    return EGSGPU_OK;
}

int egsgpu_graph_stats(egsgpu_handle *h, int32_t *launches, int32_t *iterations,
                       int32_t *captured) {
    EGS_ARG(h && launches && iterations && captured, "null argument");
    *launches = h->graph_launches;
    *iterations = h->graph_iters_used;
    *captured = h->ev_graph_iters;
    return EGSGPU_OK;
}

int egsgpu_set_event_batching(egsgpu_handle *h, int32_t photon_steps, int32_t batch,
                              int32_t max_wait, int32_t defer_below) {
    EGS_ARG(h, "handle");
    EGS_ARG(photon_steps >= 0 && batch >= 0 && batch <= 64 && max_wait >= 0 &&
            defer_below >= -1 && defer_below <= 64,
            "photon steps >= 0, batch 0-64, max_wait >= 0, defer_below -1..64 (0 = keep)");
    if (photon_steps) h->ev_ksteps = photon_steps;
    if (batch) h->ev_batch = batch;
    if (max_wait) h->ev_max_wait = max_wait;
    if (defer_below) h->ev_defer = defer_below < 0 ? 0 : defer_below;
    return EGSGPU_OK;
}

int egsgpu_autotune(egsgpu_handle *h, int64_t nhist_probe, int32_t *choice) {
    EGS_ARG(h && nhist_probe > 0, "handle, nhist_probe > 0");
    if (h->schedule != EGSGPU_SCHED_EVENT) {   // nothing to choose
        if (choice)
            for (int k = 0; k < 7; k++) choice[k] = 0;
        return EGSGPU_OK;
    }
    int rc = ensure_lanes(h);
    if (rc) return rc;
    // the probes must not change the tallies: keep them aside
    const size_t nt = (size_t)h->nvox + kExtraRegions;
    size_t nb = sizeof(double) * nt;
    double *keep = nullptr;
    EGS_HIP(hipMalloc(&keep, 2 * nb));
    auto done = [&](int r) { (void)hipFree(keep); return r; };
    if (hipMemcpyAsync(keep, h->tally.sum, nb, hipMemcpyDeviceToDevice, h->stream) != hipSuccess ||
        hipMemcpyAsync(keep + nt, h->tally.sum2, nb, hipMemcpyDeviceToDevice, h->stream) !=
            hipSuccess) {
        set_error("tally save failed");
        return done(EGSGPU_E_HIP);
    }
    int64_t nh0 = h->nhist;
    unsigned long long cnt[32];
    if (hipMemcpy(cnt, h->tally.counters, sizeof(cnt), hipMemcpyDeviceToHost) != hipSuccess) {
        set_error("counter save failed");
        return done(EGSGPU_E_HIP);
    }
    auto probe = [&](double &ms) -> int {
        int r = egsgpu_run(h, nhist_probe);
        ms = h->last_ms;
        return r;
    };
    double best, ms;
    if ((rc = probe(best))) return done(rc);   // warm-up
    if ((rc = probe(best))) return done(rc);
    // coordinate search: register budget, queue capacity (history slots
    // per stream; the banks hold 8 particles per slot), workgroup size,
    // deferral threshold (a wave hands its particles on when fewer lanes are
    // busy), interaction batch, photon steps per turn
    struct Knob { int *v; std::vector<int> opts; bool slots; };
    int wpe = h->ev_wp;
    int spl = h->nslots > 0 ? std::max(1, h->nslots / h->nlanes) : 4;
    int eb = h->ev_block ? h->ev_block : h->block;
    std::vector<Knob> knobs = {{&wpe, {2, 3}, false},
                               {&spl, {2, 4, 8}, true},
                               {&eb, {128, 256}, false},
                               {&h->ev_defer, {16, 32, 48}, false},
                               {&h->ev_batch, {8, 16, 24}, false},
                               {&h->ev_ksteps, {4, 8, 16}, false}};
    auto apply = [&]() -> int {
        h->ev_wp = h->ev_we = wpe;
        h->ev_block = eb;
        if (h->nslots != spl * h->nlanes) {
            CK(release_event(h));
            h->nslots = spl * h->nlanes;
        }
        return EGSGPU_OK;
    };
    for (auto &k : knobs) {
        int cur = *k.v, pick = cur;
        for (int o : k.opts) {
            if (o == cur) continue;
            *k.v = o;
            if ((rc = apply())) return done(rc);
            if (k.slots && (rc = probe(ms))) return done(rc);   // the new slots' first run
            if ((rc = probe(ms))) return done(rc);
            // a change must gain more than 1 % to count: the probes' noise
            // would otherwise pick settings that are no faster
            if (ms < 0.99 * best) { best = ms; pick = o; }
        }
        *k.v = pick;
        if ((rc = apply())) return done(rc);
    }
    if (hipMemcpyAsync(h->tally.sum, keep, nb, hipMemcpyDeviceToDevice, h->stream) != hipSuccess ||
        hipMemcpyAsync(h->tally.sum2, keep + nt, nb, hipMemcpyDeviceToDevice, h->stream) !=
            hipSuccess ||
        hipMemcpy(h->tally.counters, cnt, sizeof(cnt), hipMemcpyHostToDevice) != hipSuccess) {
        set_error("tally restore failed");
        return done(EGSGPU_E_HIP);
    }
    h->nhist = nh0;
    if (choice) {
        choice[0] = h->ev_wp; choice[1] = h->ev_we; choice[2] = h->ev_defer;
        choice[3] = h->ev_batch; choice[4] = h->ev_ksteps;
        choice[5] = spl; choice[6] = eb;
    }
    return done(EGSGPU_OK);
}

namespace {
// The majorant of Woodcock tracking on a grid of kWoodcockN intervals of
// ln E from the lowest AP to the highest UP of the media present: in each
// interval, factor times the largest attenuation coefficient any voxel can
// have, rho_max(m) / (gmfp_m(E) cohe_m(E)) evaluated exactly as the
// kernels do (the piecewise-linear tables), at the interval's ends and at
// every table node inside it, between which it is monotone.
constexpr int kWoodcockN = 4096;
int build_majorant(egsgpu_handle *h, double factor) {
    const Blob &b = h->tab->blob;
    const int nmed = (int)h->rho_max.size();
    double lo = 1e300, hi = -1e300;
    for (int m = 0; m < nmed; m++) {
        if (h->rho_max[m] <= 0) continue;
        lo = std::min(lo, std::log(b.f1(Blob::mk(m, "ap"))));
        const double ge0 = b.f1(Blob::mk(m, "ge0")), ge1 = b.f1(Blob::mk(m, "ge1"));
        hi = std::max(hi, (kGE - ge0) / ge1);
    }
    if (!(hi > lo)) {
        set_error("no medium with a positive density: no majorant");
        return EGSGPU_E_ARG;
    }
    const double w1 = kWoodcockN / (hi - lo), w0 = 1 - w1 * lo;
    std::vector<double> maj(kWoodcockN, 0.0);
    auto add = [&](int L, double v) {
        L = std::min(std::max(L, 1), kWoodcockN);
        maj[L - 1] = std::max(maj[L - 1], v);
    };
    for (int m = 0; m < nmed; m++) {
        const double rmax = h->rho_max[m];
        if (rmax <= 0) continue;
        const double ge0 = b.f1(Blob::mk(m, "ge0")), ge1 = b.f1(Blob::mk(m, "ge1"));
        const auto &g0 = b.F(Blob::mk(m, "gmfp0")), &g1 = b.F(Blob::mk(m, "gmfp1"));
        const auto &c0 = b.F(Blob::mk(m, "cohe0")), &c1 = b.F(Blob::mk(m, "cohe1"));
        auto mu_at = [&](double x, int node) {   // node: the table interval (1-based)
            node = std::min(std::max(node, 1), kGE);
            double gm = g1[node - 1] * x + g0[node - 1];
            double ch = h->P.rayleigh ? c1[node - 1] * x + c0[node - 1] : 1.0;
            return rmax / (gm * ch);
        };
        auto mu = [&](double x) { return mu_at(x, (int)(ge1 * x + ge0)); };
        for (int L = 1; L <= kWoodcockN; L++) {
            const double xa = (L - w0) / w1, xb = (L + 1 - w0) / w1;
            add(L, mu(xa));
            add(L, mu(xb));
        }
        for (int j = 1; j <= kGE; j++) {   // table nodes, both one-sided limits
            const double x = (j - ge0) / ge1;
            const int L = (int)(w1 * x + w0);
            add(L, mu_at(x, j));
            add(L, mu_at(x, j - 1));
        }
    }
    for (double &v : maj) v *= factor;
    if (h->fp64) {
        std::vector<double> v(maj);
        CK(upload(h, v.data(), v.size(), &h->pd.T.wmaj));
        h->pd.T.wm0 = w0; h->pd.T.wm1 = w1; h->pd.T.wmn = kWoodcockN;
    } else {
        std::vector<float> v(maj.begin(), maj.end());
        CK(upload(h, v.data(), v.size(), &h->pf.T.wmaj));
        h->pf.T.wm0 = (float)w0; h->pf.T.wm1 = (float)w1; h->pf.T.wmn = kWoodcockN;
    }
    return EGSGPU_OK;
}
}  // namespace

int egsgpu_set_photon_tracking(egsgpu_handle *h, int32_t mode, double factor) {
    EGS_ARG(h, "handle");
    EGS_ARG(mode == EGSGPU_TRACK_VOXEL || mode == EGSGPU_TRACK_WOODCOCK, "tracking mode");
    if (mode == EGSGPU_TRACK_VOXEL) {
        h->P.woodcock = 0;
        return EGSGPU_OK;
    }
    if (factor == 0) factor = 1.25;
    EGS_ARG(factor >= 1.0 && factor <= 100.0, "majorant factor 1..100 (0 = 1.25)");
    if (h->howfarless || h->P.n_split) {
        set_error("Woodcock tracking is not combined with HOWFARLESS or photon splitting, "
                  "which step photons along their own paths");
        return EGSGPU_E_UNSUPPORTED;
    }
    CK(build_majorant(h, factor));
    h->woodcock_factor = factor;
    h->P.woodcock = 1;
    return EGSGPU_OK;
}

int egsgpu_run(egsgpu_handle *h, int64_t nhist) {
    EGS_ARG(h && nhist >= 0, "handle, nhist >= 0");
    int rc = ensure_lanes(h);
    if (rc) return rc;
    CK(timing_events(h));
    EGS_HIP(hipEventRecord(h->ev0, h->stream));
    if (h->schedule == EGSGPU_SCHED_EVENT) {
        rc = h->fp64 ? run_event(h, h->pd, nhist) : run_event(h, h->pf, nhist);
        if (rc) return rc;
    } else {
        rc = h->fp64 ? launch(h, h->pd, nhist) : launch(h, h->pf, nhist);
        if (rc) return rc;
    }
    EGS_HIP(hipEventRecord(h->ev1, h->stream));
    EGS_HIP(hipStreamSynchronize(h->stream));
    if (h->schedule != EGSGPU_SCHED_EVENT) CK(take_time(h, 0));
    float ms = 0;
    EGS_HIP(hipEventElapsedTime(&ms, h->ev0, h->ev1));
    h->last_ms = ms;
    h->last_launches = h->klaunch[0] + h->klaunch[1] + h->klaunch[2] + h->klaunch[3];
    h->nhist += nhist;
    return EGSGPU_OK;
}

}  // extern "C"

// ---- egsgpu_run without allocation ---------------------------------------
namespace {
constexpr int kMaxDevices = 64;
std::mutex scratch_mu;
void *scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (taken once per
// device, with that device current).
int scratch_address(void **out) {
    int dev = 0;
    EGS_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("device %d: the static scratch is kept for devices 0..%d", dev, kMaxDevices - 1);
        return EGSGPU_E_UNSUPPORTED;
    }
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_ptr[dev]) {
        void *p = nullptr;
        EGS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) {
            set_error("hipGetSymbolAddress returned no address for the static scratch");
            return EGSGPU_E_HIP;
        }
        scratch_ptr[dev] = p;
    }
    *out = scratch_ptr[dev];
    return EGSGPU_OK;
}

// Bytes of scratch a history-schedule run of h needs: per stream its
// history table (entries and list) and its particle stack (with the entry
// past the stack that photon splitting keeps).
size_t static_bytes(const egsgpu_handle *h, size_t *tab_bytes, size_t *list_bytes) {
    const size_t rsz = h->fp64 ? sizeof(double) : sizeof(float);
    const size_t n = (size_t)h->nlanes;
    *tab_bytes = n * (size_t)h->hcap * sizeof(SlotEntry);
    *list_bytes = (n * (size_t)h->hcap * sizeof(int) + 255) / 256 * 256;
    return *tab_bytes + *list_bytes + n * (size_t)(h->mxstk + 1) * kStkF * rsz;
}

template <typename R>
int run_static(egsgpu_handle *h, Precision<R> &P, int64_t nhist, unsigned char *base) {
    size_t tb = 0, lb = 0;
    static_bytes(h, &tb, &lb);
    Precision<R> Q = P;
    Q.L.hcap = h->hcap;
    Q.L.hent = (SlotEntry *)base;
    Q.L.hlist = (int *)(base + tb);
    Q.L.stk = (R *)(base + tb + lb);
    const size_t words = tb / sizeof(uint4);
    const int blocks = (int)std::min<size_t>((words + 255) / 256, (size_t)h->ncu * 8);
    hipLaunchKernelGGL(k_zero16, dim3(std::max(blocks, 1)), dim3(256), 0, h->stream,
                       (uint4 *)base, words);
    EGS_HIP(hipGetLastError());
    return launch(h, Q, nhist, nullptr, -1, false);
}
}  // namespace

extern "C" {

int egsgpu_run_static(egsgpu_handle *h, int64_t nhist) {
    EGS_ARG(h && nhist >= 0, "handle, nhist >= 0");
    if (!h->lanes_ready) {
        set_error("egsgpu_run_static needs the handle's streams: run egsgpu_run(h, 0) once "
                  "first");
        return EGSGPU_E_ARG;
    }
    if (h->schedule == EGSGPU_SCHED_EVENT) {
        set_error("egsgpu_run_static runs the history schedules; choose one with "
                  "egsgpu_set_schedule before the first run");
        return EGSGPU_E_UNSUPPORTED;
    }
    size_t tb = 0, lb = 0;
    const size_t need = static_bytes(h, &tb, &lb);
    if (need > kScratchBytes) {
        set_error("egsgpu_run_static: %d streams need %zu bytes of scratch, the static array "
                  "holds %zu (use fewer streams: egsgpu_set_streams)",
                  h->nlanes, need, kScratchBytes);
        return EGSGPU_E_ARG;
    }
    void *base = nullptr;
    int rc = scratch_address(&base);
    if (rc) return rc;
    rc = h->fp64 ? run_static(h, h->pd, nhist, (unsigned char *)base)
                 : run_static(h, h->pf, nhist, (unsigned char *)base);
    if (rc) return rc;
    EGS_HIP(hipStreamSynchronize(h->stream));
    h->nhist += nhist;
    return EGSGPU_OK;
}

// ---- checkpoint / restart ----------------------------------------------
// Layout: "EGSGPUCK", u32 version, u32 nvox, u32 nlanes, u32 rng kind,
// i64 nhist, i32 seeds[2] (ixx, jxx or luxury, seed), f64 sum[nvox + 2],
// f64 sum2[nvox + 2] (the dose regions and the energy bookkeeping), then
// the stream records: RanmarRec[nlanes] or RanluxRec[nlanes] (streams
// that have not run yet are stored seeded).  Versions 1 (ranmar only, kind
// word 0) and 2 (sums of the dose regions only; the bookkeeping restarts)
// are still read.
namespace {
constexpr uint32_t kCkVersion = 3;
size_t ck_size(const egsgpu_handle *h, int nlanes, int kind, uint32_t version = kCkVersion) {
    const size_t nt = (size_t)h->nvox + (version >= 3 ? kExtraRegions : 0);
    return 8 + 4 * 4 + 8 + 2 * 4 + 2 * sizeof(double) * nt +
           (kind == EGSGPU_RNG_RANLUX ? sizeof(RanluxRec) : sizeof(RanmarRec)) * (size_t)nlanes;
}
}  // namespace

int egsgpu_checkpoint(egsgpu_handle *h, void *buf, size_t *size) {
    EGS_ARG(h && size, "handle, size");
    int rc = ensure_lanes(h);
    if (rc) return rc;
    size_t need = ck_size(h, h->nlanes, h->rng_kind);
    if (!buf) { *size = need; return EGSGPU_OK; }
    if (*size < need) {
        set_error("checkpoint buffer holds %zu bytes, %zu needed", *size, need);
        *size = need;
        return EGSGPU_E_ARG;
    }
    EGS_HIP(hipStreamSynchronize(h->stream));
    char *p = (char *)buf;
    memcpy(p, "EGSGPUCK", 8); p += 8;
    const bool lux = h->rng_kind == EGSGPU_RNG_RANLUX;
    uint32_t hd[4] = {kCkVersion, (uint32_t)h->nvox, (uint32_t)h->nlanes, (uint32_t)h->rng_kind};
    memcpy(p, hd, sizeof(hd)); p += sizeof(hd);
    memcpy(p, &h->nhist, 8); p += 8;
    int32_t seeds[2] = {lux ? h->lux : h->ixx, lux ? h->lux_seed : h->jxx};
    memcpy(p, seeds, 8); p += 8;
    const size_t nt = (size_t)h->nvox + kExtraRegions;
    EGS_HIP(hipMemcpy(p, h->tally.sum, sizeof(double) * nt, hipMemcpyDeviceToHost));
    p += sizeof(double) * nt;
    EGS_HIP(hipMemcpy(p, h->tally.sum2, sizeof(double) * nt, hipMemcpyDeviceToHost));
    p += sizeof(double) * nt;
    size_t rb = 0;
    void *recs = rng_records(h, &rb);
    EGS_HIP(hipMemcpy(p, recs, rb, hipMemcpyDeviceToHost));
    *size = need;
    return EGSGPU_OK;
}

int egsgpu_restore(egsgpu_handle *h, const void *buf, size_t size) {
    EGS_ARG(h && buf, "handle, buffer");
    const char *p = (const char *)buf;
    if (size < 40 || memcmp(p, "EGSGPUCK", 8) != 0) {
        set_error("not an engine checkpoint");
        return EGSGPU_E_FORMAT;
    }
    uint32_t hd[4];
    memcpy(hd, p + 8, sizeof(hd));
    if (hd[0] < 1 || hd[0] > kCkVersion) {
        set_error("checkpoint version %u", hd[0]);
        return EGSGPU_E_FORMAT;
    }
    const int kind = hd[0] == 1 ? EGSGPU_RNG_RANMAR : (int)hd[3];
    if (kind != EGSGPU_RNG_RANMAR && kind != EGSGPU_RNG_RANLUX) {
        set_error("checkpoint rng kind %d", kind);
        return EGSGPU_E_FORMAT;
    }
    if ((int)hd[1] != h->nvox) {
        set_error("checkpoint has %u voxels, the simulation %d", hd[1], h->nvox);
        return EGSGPU_E_ARG;
    }
    if (h->lanes_ready && (int)hd[2] != h->nlanes) {
        set_error("checkpoint has %u streams, the simulation %d", hd[2], h->nlanes);
        return EGSGPU_E_ARG;
    }
    if (!h->lanes_ready) h->nlanes = (int)hd[2];
    if (size < ck_size(h, (int)hd[2], kind, hd[0])) {
        set_error("checkpoint truncated");
        return EGSGPU_E_FORMAT;
    }
    int rc = ensure_lanes(h);
    if (rc) return rc;
    p += 8 + sizeof(hd);
    memcpy(&h->nhist, p, 8); p += 8;
    int32_t seeds[2];
    memcpy(seeds, p, 8); p += 8;
    h->rng_kind = kind;
    if (kind == EGSGPU_RNG_RANLUX) {
        h->lux = seeds[0];
        h->lux_seed = seeds[1];
    } else {
        h->ixx = seeds[0];
        h->jxx = seeds[1];
    }
    const size_t nt = (size_t)h->nvox + (hd[0] >= 3 ? kExtraRegions : 0);
    if (hd[0] < 3) {
        EGS_HIP(hipMemset(h->tally.sum + h->nvox, 0, sizeof(double) * kExtraRegions));
        EGS_HIP(hipMemset(h->tally.sum2 + h->nvox, 0, sizeof(double) * kExtraRegions));
    }
    EGS_HIP(hipMemcpy(h->tally.sum, p, sizeof(double) * nt, hipMemcpyHostToDevice));
    p += sizeof(double) * nt;
    EGS_HIP(hipMemcpy(h->tally.sum2, p, sizeof(double) * nt, hipMemcpyHostToDevice));
    p += sizeof(double) * nt;
    size_t rb = 0;
    void *recs = rng_records(h, &rb);
    EGS_HIP(hipMemcpy(recs, p, rb, hipMemcpyHostToDevice));
    return EGSGPU_OK;
}

int egsgpu_kernel_time(egsgpu_handle *h, int32_t kernel, double *ms, int32_t *launches) {
    EGS_ARG(h && kernel >= 0 && kernel < 4 && ms && launches, "handle, kernel 0-3, outputs");
    *ms = h->kms[kernel];
    *launches = h->klaunch[kernel];
    return EGSGPU_OK;
}

int egsgpu_density_grad(egsgpu_handle *h, int64_t nhist, const double *target,
                        double *grad_out, double *dose_out, double *loss) {
    EGS_ARG(h && nhist > 0 && target && grad_out, "handle, nhist > 0, target, grad");
    if (!h->P.kerma) {
        set_error("the density gradient is defined for the photon kerma mode (kerma = 1)");
        return EGSGPU_E_UNSUPPORTED;
    }
    if (h->schedule == EGSGPU_SCHED_EVENT) {
        set_error("the density gradient replays histories: it needs the history schedule");
        return EGSGPU_E_UNSUPPORTED;
    }
    if (h->P.n_split || h->P.nbr_split) {
        set_error("the density gradient's score function assumes unsplit photons");
        return EGSGPU_E_UNSUPPORTED;
    }
    if (h->nvox >= (1 << kGradVoxBits)) {
        set_error("the density gradient supports up to %d voxels", (1 << kGradVoxBits) - 1);
        return EGSGPU_E_UNSUPPORTED;
    }
    int rc = ensure_lanes(h);
    if (rc) return rc;
    size_t nv = (size_t)h->nvox;
    if (!h->rng_saved) {
        CK(alloc(h, (size_t)h->nlanes, &h->rng_saved));
        CK(alloc(h, (size_t)h->nlanes, &h->rlx_saved));
        DevGrad &G = h->dgrad;
        G.cap = 2048;
        CK(alloc(h, (size_t)G.cap * h->nlanes, &G.ev));
        CK(alloc(h, (size_t)kGradDepth * h->nlanes, &G.acc));
        double *w = nullptr;
        CK(alloc(h, nv, &w));
        G.w = w;
        CK(alloc(h, nv, &G.g));
        CK(alloc(h, 1, &h->dgrad_dev));
    }
    // forward pass from the current stream states
    size_t rb = 0;
    void *recs = rng_records(h, &rb);
    void *saved = h->rng_kind == EGSGPU_RNG_RANLUX ? (void *)h->rlx_saved : (void *)h->rng_saved;
    EGS_HIP(hipMemcpyAsync(saved, recs, rb, hipMemcpyDeviceToDevice, h->stream));
    CK(egsgpu_reset_tallies(h));
    CK(egsgpu_run(h, nhist));
    std::vector<double> s(nv), w(nv);
    EGS_HIP(hipMemcpy(s.data(), h->tally.sum, sizeof(double) * nv, hipMemcpyDeviceToHost));
    double L = 0;
    std::vector<double> dose(nv);
    for (size_t i = 0; i < nv; i++) {
        dose[i] = s[i] / (double)nhist / h->hmass[i];
        double r = dose[i] - target[i];
        L += r * r;
        w[i] = 2 * r / ((double)nhist * h->hmass[i]);
    }
    // replay of the same histories with the event log
    EGS_HIP(hipMemcpyAsync((void *)h->dgrad.w, w.data(), sizeof(double) * nv, hipMemcpyHostToDevice, h->stream));
    EGS_HIP(hipMemsetAsync(h->dgrad.g, 0, sizeof(double) * nv, h->stream));
    EGS_HIP(hipMemcpyAsync(h->dgrad_dev, &h->dgrad, sizeof(DevGrad), hipMemcpyHostToDevice, h->stream));
    EGS_HIP(hipMemcpyAsync(recs, saved, rb, hipMemcpyDeviceToDevice, h->stream));
    EGS_HIP(hipEventRecord(h->ev0, h->stream));
    rc = h->fp64 ? launch(h, h->pd, nhist, h->dgrad_dev, h->nhist - nhist)
                 : launch(h, h->pf, nhist, h->dgrad_dev, h->nhist - nhist);
    if (rc) return rc;
    EGS_HIP(hipEventRecord(h->ev1, h->stream));
    EGS_HIP(hipStreamSynchronize(h->stream));
    float ms = 0;
    EGS_HIP(hipEventElapsedTime(&ms, h->ev0, h->ev1));
    h->last_ms = ms;
    EGS_HIP(hipMemcpy(grad_out, h->dgrad.g, sizeof(double) * nv, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < nv; i++)
        grad_out[i] -= 2 * (dose[i] - target[i]) * dose[i] / h->hrhor[i];
    if (dose_out) std::copy(dose.begin(), dose.end(), dose_out);
    if (loss) *loss = L;
    return EGSGPU_OK;
}

int egsgpu_get_tallies(egsgpu_handle *h, double *sum, double *sum2, int64_t *nhist) {
    EGS_ARG(h, "handle");
    EGS_HIP(hipStreamSynchronize(h->stream));
    if (sum) EGS_HIP(hipMemcpy(sum, h->tally.sum, sizeof(double) * h->nvox, hipMemcpyDeviceToHost));
    if (sum2) EGS_HIP(hipMemcpy(sum2, h->tally.sum2, sizeof(double) * h->nvox, hipMemcpyDeviceToHost));
    if (nhist) *nhist = h->nhist;
    return EGSGPU_OK;
}

int egsgpu_reset_tallies(egsgpu_handle *h) {
    EGS_ARG(h, "handle");
    const size_t nt = (size_t)h->nvox + kExtraRegions;
    EGS_HIP(hipMemsetAsync(h->tally.sum, 0, sizeof(double) * nt, h->stream));
    EGS_HIP(hipMemsetAsync(h->tally.sum2, 0, sizeof(double) * nt, h->stream));
    EGS_HIP(hipMemsetAsync(h->tally.counters, 0, 256, h->stream));
    EGS_HIP(hipStreamSynchronize(h->stream));
    h->nhist = 0;
    return EGSGPU_OK;
}

int egsgpu_set_tallies(egsgpu_handle *h, const double *sum, const double *sum2,
                       int64_t nhist) {
    EGS_ARG(h && sum && sum2 && nhist >= 0, "handle, sum, sum2, nhist >= 0");
    EGS_HIP(hipMemcpyAsync(h->tally.sum, sum, sizeof(double) * h->nvox, hipMemcpyHostToDevice,
                           h->stream));
    EGS_HIP(hipMemcpyAsync(h->tally.sum2, sum2, sizeof(double) * h->nvox,
                           hipMemcpyHostToDevice, h->stream));
    // the energy bookkeeping restarts with the run that follows
    EGS_HIP(hipMemsetAsync(h->tally.sum + h->nvox, 0, sizeof(double) * kExtraRegions, h->stream));
    EGS_HIP(hipMemsetAsync(h->tally.sum2 + h->nvox, 0, sizeof(double) * kExtraRegions, h->stream));
    EGS_HIP(hipStreamSynchronize(h->stream));
    h->nhist = nhist;
    return EGSGPU_OK;
}

int egsgpu_get_energy_balance(egsgpu_handle *h, double out[4]) {
    EGS_ARG(h && out, "handle, out");
    EGS_HIP(hipStreamSynchronize(h->stream));
    double s[kExtraRegions], s2[kExtraRegions];
    EGS_HIP(hipMemcpy(s, h->tally.sum + h->nvox, sizeof(s), hipMemcpyDeviceToHost));
    EGS_HIP(hipMemcpy(s2, h->tally.sum2 + h->nvox, sizeof(s2), hipMemcpyDeviceToHost));
    for (int k = 0; k < 4; k++) out[k] = 0;
    for (int c = 0; c < kExtraCopies; c++) {   // region 0 escaped, 1 incident
        out[2] += s[c];
        out[3] += s2[c];
        out[0] += s[kExtraCopies + c];
        out[1] += s2[kExtraCopies + c];
    }
    return EGSGPU_OK;
}

int egsgpu_get_counters(egsgpu_handle *h, int64_t counters[8]) {
    EGS_ARG(h && counters, "handle, counters");
    int64_t all[32];
    EGS_HIP(hipMemcpy(all, h->tally.counters, sizeof(all), hipMemcpyDeviceToHost));
    for (int i = 0; i < 8; i++) counters[i] = all[i] + all[8 + i] + all[16 + i] + all[24 + i];
    return EGSGPU_OK;
}

int egsgpu_kernel_counters(egsgpu_handle *h, int32_t kernel, int64_t counters[8]) {
    EGS_ARG(h && counters && kernel >= 0 && kernel < 4, "handle, kernel 0-3, counters");
    EGS_HIP(hipMemcpy(counters, h->tally.counters + 8 * kernel, 64, hipMemcpyDeviceToHost));
    return EGSGPU_OK;
}

int egsgpu_last_kernel_ms(egsgpu_handle *h, double *ms, int32_t *launches) {
    EGS_ARG(h && ms && launches, "handle, outputs");
    *ms = h->last_ms;
    *launches = h->last_launches;
    return EGSGPU_OK;
}

}  // extern "C"
