// egs_gpu_app.cpp -- host application for the voxel dose path.
//
// The egs++ application surface the engine drops into: an
// EGS_AdvancedApplication-like object with initSimulation / runSimulation /
// finishSimulation (egs++/egs_application.cpp:753-802, 839-917, 1048-1085)
// reading the reference's .egsinp blocks:
//   geometry definition   EGS_XYZGeometry (egs_nd_geometry.cpp:1162-1300):
//                         x/y/z-planes or x/y/z-slabs, media input with
//                         "set medium" and "set relative density"
//                         (egs_base_geometry.cpp:742-860)
//   source definition     egs_parallel_beam (egs_parallel_beam.cpp:40-117)
//                         with point/rectangle shape and monoenergetic or
//                         tabulated spectrum (egs_spectra.cpp:482-760)
//   MC transport parameter (egs_advanced_application.cpp:482-639)
//   run control           ncase / nbatch / calculation = first, restart,
//                         analyze or combine (egs_run_control.cpp:64-120,
//                         185-215), with the .egsdat state files of
//                         EGS_Application::outputData / readData /
//                         combineResults (egs_application.cpp:482-675)
//   rng definition        ranmar initial seeds (egs_rndm.cpp:409-463)
//   ausgab object         egs_dose_scoring .3ddose output
//                         (egs_dose_scoring.cpp:416-479);
//                         egs_radiative_splitting (radiative splitting and
//                         Russian roulette, egs_radiative_splitting.cpp:80-132)
// or, for a DOSXYZnrc input file (recognised by its records), the
// DOSXYZnrc front end: phantom records or an .egsphant CT phantom,
// isource 0, 1 and 3 with monoenergetic or .spectrum energies, NCASE and the
// ranmar seeds of record 13, the EGSnrc transport parameter block, and the
// .3ddose of DOSXYZnrc normalised per incident fluence
// (dosxyznrc.mortran:2997-3002, 3100-3130).
// Transport runs through the C ABI of include/egsgpu.h.
//
// usage: egs_gpu_app -i input.egsinp -p tables.egstab [-o output] [--parse-only]
#include <cctype>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <initializer_list>
#include <memory>
#include <sstream>
#include <algorithm>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/egsgpu.h"
#include "dosxyz_input.h"
#include "egs_dat.h"
#include "egs_input.h"
#include "egs_phantom.h"
#include "egs_phsp.h"

using namespace egshost;

namespace {

std::string engine_error() {
    char buf[1024];
    egsgpu_last_error(buf, sizeof(buf));
    return buf;
}

void check(int rc, const char *what) {
    if (rc) throw std::runtime_error(std::string(what) + ": " + engine_error());
}

struct Planes {
    std::vector<double> p;
    // EGS_PlanesT(xo, dx, np): p[j+1] = p[j] + dx (egs_planes.h:264-283)
    static Planes slabs(const std::vector<double> &s) {
        if (s.size() != 3) throw std::runtime_error("slab input needs 3 values");
        int n = (int)(s[2] + 0.1);
        if (n < 1 || s[1] <= 0) throw std::runtime_error("bad slab input");
        Planes r;
        r.p.resize(n + 1);
        r.p[0] = s[0];
        for (int j = 0; j < n; j++) r.p[j + 1] = r.p[j] + s[1];
        return r;
    }
};

}  // namespace

class EGS_GPUApplication {
  public:
    std::string input_file, tables_file, output;
    std::unique_ptr<InputItem> input;
    egsgpu_tables *tab = nullptr;
    egsgpu_handle *h = nullptr;
    Planes xp, yp, zp;
    std::vector<int32_t> medium;
    std::vector<double> rhor;
    std::vector<std::string> media;    // geometry media names
    std::vector<double> spec_e, spec_w;
    EgsSource src{};
    EgsTransport tp{};
    int64_t ncase = 0;
    int nbatch = 10;
    int ixx = 1802, jxx = 9373;
    double area = 0;
    std::string dose_file;
    // run control "calculation": 0 first, 1 restart, 2 analyze, 3 combine
    int calculation = 0;
    // the scoring arrays of the egs_dose_scoring object in the state file:
    // region dose (default yes), dose file (an "output dose file" item)
    bool region_dose = true, dose_file_array = false, medium_dose = false;
    double cpu_before = 0, cpu_run = 0;
    int64_t ndone_before = 0;
    EgsDat state;                 // analyze / combine: the state read
    // DOSXYZnrc front end
    bool dosxyz = false;
    bool parse_only = false;
    int zeroairdose = 0;
    DosxyzInput dz;

    ~EGS_GPUApplication() {
        if (h) egsgpu_destroy(h);
        if (tab) egsgpu_tables_free(tab);
    }

    int parse_args(int argc, char **argv) {
        for (int i = 1; i < argc; i++) {
            std::string a = argv[i];
            auto next = [&]() -> std::string {
                if (i + 1 >= argc) throw std::runtime_error("missing value for " + a);
                return argv[++i];
            };
            if (a == "-i") input_file = next();
            else if (a == "-p") tables_file = next();
            else if (a == "-o") output = next();
            else if (a == "--parse-only") parse_only = true;
            else throw std::runtime_error("unknown argument " + a);
        }
        if (input_file.empty() || (tables_file.empty() && !parse_only))
            throw std::runtime_error(
                "usage: egs_gpu_app -i input.egsinp -p tables.egstab [-o output] [--parse-only]");
        if (output.empty()) {
            output = input_file;
            size_t d = output.rfind(".egsinp");
            if (d != std::string::npos) output = output.substr(0, d);
        }
        return 0;
    }

    void init_geometry() {
        const InputItem *gd = input->block("geometry definition");
        if (!gd) throw std::runtime_error("no geometry definition");
        std::string simname;
        gd->get("simulation geometry", simname);
        const InputItem *g = nullptr;
        for (auto *b : gd->blocks("geometry")) {
            std::string nm;
            b->get("name", nm);
            if (simname.empty() || nm == simname) g = b;
        }
        if (!g) throw std::runtime_error("simulation geometry not found");
        std::string type;
        g->get("type", type);
        if (!InputItem::same(type, "EGS_XYZGeometry"))
            throw std::runtime_error("only EGS_XYZGeometry is supported, got " + type);
        std::string phant;
        if (!g->get("egsphant file", phant)) {
            init_egsphant(g, phant);
            return;
        }
        std::vector<double> v;
        const char *axes[3][2] = {{"x-planes", "x-slabs"}, {"y-planes", "y-slabs"}, {"z-planes", "z-slabs"}};
        Planes *P[3] = {&xp, &yp, &zp};
        for (int a = 0; a < 3; a++) {
            if (!g->get(axes[a][0], v)) P[a]->p = v;
            else if (!g->get(axes[a][1], v)) *P[a] = Planes::slabs(v);
            else throw std::runtime_error(std::string("missing ") + axes[a][0]);
        }
        int nx = (int)xp.p.size() - 1, ny = (int)yp.p.size() - 1, nz = (int)zp.p.size() - 1;
        int nreg = nx * ny * nz;
        medium.assign(nreg, 0);
        rhor.assign(nreg, 1.0);
        const InputItem *mi = g->block("media input");
        if (!mi) throw std::runtime_error("geometry has no media input");
        mi->get("media", media);
        if (media.empty()) throw std::runtime_error("media input lists no media");
        // setMedia (egs_base_geometry.cpp:786-812)
        for (auto &c : mi->children) {
            if (c->is_block) continue;
            if (InputItem::same(c->key, "set medium")) {
                std::string val = c->value;
                for (char &ch : val) if (ch == ',') ch = ' ';
                std::istringstream is(val);
                std::vector<int> s;
                int x;
                while (is >> x) s.push_back(x);
                int start, end, med, delta = 1;
                if (s.size() == 2) { start = end = s[0]; med = s[1]; }
                else if (s.size() == 3) { start = s[0]; end = s[1]; med = s[2]; }
                else if (s.size() == 4) { start = s[0]; end = s[1]; med = s[2]; delta = s[3]; }
                else throw std::runtime_error("set medium needs 2-4 values");
                if (med < 0 || med >= (int)media.size()) throw std::runtime_error("set medium index out of range");
                start = std::max(start, 0);
                end = std::min(end, nreg - 1);
                for (int j = start; j <= end; j += std::max(delta, 1)) medium[j] = med;
            } else if (InputItem::same(c->key, "set relative density")) {
                std::string val = c->value;
                for (char &ch : val) if (ch == ',') ch = ' ';
                std::istringstream is(val);
                std::vector<double> s;
                double x;
                while (is >> x) s.push_back(x);
                int start, end;
                double rho;
                if (s.size() == 2) { start = end = (int)(s[0] + 0.1); rho = s[1]; }
                else if (s.size() == 3) { start = (int)(s[0] + 0.1); end = (int)(s[1] + 0.1); rho = s[2]; }
                else throw std::runtime_error("set relative density needs 2 or 3 values");
                start = std::max(start, 0);
                end = std::min(end, nreg - 1);
                for (int j = start; j <= end; j++) rhor[j] = rho;
            }
        }
    }

    // A file named in the input: as given, else next to the input file.
    std::string resolve(const std::string &f) const {
        if (std::ifstream(f).good()) return f;
        size_t d = input_file.rfind('/');
        if (d != std::string::npos) {
            std::string g = input_file.substr(0, d + 1) + f;
            if (std::ifstream(g).good()) return g;
        }
        return f;
    }

    // EGS_XYZGeometry from an .egsphant file (egs_nd_geometry.cpp:303-772).
    // With a CT ramp, media and relative densities follow from the mass
    // densities (first ramp medium with rho_min <= rho < rho_max, relative
    // density rho/rho_default unless within epsilon of 1); without one,
    // the phantom's own medium characters are used, as DOSXYZnrc does, and
    // the relative density is rho over the medium's nominal density.
    void init_egsphant(const InputItem *g, const std::string &file) {
        EgsPhantom ph = read_egsphant(resolve(file));
        xp.p = ph.xb; yp.p = ph.yb; zp.p = ph.zb;
        size_t nreg = ph.density.size();
        medium.assign(nreg, 0);
        rhor.assign(nreg, 1.0);
        std::string ramp_file;
        if (!g->get("ct ramp", ramp_file)) {
            std::vector<CtRampEntry> ramp = read_ct_ramp(resolve(ramp_file));
            media.clear();
            for (auto &e : ramp) media.push_back(e.name);
            const double eps = 1.0 / (double)(1ULL << 50);
            for (size_t j = 0; j < nreg; j++) {
                double rho = ph.density[j];
                size_t i = 0;
                while (i < ramp.size() && !(rho >= ramp[i].rho_min && rho < ramp[i].rho_max)) i++;
                if (i == ramp.size())
                    throw std::runtime_error("voxel " + std::to_string(j) + " of density " +
                                             std::to_string(rho) +
                                             " has no medium in the CT ramp (vacuum voxels are "
                                             "not supported)");
                medium[j] = (int32_t)i;
                if (ramp[i].rho_def > 0) {
                    double r = rho / ramp[i].rho_def;
                    if (std::fabs(r - 1) > eps) rhor[j] = r;
                }
            }
        } else {
            media = ph.media;
            for (size_t j = 0; j < nreg; j++) medium[j] = ph.medium[j];
            phantom_density = std::move(ph.density);
        }
    }

    void init_source() {
        const InputItem *sd = input->block("source definition");
        if (!sd) throw std::runtime_error("no source definition");
        std::string simname;
        sd->get("simulation source", simname);
        const InputItem *s = nullptr;
        for (auto *b : sd->blocks("source")) {
            std::string nm;
            b->get("name", nm);
            if (simname.empty() || nm == simname) s = b;
        }
        if (!s) throw std::runtime_error("simulation source not found");
        std::string lib;
        s->get("library", lib);
        // egs_transformed_source around another source (translations only)
        double tx = 0, ty = 0, tz = 0;
        if (InputItem::same(lib, "egs_transformed_source")) {
            std::string inner;
            s->get("source name", inner);
            std::vector<double> tr;
            if (const InputItem *t = s->block("transformation")) {
                if (!t->get("translation", tr) && tr.size() == 3) { tx = tr[0]; ty = tr[1]; tz = tr[2]; }
                std::vector<double> rot;
                if (!t->get("rotation", rot) || !t->get("rotation vector", rot))
                    throw std::runtime_error("egs_transformed_source: only translations are supported");
            }
            const InputItem *in = nullptr;
            for (auto *b : sd->blocks("source")) {
                std::string nm;
                b->get("name", nm);
                if (nm == inner) in = b;
            }
            if (!in) throw std::runtime_error("egs_transformed_source: source " + inner + " not found");
            s = in;
            lib.clear();
            s->get("library", lib);
        }
        if (InputItem::same(lib, "egs_phsp_source") || InputItem::same(lib, "iaea_phsp_source")) {
            init_phsp_source(s, InputItem::same(lib, "iaea_phsp_source"));
            src.x0 += tx; src.y0 += ty; src.z += tz;
            return;
        }
        const bool fano = InputItem::same(lib, "egs_fano_source");
        if (!InputItem::same(lib, "egs_parallel_beam") && !fano)
            throw std::runtime_error("only egs_parallel_beam, egs_fano_source, egs_phsp_source and "
                                     "iaea_phsp_source are supported, got " + lib);
        int q = 0;
        s->get("charge", q);
        src.charge = q;
        if (fano) {
            // egs_fano_source (egs_fano_source.cpp input): a box shape, the
            // maximum mass density and an isotropic direction
            src.kind = EGSGPU_SOURCE_FANO;
            if (s->get("max mass density", src.rho_max) || !(src.rho_max > 0))
                throw std::runtime_error("egs_fano_source needs max mass density > 0");
            const InputItem *sh = s->block("shape");
            if (!sh) throw std::runtime_error("egs_fano_source needs a shape");
            std::string st;
            sh->get("type", st);
            if (st.empty()) sh->get("library", st);
            if (!InputItem::same(st, "box") && !InputItem::same(st, "egs_box"))
                throw std::runtime_error("egs_fano_source: only a box shape is supported");
            std::vector<double> bs, tr{0, 0, 0};
            if (sh->get("box size", bs) || (bs.size() != 1 && bs.size() != 3))
                throw std::runtime_error("box shape needs box size (1 or 3 values)");
            if (bs.size() == 1) bs = {bs[0], bs[0], bs[0]};
            if (const InputItem *t = sh->block("transformation")) t->get("translation", tr);
            if (tr.size() != 3) throw std::runtime_error("box translation needs 3 values");
            src.x0 = tr[0] - bs[0] / 2 + tx; src.x1 = tr[0] + bs[0] / 2 + tx;
            src.y0 = tr[1] - bs[1] / 2 + ty; src.y1 = tr[1] + bs[1] / 2 + ty;
            src.z = tr[2] - bs[2] / 2 + tz; src.z1 = tr[2] + bs[2] / 2 + tz;
            src.ux = 0; src.uy = 0; src.uz = 1;
            area = 0;
        }
        std::vector<double> dir{0, 0, 1};
        s->get("direction", dir);
        double nrm = std::sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
        src.ux = dir[0] / nrm; src.uy = dir[1] / nrm; src.uz = dir[2] / nrm;
        const InputItem *sh = s->block("shape");
        if (!sh) throw std::runtime_error("parallel beam needs a shape");
        std::string st;
        sh->get("type", st);
        if (st.empty()) sh->get("library", st);
        std::vector<double> pos;
        if (fano) {
            // the box was read above
        } else if (InputItem::same(st, "point")) {
            if (sh->get("position", pos) || pos.size() != 3) throw std::runtime_error("point shape needs position");
            src.shape = EGSGPU_SHAPE_POINT;
            src.x0 = src.x1 = pos[0]; src.y0 = src.y1 = pos[1]; src.z = pos[2];
            area = 0;
        } else if (InputItem::same(st, "egs_rectangle") || InputItem::same(st, "rectangle")) {
            if (sh->get("rectangle", pos) || pos.size() != 4) throw std::runtime_error("rectangle needs 4 values");
            src.shape = EGSGPU_SHAPE_RECTANGLE;
            src.x0 = std::min(pos[0], pos[2]); src.x1 = std::max(pos[0], pos[2]);
            src.y0 = std::min(pos[1], pos[3]); src.y1 = std::max(pos[1], pos[3]);
            std::vector<double> tr;
            src.z = 0;
            if (const InputItem *t = sh->block("transformation"))
                if (!t->get("translation", tr) && tr.size() == 3) {
                    src.x0 += tr[0]; src.x1 += tr[0]; src.y0 += tr[1]; src.y1 += tr[1]; src.z = tr[2];
                }
            area = (src.x1 - src.x0) * (src.y1 - src.y0);
        } else {
            throw std::runtime_error("unsupported shape " + st);
        }
        if (!fano) { src.x0 += tx; src.x1 += tx; src.y0 += ty; src.y1 += ty; src.z += tz; }
        const InputItem *sp = s->block("spectrum");
        if (!sp) throw std::runtime_error("source needs a spectrum");
        std::string stype;
        sp->get("type", stype);
        if (InputItem::same(stype, "monoenergetic")) {
            src.spectrum = EGSGPU_SPEC_MONO;
            if (sp->get("energy", src.energy)) throw std::runtime_error("monoenergetic spectrum needs energy");
        } else if (InputItem::same(stype, "tabulated spectrum")) {
            int mode = 1;
            if (sp->get("energies", spec_e) || sp->get("probabilities", spec_w))
                throw std::runtime_error("tabulated spectrum needs energies and probabilities");
            if (sp->get("spectrum mode", mode)) sp->get("spectrum type", mode);
            if (mode == 1 || mode == 0) {
                if (spec_e.size() != spec_w.size() + 1 && spec_e.size() == spec_w.size()) {
                    // per-bin upper edges given with a lower first edge missing
                    throw std::runtime_error("histogram spectrum needs nbins+1 energies");
                }
                if (mode == 0)   // counts per bin -> per MeV
                    for (size_t k = 0; k < spec_w.size(); k++) spec_w[k] /= (spec_e[k + 1] - spec_e[k]);
                src.spectrum = EGSGPU_SPEC_HISTOGRAM;
            } else if (mode == 2) {
                src.spectrum = EGSGPU_SPEC_LINE;
            } else {
                throw std::runtime_error("spectrum mode 3 (interpolated) is not supported");
            }
            src.nbins = (int)spec_w.size();
            src.spec_e = spec_e.data();
            src.spec_w = spec_w.data();
        } else {
            throw std::runtime_error("unsupported spectrum type " + stype);
        }
    }

    // egs_phsp_source (egs_phsp_source.cpp:455-560 input; rejectParticle
    // :433-470): particle type and cutout filters give rejected particles
    // weight 0 (their histories score nothing); the phase-space plane is
    // z = 0 unless an egs_transformed_source translates it
    EgsPhsp phsp;
    // iaea_phsp_source (iaea_phsp_source.cpp:225-310 input): the same
    // particle-type and cutout filters on an IAEA file, whose particles
    // carry their z plane
    void init_phsp_source(const InputItem *s, bool iaea) {
        std::string file;
        double zplane = 0;
        if (iaea) {
            if (s->get("iaea phase space file", file))
                throw std::runtime_error("iaea_phsp_source needs an iaea phase space file");
            const std::string hb = resolve(file + ".IAEAheader");   // base name given
            phsp = read_iaea_phsp(hb.substr(0, hb.size() - 11), zplane);
        } else {
            if (s->get("phase space file", file))
                throw std::runtime_error("egs_phsp_source needs a phase space file");
            phsp = read_egsphsp(resolve(file));
        }
        std::string ptype = "all";
        s->get("particle type", ptype);
        int want = 2;   // 2 all, 3 charged, else the charge
        if (InputItem::same(ptype, "photons")) want = 0;
        else if (InputItem::same(ptype, "electrons")) want = -1;
        else if (InputItem::same(ptype, "positrons")) want = 1;
        else if (InputItem::same(ptype, "charged")) want = 3;
        else if (!InputItem::same(ptype, "all"))
            throw std::runtime_error("unknown particle type " + ptype);
        std::vector<double> cut;
        bool cutout = !s->get("cutout", cut) && cut.size() == 4;
        double rec = 0;
        if (!s->get("recycle photons", rec) && rec > 0)
            throw std::runtime_error("phase-space recycling is not supported");
        if (!s->get("recycle electrons", rec) && rec > 0)
            throw std::runtime_error("phase-space recycling is not supported");
        for (int64_t k = 0; k < phsp.n; k++) {
            bool rej = (want < 2 && phsp.q[k] != want) || (want == 3 && !phsp.q[k]) ||
                       ((uint32_t)phsp.latch[k] & 2147483648u);
            if (cutout && (phsp.x[k] < cut[0] || phsp.x[k] > cut[1] || phsp.y[k] < cut[2] ||
                           phsp.y[k] > cut[3]))
                rej = true;
            if (rej) phsp.wt[k] = phsp.wt[k] > 0 ? 0.0f : -0.0f;
        }
        src = EgsSource{};
        src.kind = EGSGPU_SOURCE_PHSP;
        src.spectrum = EGSGPU_SPEC_MONO;
        src.charge = 0;
        src.z = zplane;
        area = 0;
        fprintf(stderr, "phase space: %lld particles (%lld photons), %g incident particles, "
                        "E %g-%g MeV\n", (long long)phsp.n, (long long)phsp.nphot,
                (double)phsp.pinc, (double)phsp.emin, (double)phsp.emax);
    }

    void init_transport() {
        tp = EgsTransport{};
        tp.rayleigh = 1; tp.bound_compton = 3; tp.photoelectron_angle = 1;
        tp.atomic_relaxation = 1; tp.pair_angle = 1; tp.brems_angle = 1;
        tp.spin = 1; tp.exact_bca = 1; tp.estepe = 0.25; tp.ximax = 0.5;
        tp.skin_depth = 3; tp.kerma = 0; tp.fp64 = 0;
        const InputItem *t = input->block("MC transport parameter");
        if (!t) return;
        auto onoff = [&](const char *k, int &dst) {
            std::string v;
            if (!t->get(k, v)) dst = !(InputItem::same(v, "off") || InputItem::same(v, "no") || v == "0");
        };
        t->get("Global ECUT", tp.ecut);
        t->get("Global PCUT", tp.pcut);
        onoff("Spin effects", tp.spin);
        onoff("Photoelectron angular sampling", tp.photoelectron_angle);
        std::string v;
        // the option lists of egs_advanced_application.cpp:497-531
        auto choose = [&](const char *key, std::initializer_list<const char *> opts,
                          std::initializer_list<int> values, int &dst) {
            std::string val;
            if (t->get(key, val)) return;
            auto vi = values.begin();
            for (const char *o : opts) {
                if (InputItem::same(val, o)) { dst = *vi; return; }
                ++vi;
            }
            throw std::runtime_error(std::string(key) + " = " + val + " is not supported");
        };
        // Rayleigh: custom form factors (iraylr 2) are not built
        choose("Rayleigh scattering", {"Off", "On", "no", "yes", "0", "1"}, {0, 1, 0, 1, 0, 1},
               tp.rayleigh);
        // relaxation: On means EADL; the pre-EADL "simple" RELAX is not built
        choose("Atomic relaxations", {"Off", "On", "eadl", "no", "yes", "0", "1"},
               {0, 1, 1, 0, 1, 0, 1}, tp.atomic_relaxation);
        choose("Bound Compton scattering", {"Off", "On", "Simple", "norej"}, {0, 1, 2, 3},
               tp.bound_compton);
        // electron impact ionisation: Off, On (the ik data) or a data set
        // name (setEIIData, egs_advanced_application.cpp:841-859); the blob
        // must have been built with it (egsgpu_create checks)
        if (!t->get("Electron impact ionization", v))
            std::snprintf(tp.eii_xsec, sizeof tp.eii_xsec, "%s", v.c_str());
        // the photon data set; the table blob given with -p must have been
        // built from it (egsgpu_create checks)
        if (!t->get("Photon cross sections", v)) {
            std::string px;
            for (char c : v) px.push_back((char)std::tolower((unsigned char)c));
            if (px != "xcom" && px != "epdl" && px != "si")
                throw std::runtime_error("Photon cross sections = " + v + " is not supported "
                                         "(xcom, epdl or si)");
            std::snprintf(tp.photon_xsec, sizeof tp.photon_xsec, "%s", px.c_str());
        }
        // egs_advanced_application.cpp:552-612: the step algorithm takes
        // EGSnrc / PRESTA-I / PRESTA-II / default (all but PRESTA-I mean
        // PRESTA-II), boundary crossing Exact / PRESTA-I
        if (!t->get("Boundary crossing algorithm", v)) {
            if (InputItem::same(v, "Exact")) tp.exact_bca = 1;
            else if (InputItem::same(v, "PRESTA-I")) tp.exact_bca = 0;
            else throw std::runtime_error("Boundary crossing algorithm = " + v +
                                          ": Exact or PRESTA-I");
        }
        if (!t->get("Electron-step algorithm", v)) {
            if (InputItem::same(v, "PRESTA-I")) tp.step_algorithm = 1;
            else if (InputItem::same(v, "PRESTA-II") || InputItem::same(v, "EGSnrc") ||
                     InputItem::same(v, "default"))
                tp.step_algorithm = 0;
            else throw std::runtime_error("Electron-step algorithm = " + v +
                                          ": PRESTA-II, PRESTA-I, EGSnrc or default");
        }
        if (!t->get("Brems angular sampling", v)) tp.brems_angle = InputItem::same(v, "Simple") ? 0 : 1;
        // egs_advanced_application.cpp:546-551: Off, Simple, KM, Uniform, Blend
        if (!t->get("Pair angular sampling", v)) {
            const char *opt[] = {"Off", "Simple", "KM", "Uniform", "Blend"};
            int k = 0;
            while (k < 5 && !InputItem::same(v, opt[k])) k++;
            if (k == 5) throw std::runtime_error("Pair angular sampling = " + v +
                                                 ": Off, Simple, KM, Uniform or Blend");
            tp.pair_angle = k;
        }
        if (!t->get("Pair cross sections", v)) {
            if (InputItem::same(v, "BH")) tp.pair_xsec = 0;
            else if (InputItem::same(v, "NRC")) tp.pair_xsec = 1;
            else throw std::runtime_error("Pair cross sections = " + v + ": BH or NRC");
        }
        if (!t->get("Brems cross sections", v)) {
            if (InputItem::same(v, "BH")) tp.brems_xsec = 0;
            else if (InputItem::same(v, "NIST")) tp.brems_xsec = 1;
            else if (InputItem::same(v, "NRC")) tp.brems_xsec = 2;
            else throw std::runtime_error("Brems cross sections = " + v +
                                          ": BH, NIST or NRC");
        }
        t->get("ESTEPE", tp.estepe);
        t->get("Ximax", tp.ximax);
        t->get("Skin depth for BCA", tp.skin_depth);
        int kerma = 0;
        if (!t->get("kerma approximation", v)) kerma = InputItem::same(v, "yes") || InputItem::same(v, "on");
        tp.kerma = kerma;
    }

    // egs_radiative_splitting ausgab objects (egs_radiative_splitting.h
    // :110-121, .cpp:80-132): 'splitting = n' with n > 1 splits radiative
    // events n-fold and plays Russian roulette with the charged particles
    // of photon interactions (probability 1/n); n < -1 splits without
    // roulette; 1 and -1 leave the transport alone
    void init_variance_reduction() {
        const InputItem *ad = input->block("ausgab object definition");
        if (!ad) return;
        for (const InputItem *a : ad->blocks("ausgab object")) {
            std::string lib;
            if (a->get("library", lib) || !InputItem::same(lib, "egs_radiative_splitting"))
                continue;
            double fs = 1;
            a->get("splitting", fs);
            int n = (int)fs;
            if (n == 0)
                throw std::runtime_error("egs_radiative_splitting with splitting = 0 (no radiative "
                                         "events) is not supported");
            if (n > 1) { tp.nbr_split = n; tp.play_rr = 1; }
            else if (n < -1) { tp.nbr_split = -n; tp.play_rr = 0; }
            else { tp.nbr_split = 0; tp.play_rr = 0; }
        }
    }

    // ---- DOSXYZnrc front end ---------------------------------------------
    void init_dosxyz() {
        dosxyz = true;
        dz = read_dosxyz_input(input_file);
        if (dz.nmed == 0) {   // CT phantom (dosxyznrc.mortran:1533-1640)
            EgsPhantom ph = read_egsphant(resolve(dz.phantom_file));
            xp.p = ph.xb; yp.p = ph.yb; zp.p = ph.zb;
            media = ph.media;
            medium.assign(ph.medium.begin(), ph.medium.end());
            rhor.assign(medium.size(), 1.0);
            phantom_density = std::move(ph.density);
            zeroairdose = dz.zeroairdose;
        } else {
            xp.p = dz.xb; yp.p = dz.yb; zp.p = dz.zb;
            media = dz.media;
            medium.resize(dz.med.size());
            rhor.assign(dz.med.size(), 1.0);
            phantom_density.assign(dz.med.size(), 0.0);
            for (size_t i = 0; i < dz.med.size(); i++) {
                if (dz.med[i] < 1)
                    throw std::runtime_error("vacuum voxels (medium 0) are not supported");
                medium[i] = dz.med[i] - 1;
                phantom_density[i] = dz.rho[i];   // 0: the medium's default density
            }
        }
        init_dosxyz_source();
        input = parse_input(dz.transport_text);
        init_transport();
        // ECUTIN / PCUTIN win when above the EGSnrc parameters (record 4)
        tp.ecut = std::max(tp.ecut, dz.ecutin);
        tp.pcut = std::max(tp.pcut, dz.pcutin);
        ncase = dz.ncase;
        if (ncase <= 0) throw std::runtime_error("NCASE must be > 0");
        nbatch = 10;   // $NBATCH
        // seeds (dosxyznrc.mortran:1905-1916; DOSXYZnrc is built with ranmar)
        ixx = dz.inseed1;
        jxx = dz.inseed2;
        if (ixx <= 0 || ixx > 31328) ixx = 1802;
        if (jxx <= 0 || jxx > 30081) jxx = 9373;
        if (dz.ihowfarless) {   // only meaningful in a homogeneous phantom
            bool homog = true;
            for (size_t i = 1; i < medium.size() && homog; i++)
                homog = medium[i] == medium[0] && phantom_density[i] == phantom_density[0];
            if (homog) tp.howfarless = 1;
            else fprintf(stderr, "note: HOWFARLESS requested for a phantom that is not "
                                 "homogeneous; every voxel boundary is tracked\n");
        }
        // charged-particle range rejection against voxel boundaries below
        // ESAVE_GLOBAL (dosxyznrc.mortran:1050-1064, 2514-2545)
        if (dz.ireject == 1) {
            if (tp.howfarless)
                fprintf(stderr, "note: range rejection is not applied with HOWFARLESS\n");
            else if (dz.esave_global <= tp.ecut)
                fprintf(stderr, "note: ESAVE_GLOBAL <= ECUT, so no range rejection\n");
            else { tp.ireject = 1; tp.esave = dz.esave_global; }
        }
        // photon splitting with Russian roulette of radiative photons
        // (dosxyznrc.mortran:855-1030, 2557-2562)
        if (dz.n_split > 1) tp.n_split = dz.n_split;
        if (dz.i_phsp_out)
            fprintf(stderr, "note: phase-space output requested; not written\n");
        if (dz.irestart)
            fprintf(stderr, "note: IRESTART = %d requested; the run starts afresh\n", dz.irestart);
        dose_file = output + ".3ddose";
    }

    // srcinit for isource 0, 1 and 3 (srcxyznrc.mortran:2442-2586, 2771-2789)
    void init_dosxyz_source() {
        const double deg = 3.141593 / 180.;   // as in srcinit
        src = EgsSource{};
        src.charge = dz.iqin;
        const int nx = (int)xp.p.size() - 1, ny = (int)yp.p.size() - 1, nz = (int)zp.p.size() - 1;
        const std::vector<double> &s = dz.src;
        if (dz.isource == 3) {   // frontal point source (:1100-1109, 2550-2586)
            double xinl = s[0], xinu = s[1], yinl = s[2], yinu = s[3], ssd = s[4];
            if (xinl < xp.p[0]) xinl = xp.p[0];
            if (xinu <= xinl) xinu = xinl;
            if (xinu > xp.p[nx]) xinu = xp.p[nx];
            if (xinl > xp.p[nx]) xinl = xp.p[nx];
            if (yinl < yp.p[0]) yinl = yp.p[0];
            if (yinu <= yinl) yinu = yinl;
            if (yinu > yp.p[ny]) yinu = yp.p[ny];
            if (yinl > yp.p[ny]) yinl = yp.p[ny];
            if (!(ssd > 0)) throw std::runtime_error("source 3: SSD must be > 0");
            src.kind = EGSGPU_SOURCE_DOSXYZ_POINT;
            src.x0 = xinl; src.x1 = xinu; src.y0 = yinl; src.y1 = yinu; src.z = zp.p[0];
            src.dsource = ssd;
            src.ux = 0; src.uy = 0; src.uz = 1;
            area = (xinu - xinl) * (yinu - yinl);   // beamarea
        } else if (dz.isource == 0) {
            double xinl = s[0], xinu = s[1], yinl = s[2], yinu = s[3];
            double thetax = s[4], thetay = s[5], thetaz = s[6];
            if (xinl < xp.p[0]) xinl = xp.p[0];
            if (xinu <= xinl) xinu = xinl;
            if (xinu > xp.p[nx]) xinu = xp.p[nx];
            if (xinl > xp.p[nx]) xinl = xp.p[nx];
            if (yinl < yp.p[0]) yinl = yp.p[0];
            if (yinu <= yinl) yinu = yinl;
            if (yinu > yp.p[ny]) yinu = yp.p[ny];
            if (yinl > yp.p[ny]) yinl = yp.p[ny];
            double winc = std::cos(thetaz * deg), uinc = std::cos(thetax * deg),
                   vinc = std::cos(thetay * deg);
            if (thetaz > 90. || thetaz < -90. || thetax > 180. || thetax < 0. || thetay > 180. ||
                thetay < 0.)
                throw std::runtime_error("source 0: the angles do not point into the front face");
            double ss = uinc * uinc + vinc * vinc + winc * winc;
            if (ss > 1.001 || ss < 0.009)
                throw std::runtime_error("source 0: direction cosines must have unit sum of squares");
            winc = std::sqrt(1 - uinc * uinc - vinc * vinc);
            src.kind = EGSGPU_SOURCE_DOSXYZ_FRONT;
            src.x0 = xinl; src.x1 = xinu; src.y0 = yinl; src.y1 = yinu; src.z = zp.p[0];
            src.ux = uinc; src.uy = vinc; src.uz = winc;
            area = (xinu - xinl) * (yinu - yinl);
        } else {
            double xiso = s[0], yiso = s[1], ziso = s[2], theta = s[3], phi = s[4];
            double xcol = s[5], ycol = s[6], phicol = s[7];
            double ct = std::cos(theta * deg), st = std::sin(theta * deg);
            double cc = std::cos(phicol * deg), sc = std::sin(phicol * deg);
            double cp = std::cos(phi * deg), sp = std::sin(phi * deg);
            double r[9] = { ct * cp * cc + sp * sc, -ct * cp * sc + sp * cc, -st * cp,
                            ct * sp * cc - cp * sc, -ct * sp * sc - cp * cc, -st * sp,
                           -st * cc,                st * sc,                -ct};
            double d = 0;
            for (int a = 0; a < 2; a++)
                for (int b = 0; b < 2; b++)
                    for (int c = 0; c < 2; c++) {
                        double dx = xiso - xp.p[a ? nx : 0], dy = yiso - yp.p[b ? ny : 0],
                               dzz = ziso - zp.p[c ? nz : 0];
                        d = std::max(d, dx * dx + dy * dy + dzz * dzz);
                    }
            src.kind = EGSGPU_SOURCE_DOSXYZ_BEAM;
            for (int k = 0; k < 9; k++) src.rot[k] = r[k];
            src.iso[0] = xiso; src.iso[1] = yiso; src.iso[2] = ziso;
            src.xcol = xcol; src.ycol = ycol;
            src.dsource = -std::sqrt(d + xcol * xcol + ycol * ycol);
            src.ux = r[2]; src.uy = r[5]; src.uz = r[8];
            area = xcol * ycol;
        }
        if (dz.enflag == 0) {
            src.spectrum = EGSGPU_SPEC_MONO;
            src.energy = dz.ein;
        } else {
            dosxyz_spectrum_cdfinv(resolve(dz.spectrum_file), 1000, spec_e, spec_w, nullptr);
            src.spectrum = EGSGPU_SPEC_CDFINV;
            src.nbins = (int)spec_e.size();
            src.spec_e = spec_e.data();
            src.spec_w = spec_w.data();
        }
    }

    void print_parsed() const {
        int nx = (int)xp.p.size() - 1, ny = (int)yp.p.size() - 1, nz = (int)zp.p.size() - 1;
        printf("format: %s\n", dosxyz ? "dosxyznrc" : "egs++");
        printf("grid: %d %d %d\n", nx, ny, nz);
        printf("x: %.6g %.6g\ny: %.6g %.6g\nz: %.6g %.6g\n", xp.p[0], xp.p[nx], yp.p[0],
               yp.p[ny], zp.p[0], zp.p[nz]);
        printf("media:");
        for (auto &m : media) printf(" %s", m.c_str());
        printf("\n");
        std::vector<long> count(media.size(), 0);
        for (int m : medium) if (m >= 0 && m < (int)count.size()) count[m]++;
        printf("voxels per medium:");
        for (long c : count) printf(" %ld", c);
        printf("\n");
        long ndens = 0;
        for (double d : phantom_density) ndens += d > 0;
        printf("voxels with a density: %ld\n", ndens);
        printf("source: kind %d charge %d spectrum %d energy %.6g area %.6g\n", src.kind,
               src.charge, src.spectrum, src.energy, area);
        printf("source rectangle: %.6g %.6g %.6g %.6g z %.6g\n", src.x0, src.x1, src.y0, src.y1,
               src.z);
        printf("direction: %.9g %.9g %.9g\n", src.ux, src.uy, src.uz);
        printf("dsource: %.9g\n", src.dsource);
        if (src.kind == EGSGPU_SOURCE_PHSP) {
            long long nrej = 0, nq[3] = {0, 0, 0};
            for (int64_t k = 0; k < phsp.n; k++) {
                nrej += phsp.wt[k] == 0;
                nq[phsp.q[k] + 1]++;
            }
            printf("phsp: %lld %lld %lld %lld %lld %.9g\n", (long long)phsp.n, nq[0], nq[1], nq[2],
                   nrej, (double)phsp.pinc);
            printf("phsp plane: %.9g %.9g %.9g\n", src.x0, src.y0, src.z);
            double se = 0, sx = 0, sy = 0, su = 0, sv = 0, sw = 0;
            for (int64_t k = 0; k < phsp.n; k++) {
                se += phsp.e[k]; sx += phsp.x[k]; sy += phsp.y[k];
                su += phsp.u[k]; sv += phsp.v[k]; sw += phsp.wt[k];
            }
            printf("phsp sums: %.9g %.9g %.9g %.9g %.9g %.9g\n", se, sx, sy, su, sv, sw);
        }
        if (src.spectrum == EGSGPU_SPEC_CDFINV && !spec_e.empty()) {
            size_t n = spec_e.size(), first = 0;
            while (first < n && spec_e[first] == spec_e[0]) first++;
            printf("cdfinv: %zu %zu %.9g %.9g %.9g %.9g\n", n, first, spec_e[0], spec_w[0],
                   spec_e[n - 1], spec_w[n - 1]);
        }
        printf("ncase: %lld\nseeds: %d %d\n", (long long)ncase, ixx, jxx);
        printf("ecut: %.6g\npcut: %.6g\n", tp.ecut, tp.pcut);
        printf("transport: rayleigh %d bound_compton %d pe_angle %d relax %d pair_angle %d "
               "brems_angle %d spin %d estepe %.6g ximax %.6g skin %.6g howfarless %d "
               "n_split %d nbr_split %d play_rr %d ireject %d esave %.6g exact_bca %d "
               "step_algorithm %d brems_xsec %d pair_xsec %d photon_xsec %s eii %s\n",
               tp.rayleigh, tp.bound_compton, tp.photoelectron_angle, tp.atomic_relaxation,
               tp.pair_angle, tp.brems_angle, tp.spin, tp.estepe, tp.ximax, tp.skin_depth,
               tp.howfarless, tp.n_split, tp.nbr_split, tp.play_rr, tp.ireject, tp.esave,
               tp.exact_bca, tp.step_algorithm, tp.brems_xsec, tp.pair_xsec,
               tp.photon_xsec[0] ? tp.photon_xsec : "default",
               tp.eii_xsec[0] ? tp.eii_xsec : "off");
    }

    void init_run() {
        if (const InputItem *rc = input->block("run control")) {
            double nc = 0;
            if (rc->get("number of histories", nc)) rc->get("ncase", nc);
            ncase = (int64_t)nc;
            rc->get("nbatch", nbatch);
            std::string calc;
            if (!rc->get("calculation", calc)) {
                const char *kinds[] = {"first", "restart", "analyze", "combine"};
                calculation = -1;
                for (int k = 0; k < 4; k++)
                    if (InputItem::same(calc, kinds[k])) calculation = k;
                if (calculation < 0) throw std::runtime_error("calculation = " + calc);
            }
        }
        if (const InputItem *ad = input->block("ausgab object definition")) {
            for (const InputItem *a : ad->blocks("ausgab object")) {
                std::string lib, v;
                if (a->get("library", lib) || !InputItem::same(lib, "egs_dose_scoring")) continue;
                if (!a->get("region dose", v)) region_dose = !InputItem::same(v, "no");
                if (!a->get("medium dose", v)) medium_dose = InputItem::same(v, "yes");
                dose_file_array = a->item("output dose file") || a->block("output dose file");
            }
        }
        if (ncase <= 0 && calculation < 2)
            throw std::runtime_error("run control needs ncase > 0");
        if (nbatch < 1) nbatch = 1;
        if (const InputItem *r = input->block("rng definition")) {
            std::vector<int> s;
            if (!r->get("initial seeds", s) && s.size() == 2) { ixx = s[0]; jxx = s[1]; }
        }
        dose_file = output + ".3ddose";
    }

    // EGS_AdvancedApplication::initSimulation analogue
    int initSimulation() {
        if (is_dosxyz_input(input_file)) {
            init_dosxyz();
        } else {
            input = parse_input_file(input_file);
            init_geometry();
            init_source();
            init_transport();
            init_variance_reduction();
            init_run();
        }
        if (parse_only) {
            print_parsed();
            return 1;
        }
        check(egsgpu_tables_load(tables_file.c_str(), &tab), "loading tables");
        // map geometry media names to table indices (egsAddMedium)
        int32_t ntab = 0;
        check(egsgpu_tables_nmedia(tab, &ntab), "table media");
        std::vector<int> map(media.size(), -1);
        mdens.assign(media.size(), 1.0);
        for (size_t k = 0; k < media.size(); k++) {
            for (int t = 0; t < ntab; t++) {
                char nm[64];
                double rho;
                check(egsgpu_tables_medium_info(tab, t, nm, sizeof(nm), &rho), "medium info");
                if (media[k] == nm) { map[k] = t; mdens[k] = rho; }
            }
            if (map[k] < 0) throw std::runtime_error("medium " + media[k] + " is not in the table blob");
        }
        std::vector<int32_t> tmed(medium.size());
        for (size_t i = 0; i < medium.size(); i++) tmed[i] = map[medium[i]];
        if (!phantom_density.empty())   // mass densities (.egsphant, DOSXYZnrc records)
            for (size_t i = 0; i < medium.size(); i++)
                if (phantom_density[i] > 0) rhor[i] = phantom_density[i] / mdens[medium[i]];
        if (!dosxyz && calculation >= 2) return load_states();   // no transport
        EgsXYZGeometry g{};
        g.nx = (int)xp.p.size() - 1; g.ny = (int)yp.p.size() - 1; g.nz = (int)zp.p.size() - 1;
        g.xb = xp.p.data(); g.yb = yp.p.data(); g.zb = zp.p.data();
        g.medium = tmed.data();
        g.rhor = rhor.data();
        g.boundary_tolerance = 0;
        if (tp.spin) {
            fprintf(stderr, "note: spin effects requested but the spin data are not available; running with Spin effects = Off\n");
            tp.spin = 0;
        }
        check(egsgpu_create(tab, &g, &src, &tp, &h), "creating the simulation");
        if (src.kind == EGSGPU_SOURCE_PHSP)
            check(egsgpu_set_phsp(h, phsp.n, phsp.x.data(), phsp.y.data(), phsp.u.data(),
                                  phsp.v.data(), phsp.e.data(), phsp.wt.data(), phsp.q.data()),
                  "phase-space particles");
        check(egsgpu_seed(h, ixx, jxx), "seeding");
        if (!dosxyz && calculation == 1) restart();
        return 0;
    }

    // ---- .egsdat state files ----------------------------------------------
    int narrays() const {
        if (medium_dose)
            throw std::runtime_error("state files with medium dose are not supported (the "
                                     "per-medium history sums are not kept)");
        return (region_dose ? 1 : 0) + (dose_file_array ? 1 : 0);
    }
    int nvox() const {
        return ((int)xp.p.size() - 1) * ((int)yp.p.size() - 1) * ((int)zp.p.size() - 1);
    }
    const EgsScoringArray &voxel_array(const EgsDat &d, const std::string &f) const {
        if (d.arrays.empty()) throw std::runtime_error(f + ": no scoring array");
        const EgsScoringArray &a = d.arrays.back();
        if (a.nreg != nvox()) throw std::runtime_error(f + ": scoring array of another size");
        return a;
    }
    // calculation = restart: the engine's own checkpoint (every stream
    // continues exactly) if there is one, else the .egsdat sums, with the
    // streams seeded apart from those of the run that made them
    void restart() {
        std::ifstream ck(output + ".egsgpu", std::ios::binary);
        if (ck) {
            std::vector<char> buf((std::istreambuf_iterator<char>(ck)),
                                  std::istreambuf_iterator<char>());
            check(egsgpu_restore(h, buf.data(), buf.size()), "restoring the engine checkpoint");
            EgsDat d = read_egsdat(output + ".egsdat", narrays());
            cpu_before = d.cpu;
            printf("  restart from %s.egsgpu: %lld histories done\n", output.c_str(),
                   (long long)d.ndone);
            return;
        }
        EgsDat d = read_egsdat(output + ".egsdat", narrays());
        const EgsScoringArray &a = voxel_array(d, output + ".egsdat");
        check(egsgpu_set_tallies(h, a.sum.data(), a.sum2.data(), d.ndone), "setting tallies");
        int is, js;
        sequence_seeds(ixx, jxx, int64_t(1) << 20, is, js);
        check(egsgpu_seed(h, is, js), "seeding");
        cpu_before = d.cpu;
        printf("  restart from %s.egsdat: %lld histories done\n", output.c_str(),
               (long long)d.ndone);
    }
    // the seeds of ranmar sequence s after (ixx, jxx), as createRNG hands
    // them to parallel jobs (egs_rndm.cpp:438-446)
    static void sequence_seeds(int ixx, int jxx, int64_t s, int &is, int &js) {
        if (ixx <= 0 || ixx >= 31328) ixx = 1802;
        if (jxx <= 0 || jxx >= 30081) jxx = 9373;
        int64_t t = (int64_t)jxx + s, carry = (t - 1) / 30080;
        js = (int)(1 + (t - 1) % 30080);
        is = (int)(1 + ((int64_t)ixx - 1 + carry) % 31327);
    }
    // calculation = analyze / combine
    int load_states() {
        const int na = narrays();
        if (calculation == 2) {
            state = read_egsdat(output + ".egsdat", na);
        } else {   // <output>_w<j>.egsdat of parallel jobs j (combineResults)
            int nfile = 0;
            for (int j = 1; j <= 1024; j++) {
                std::string f = output + "_w" + std::to_string(j) + ".egsdat";
                std::ifstream probe(f);
                if (!probe) continue;
                EgsDat d = read_egsdat(f, na);
                voxel_array(d, f);
                printf("  %2d %-30s ncase=%lld\n", ++nfile, f.c_str(), (long long)d.ndone);
                add_egsdat(state, d);
            }
            if (!nfile) throw std::runtime_error("no " + output + "_w<j>.egsdat files to combine");
        }
        voxel_array(state, output + ".egsdat");
        return 0;
    }
    // spectrum sums of the source state: exact for monoenergetic sources,
    // expectations for tabulated spectra
    void spectrum_moments(double &m1, double &m2) const {
        m1 = m2 = 0;
        if (src.spectrum == EGSGPU_SPEC_MONO) { m1 = src.energy; m2 = m1 * m1; return; }
        double w = 0;
        for (size_t k = 0; k < spec_w.size(); k++) {
            if (src.spectrum == EGSGPU_SPEC_HISTOGRAM) {
                double a = spec_e[k], b = spec_e[k + 1], p = spec_w[k] * (b - a);
                w += p; m1 += p * 0.5 * (a + b); m2 += p * (a * a + a * b + b * b) / 3;
            } else if (src.spectrum == EGSGPU_SPEC_LINE) {
                w += spec_w[k]; m1 += spec_w[k] * spec_e[k];
                m2 += spec_w[k] * spec_e[k] * spec_e[k];
            }
        }
        if (w > 0) { m1 /= w; m2 /= w; }
    }
    void write_state(const std::vector<double> &s, const std::vector<double> &s2, int64_t n) {
        EgsDat d;
        d.ndone = n;
        d.cpu = cpu_before + cpu_run;
        d.current_case = n;
        // a ranmar sequence no stream of this engine has used, so that a
        // CPU restart from this file draws fresh numbers
        int is, js;
        sequence_seeds(ixx, jxx, int64_t(1) << 21, is, js);
        ranmar_initial_state(is, js, d);
        d.src_count = n;
        d.spec_count = n;
        double m1, m2;
        spectrum_moments(m1, m2);
        d.sum_E = m1 * (double)n;
        d.sum_E2 = m2 * (double)n;
        d.last_case = n;
        EgsScoringArray a;
        a.nreg = (int)s.size();
        a.current_ncase_short = (int)(n % 65536);
        a.current_ncase = n;
        a.current_ncase_65536 = n - n % 65536;
        a.last.assign(a.nreg, n);
        a.sum = s;
        a.sum2 = s2;
        for (int k = 0; k < narrays(); k++) d.arrays.push_back(a);
        write_egsdat(output + ".egsdat", d);
        if (h) {
            size_t size = 0;
            check(egsgpu_checkpoint(h, nullptr, &size), "checkpoint size");
            std::vector<char> buf(size);
            check(egsgpu_checkpoint(h, buf.data(), &size), "checkpoint");
            std::ofstream ck(output + ".egsgpu", std::ios::binary);
            ck.write(buf.data(), (std::streamsize)size);
        }
        printf("  state written to %s.egsdat%s\n", output.c_str(), h ? " (+ .egsgpu)" : "");
    }

    // runSimulation: nbatch batches of ncase/nbatch histories
    int runSimulation() {
        if (!h) return 0;   // analyze / combine
        int64_t per = std::max<int64_t>(1, ncase / nbatch);
        int64_t done = 0;
        auto t0 = std::chrono::steady_clock::now();
        printf("\n  batch   histories   elapsed (s)\n");
        for (int b = 0; b < nbatch && done < ncase; b++) {
            int64_t n = (b == nbatch - 1) ? ncase - done : std::min(per, ncase - done);
            check(egsgpu_run(h, n), "run");
            done += n;
            double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            printf("  %5d  %10lld  %10.3f\n", b + 1, (long long)done, el);
            cpu_run = el;
        }
        return 0;
    }

    // finishSimulation: .3ddose output (EGS_DoseScoring::outputDoseFile)
    int finishSimulation() {
        int nx = (int)xp.p.size() - 1, ny = (int)yp.p.size() - 1, nz = (int)zp.p.size() - 1;
        int nreg = nx * ny * nz;
        std::vector<double> s(nreg), s2(nreg);
        int64_t n = 0;
        if (h) {
            check(egsgpu_get_tallies(h, s.data(), s2.data(), &n), "tallies");
        } else {   // analyze / combine: the sums of the state files
            const EgsScoringArray &a = voxel_array(state, output + ".egsdat");
            s = a.sum;
            s2 = a.sum2;
            n = state.ndone;
            cpu_before = state.cpu;
        }
        if (!dosxyz && calculation != 2) write_state(s, s2, n);
        double normE = area > 0 ? area : 1.0;   // per fluence for beams with an area
        double normD = 1.602e-10 * normE;
        // DOSXYZnrc: incident fluence counts the beam particles that missed
        // the phantom (nmissm, dosxyznrc.mortran:2998-3001); they are
        // histories with nothing scored
        int64_t nmiss = 0;
        if (dosxyz) {
            int64_t c[8];
            check(egsgpu_get_counters(h, c), "counters");
            nmiss = c[7];
        }
        const int64_t nhist = n;
        n += nmiss;
        std::ofstream f(dose_file);
        f << nx << " " << ny << " " << nz << "\n";
        for (auto *P : {&xp, &yp, &zp}) {
            for (double b : P->p) f << (float)b << " ";
            f << "\n";
        }
        std::vector<float> dose(nreg), unc(nreg);
        double etot = 0;
        for (int i = 0; i < nreg; i++) {
            double r = s[i] / n, r2 = s2[i] / n;
            double dr = r2 - r * r;
            dr = dr > 0 ? std::sqrt(dr / (n - 1)) : 0;
            int iz = i / (nx * ny), iy = (i / nx) % ny, ix = i % nx;
            double vol = (xp.p[ix + 1] - xp.p[ix]) * (yp.p[iy + 1] - yp.p[iy]) * (zp.p[iz + 1] - zp.p[iz]);
            double mass = vol * rhor[i] * medium_density(medium[i]);
            dose[i] = (float)(r * normD / mass);
            unc[i] = (float)(r > 0 ? dr / r : 1.0);
            if (zeroairdose && rhor[i] * medium_density(medium[i]) <= 0.044) {
                dose[i] = 0;   // zeroairdose (dosxyznrc.mortran:3395-3405)
                unc[i] = 1.0;
            }
            etot += r;
        }
        for (float d : dose) f << d << " ";
        f << "\n";
        for (float u : unc) f << u << " ";
        f << "\n";
        printf("\n  histories: %lld   energy deposited per history: %.6g MeV\n", (long long)nhist,
               etot);
        if (nmiss) printf("  beam particles that missed the phantom: %lld\n", (long long)nmiss);
        printf("  dose written to %s\n", dose_file.c_str());
        return 0;
    }

    std::vector<double> mdens;   // nominal density of each geometry medium
    std::vector<double> phantom_density;   // .egsphant mass densities (no CT ramp)
    double medium_density(int m) { return m < (int)mdens.size() ? mdens[m] : 1.0; }
};

int main(int argc, char **argv) {
    try {
        EGS_GPUApplication app;
        app.parse_args(argc, argv);
        if (app.initSimulation()) return 0;   // --parse-only
        app.runSimulation();
        app.finishSimulation();
    } catch (const std::exception &e) {
        fprintf(stderr, "egs_gpu_app: %s\n", e.what());
        return 1;
    }
    return 0;
}
