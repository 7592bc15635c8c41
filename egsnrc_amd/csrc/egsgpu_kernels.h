// egsgpu_kernels.h -- the transport kernels (templates) and their argument
// records.  Each kernel instantiation is compiled in exactly one of the
// egsgpu_kern_*.hip files; the engine declares them extern (the
// EGSGPU_FOR_EACH_KERNEL list below) and launches them by address.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "../../include/egsgpu.h"
#include "egs_tables.h"
#include "egsgpu_common.h"
#include "egsgpu_device.h"
#include "egsgpu_engine.h"
#include "egsgpu_transport.h"

namespace egsgpu {

template <typename R>
struct KArgs {
    DevTables<R> T;
    GeomView<R> G;
    const VoxData<R> *vox;
    DevParams P;
    DevLanes<R> L;
    DevTally tally;
    DevSource<R> S;
    int64_t base;      // histories per lane
    int rem;           // lanes < rem do one more
    const DevGrad *grad;   // non-null: density-gradient replay
    long long hist0;   // index of the launch's first history in the simulation
};

// EGS_AliasTable::sample (egs_alias_table.cpp:268-300), types 0 and 1
template <typename R, bool EV, bool LUX, bool HFL, bool SPLIT>
__device__ __forceinline__ R source_energy(const DevSource<R> &S,
                                          Shower<R, EV, LUX, HFL, SPLIT> &sh) {
    if (S.spectrum == 0) return S.energy;
    if (S.spectrum == EGSGPU_SPEC_CDFINV) {
        // DOSXYZnrc srchst (srcxyznrc.mortran:3440-3446): grid point
        // K = min(DELTAK*r1 + 1, DELTAK), E = CDFINV(K,1) + r2*CDFINV(K,2)
        double r1 = (double)sh.rnd();
        double r2 = (double)sh.rnd();
        double kk = (double)S.np * r1 + 1.0;
        int k = (int)(kk < (double)S.np ? kk : (double)S.np);
        return (R)(S.xi[k - 1] + r2 * S.wi[k - 1]);
    }
    double r1 = (double)sh.rnd();
    double aj = r1 * S.np;
    int j = (int)aj;
    aj -= j;
    if (aj > S.wi[j]) j = S.bin[j];
    if (!S.type) return (R)S.xi[j];
    double x = S.xi[j], dx = S.xi[j + 1] - x;
    double r2 = (double)sh.rnd();
    return (R)(x + dx * r2);
}

// Plane index i with p[i] <= x < p[i+1] (the region search of DOSXYZnrc's
// srchst, srcxyznrc.mortran:3427-3437), clamped to the grid.
template <typename R>
__device__ __forceinline__ int planes_lower(double x, int n, const R *p) {
    int lo = 0, hi = n;   // p[lo] <= x < p[hi] is kept
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if ((double)p[mid] <= x) lo = mid; else hi = mid;
    }
    return lo;
}

// Region index of the first plane j with p[j] <= x <= p[j+1] (the entry
// search of srcxyznrc.mortran:4255-4392, both ends inclusive).
template <typename R>
__device__ __forceinline__ int planes_first(double x, int n, const R *p) {
    int lo = 0, hi = n;   // the first j with x <= p[j+1]
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (x <= (double)p[mid + 1]) hi = mid; else lo = mid + 1;
    }
    return lo < n ? lo : n - 1;
}

// DOSXYZnrc isource 0, srchst (srcxyznrc.mortran:3416-3448): position on
// the front face z = zbound(1) (x drawn, then y, each only if its range
// is not empty), fixed direction, then the energy.
template <typename R, bool EV, bool LUX, bool HFL, bool SPLIT>
__device__ __forceinline__ bool sample_dosxyz_front(const DevSource<R> &S, const GeomView<R> &G,
                                                    Shower<R, EV, LUX, HFL, SPLIT> &sh, Part<R> &q) {
    double x = (double)S.x0, y = (double)S.y0;
    double dx = (double)S.x1 - (double)S.x0, dy = (double)S.y1 - (double)S.y0;
    if (dx != 0) x = (double)S.x0 + (double)sh.rnd() * dx;
    if (dy != 0) y = (double)S.y0 + (double)sh.rnd() * dy;
    R E = source_energy(S, sh);
    int ix = planes_lower(x, G.nx, G.xp), iy = planes_lower(y, G.ny, G.yp);
    q.e = S.charge ? E + (R)0.5109989461 : E;
    q.x = (R)x; q.y = (R)y; q.z = G.zp[0];
    q.u = S.ux; q.v = S.uy; q.w = S.uz;
    q.wt = 1; q.dnear = 0; q.iq = S.charge; q.ir = ix + iy * G.nx;
    return true;
}

// DOSXYZnrc isource 3, srchst (srcxyznrc.mortran:3451-3495): a point
// of the rectangle on the front face, accepted with probability
// (SSD/r)^3 for a point source at distance SSD on the z axis (a pencil for
// a rectangle of zero width), then the energy; direction from the source
// point through it.
template <typename R, bool EV, bool LUX, bool HFL, bool SPLIT>
__device__ __forceinline__ bool sample_dosxyz_point(const DevSource<R> &S, const GeomView<R> &G,
                                                    Shower<R, EV, LUX, HFL, SPLIT> &sh, Part<R> &q) {
    const double ssd = S.dsource;
    double x = (double)S.x0, y = (double)S.y0;
    const double dx = (double)S.x1 - x, dy = (double)S.y1 - y;
    double rxyz = sqrt(ssd * ssd + x * x + y * y);
    if (dx != 0 && dy != 0) {
        for (int ntry = 0; ntry < 1000000; ntry++) {
            x = (double)sh.rnd() * dx + (double)S.x0;
            y = (double)sh.rnd() * dy + (double)S.y0;
            double r1 = (double)sh.rnd();
            rxyz = sqrt(ssd * ssd + x * x + y * y);
            double win = ssd / rxyz;
            if (r1 < win * win * win) break;
        }
    }
    R E = source_energy(S, sh);
    int ix = planes_lower(x, G.nx, G.xp), iy = planes_lower(y, G.ny, G.yp);
    q.e = S.charge ? E + (R)0.5109989461 : E;
    q.x = (R)x; q.y = (R)y; q.z = G.zp[0];
    q.u = (R)(x / rxyz); q.v = (R)(y / rxyz); q.w = (R)(ssd / rxyz);
    q.wt = 1; q.dnear = 0; q.iq = S.charge; q.ir = ix + iy * G.nx;
    return true;
}

// DOSXYZnrc isource 1, srchst (srcxyznrc.mortran:3551-3571, 4157-4165,
// 4255-4392): a point of the collimator rectangle and the energy, rotated
// into position at distance dsource from the isocentre, then carried to
// the first face of the grid it meets; a beam particle that misses the
// grid is counted (sh.c_miss) and sampled again.
template <typename R, bool EV, bool LUX, bool HFL, bool SPLIT>
__device__ __forceinline__ bool sample_dosxyz_beam(const DevSource<R> &S, const GeomView<R> &G,
                                                   Shower<R, EV, LUX, HFL, SPLIT> &sh, Part<R> &q) {
    const double *r = S.rot;
    const double u = r[2], v = r[5], w = r[8];
    const double x0 = G.xp[0], x1 = G.xp[G.nx], y0 = G.yp[0], y1 = G.yp[G.ny];
    const double z0 = G.zp[0], z1 = G.zp[G.nz];
    for (int ntry = 0; ntry < 100000; ntry++) {
        double xs = ((double)sh.rnd() - 0.5) * S.xcol;
        double ys = ((double)sh.rnd() - 0.5) * S.ycol;
        R E = source_energy(S, sh);
        double xp = r[0] * xs + r[1] * ys + r[2] * S.dsource + S.iso[0];
        double yp = r[3] * xs + r[4] * ys + r[5] * S.dsource + S.iso[1];
        double zp = r[6] * xs + r[7] * ys + r[8] * S.dsource + S.iso[2];
        double xi = 0, yi = 0, zi = 0;
        int ix = -1, iy = -1, iz = -1;
        if (u != 0) {
            double xf = u > 0 ? x0 : x1;
            double d = (xf - xp) / u;
            double yy = yp + d * v, zz = zp + d * w;
            if (yy >= y0 && yy <= y1 && zz >= z0 && zz <= z1) {
                xi = xf; yi = yy; zi = zz;
                ix = u > 0 ? 0 : G.nx - 1;
                iy = planes_first(yy, G.ny, G.yp);
                iz = planes_first(zz, G.nz, G.zp);
            }
        }
        if (ix < 0 && v != 0) {
            double yf = v > 0 ? y0 : y1;
            double d = (yf - yp) / v;
            double xx = xp + d * u, zz = zp + d * w;
            if (xx >= x0 && xx <= x1 && zz >= z0 && zz <= z1) {
                xi = xx; yi = yf; zi = zz;
                iy = v > 0 ? 0 : G.ny - 1;
                ix = planes_first(xx, G.nx, G.xp);
                iz = planes_first(zz, G.nz, G.zp);
            }
        }
        if (ix < 0 && w != 0) {
            double zf = w > 0 ? z0 : z1;
            double d = (zf - zp) / w;
            double xx = xp + d * u, yy = yp + d * v;
            if (xx >= x0 && xx <= x1 && yy >= y0 && yy <= y1) {
                xi = xx; yi = yy; zi = zf;
                iz = w > 0 ? 0 : G.nz - 1;
                ix = planes_first(xx, G.nx, G.xp);
                iy = planes_first(yy, G.ny, G.yp);
            }
        }
        if (ix < 0) { sh.c_miss++; continue; }
        q.e = S.charge ? E + (R)0.5109989461 : E;
        q.x = (R)xi; q.y = (R)yi; q.z = (R)zi;
        q.u = (R)u; q.v = (R)v; q.w = (R)w;
        q.wt = 1; q.dnear = 0; q.iq = S.charge;
        q.ir = ix + iy * G.nx + iz * G.nxy;
        return true;
    }
    return false;
}

// Phase-space particle hidx mod n (egs_phsp_source.cpp:285-323): no random
// numbers; entry into the grid along the particle's direction.
template <typename R>
__device__ __forceinline__ bool sample_phsp(const DevSource<R> &S, const GeomView<R> &G,
                                            long long hidx, Part<R> &q) {
    const long long k = hidx % S.nphsp;
    R x = (R)S.px[k] + S.x0, y = (R)S.py[k] + S.y0, z = S.z;
    R u = (R)S.pu[k], v = (R)S.pv[k];
    R aux = u * u + v * v;
    aux = aux < (R)1 ? dsqrt((R)1 - aux) : (R)0;
    const float wts = S.pw[k];
    R w = wts > 0 ? aux : -aux;
    int ireg = geom_is_where(G, x, y, z);
    if (ireg < 0) {
        R t = (R)1e30;
        ireg = geom_howfar_out(G, x, y, z, u, v, w, t);
        if (ireg >= 0) { x += u * t; y += v * t; z += w * t; }
    }
    const int iq = S.pq[k];
    const R E = (R)S.pe[k];
    q.e = iq ? E + (R)0.5109989461 : E;
    q.x = x; q.y = y; q.z = z;
    q.u = u; q.v = v; q.w = w;
    q.wt = (R)fabsf(wts); q.dnear = 0; q.iq = iq; q.ir = ireg;
    return ireg >= 0 && q.wt > 0;
}

// One primary of EGS_Application::simulateSingleShower
// (egs_application.cpp:919-959) from the parallel-beam source
// (egs_parallel_beam.cpp:40-117): position on the source shape, entry into
// the grid along the beam direction.  Returns false if it misses the grid.
// The DOSXYZnrc sources (isource 0 and 1) have their own sampling; a
// phase-space source takes particle hidx (the history's index in the
// simulation).
// egs_fano_source (egs_fano_source.h:128-182, with the energy drawn first
// as EGS_BaseSimpleSource::getNextParticle does, egs_base_source.h:684-692):
// a point uniform in the box, drawn again while it lies outside the grid,
// kept with probability rho/rho_max (rho the mass density of its voxel:
// the medium's density times the voxel's relative density), so that
// starting points are distributed in proportion to mass; then an
// isotropic direction (cos theta uniform in [-1, 1], phi uniform).
template <typename R, bool EV, bool LUX, bool HFL, bool SPLIT>
__device__ __forceinline__ bool sample_fano(const DevSource<R> &S, const GeomView<R> &G,
                                            const VoxData<R> *vox,
                                            Shower<R, EV, LUX, HFL, SPLIT> &sh, Part<R> &q) {
    const R RM = (R)0.5109989461;
    const R E = source_energy(S, sh);
    R x = 0, y = 0, z = 0;
    int ireg = -1;
    for (int ntry = 0; ntry < 1000000; ntry++) {
        x = S.x0 + (S.x1 - S.x0) * sh.rnd();
        y = S.y0 + (S.y1 - S.y0) * sh.rnd();
        z = S.z + (S.z1 - S.z) * sh.rnd();
        const int ir = geom_is_where(G, x, y, z);
        if (ir < 0) continue;
        const VoxData<R> d = vox[ir];
        const R rho = sh.T.ms[d.med].rho * d.rho;
        if (!(sh.rnd() * S.rho_max > rho)) { ireg = ir; break; }
    }
    if (ireg < 0) return false;
    const R w = 1 - 2 * sh.rnd();
    R u = 0, v = 0, sinz = 1 - w * w;
    if (sinz > (R)1e-10) {
        sinz = dsqrt(sinz);
        const R phi = (R)6.283185307179586 * sh.rnd();
        u = sinz * cos(phi);
        v = sinz * sin(phi);
    }
    q.e = S.charge ? E + RM : E;
    q.x = x; q.y = y; q.z = z;
    q.u = u; q.v = v; q.w = w;
    q.wt = 1; q.dnear = 0; q.iq = S.charge; q.ir = ireg;
    return true;
}

template <typename R, bool EV, bool LUX, bool HFL, bool SPLIT>
__device__ __forceinline__ bool sample_primary(const DevSource<R> &S, const GeomView<R> &G,
                                               const VoxData<R> *vox,
                                               Shower<R, EV, LUX, HFL, SPLIT> &sh, Part<R> &q,
                                               long long hidx) {
    if (S.kind == EGSGPU_SOURCE_FANO) return sample_fano(S, G, vox, sh, q);
    if (S.kind == EGSGPU_SOURCE_DOSXYZ_FRONT) return sample_dosxyz_front(S, G, sh, q);
    if (S.kind == EGSGPU_SOURCE_DOSXYZ_BEAM) return sample_dosxyz_beam(S, G, sh, q);
    if (S.kind == EGSGPU_SOURCE_DOSXYZ_POINT) return sample_dosxyz_point(S, G, sh, q);
    if (S.kind == EGSGPU_SOURCE_PHSP) return sample_phsp(S, G, hidx, q);
    const R RM = (R)0.5109989461;
    int ireg = -1;
    for (int ntry = 0; ntry < 100000 && ireg < 0; ntry++) {
        R E = source_energy(S, sh);
        R x = S.x0, y = S.y0, z = S.z;
        if (S.shape == 1) {
            x = S.x0 + (S.x1 - S.x0) * sh.rnd();
            y = S.y0 + (S.y1 - S.y0) * sh.rnd();
        }
        ireg = geom_is_where(G, x, y, z);
        if (ireg < 0) {
            R t = (R)1e30;
            ireg = geom_howfar_out(G, x, y, z, S.ux, S.uy, S.uz, t);
            if (ireg >= 0) { x += S.ux * t; y += S.uy * t; z += S.uz * t; }
        }
        q.e = S.charge ? E + RM : E;
        q.x = x; q.y = y; q.z = z;
        q.u = S.ux; q.v = S.uy; q.w = S.uz;
        q.wt = 1; q.dnear = 0; q.iq = S.charge; q.ir = ireg;
    }
    return ireg >= 0;
}

// Sum over the 64 lanes of the wave (all lanes must be active).
__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
    for (int off = 32; off > 0; off >>= 1) {
        unsigned lo = __shfl_xor((unsigned)v, off);
        unsigned hi = __shfl_xor((unsigned)(v >> 32), off);
        v += ((unsigned long long)hi << 32) | lo;
    }
    return v;
}

// Add this wave's event counters to the kernel's totals: one atomic per
// counter and wave.  Called by all lanes of every wave.
template <typename R, bool EV, bool LUX, bool HFL, bool SPLIT>
__device__ __forceinline__ void add_counters(const DevTally &t,
                                             const Shower<R, EV, LUX, HFL, SPLIT> &sh,
                                             int kernel) {
    unsigned long long *c = t.counters + 8 * kernel;
    const unsigned v[8] = {sh.c_psteps, sh.c_cross, sh.c_dep, sh.c_flush,
                           sh.c_esteps, sh.c_ovf, sh.c_hovf, sh.c_miss};
#pragma unroll
    for (int k = 0; k < 8; k++) {
        unsigned long long w = wave_sum(v[k]);
        if (__lane_id() == 0 && w) atomicAdd(c + k, w);
    }
}

// History schedule: lane s runs its share of the histories one after the
// other with ranmar stream s, like egs-parallel job s.  SPLIT: every photon
// goes through DOSXYZnrc's photon splitting.
// Resident waves per SIMD the history kernels are compiled for: 2 (a
// 256-register budget with a few hundred bytes of scratch) ran the slab
// workload 1.45x and HOWFARLESS 1.6x faster than 1 (344 registers, no
// scratch); 3 was slower than either.
constexpr int kHistWaves = 2;

template <typename R, bool LUX, bool HFL = false, bool SPLIT = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kHistWaves)))
k_transport(KArgs<R> a) {
    extern __shared__ float rbuf_all[];
    int lane = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = lane < a.L.nlanes;   // lanes past the last stream idle
    using S = Shower<R, false, LUX, HFL, SPLIT>;
    S sh(a.T, a.G, a.vox, a.P, a.L, a.tally, live ? lane : 0, rbuf_all + threadIdx.x);
    sh.grad = a.grad;
    int64_t nh = live ? a.base + (lane < a.rem ? 1 : 0) : 0;
    for (int64_t k = 0; k < nh; k++) {
        Part<R> q;
        if (!sample_primary(a.S, a.G, a.vox, sh, q, a.hist0 + k * a.L.nlanes + lane)) continue;
        sh.p = q;
        sh.np = 1;
        sh.eph = S::kNew;
        sh.score_extra(S::kIncident, q);
        while (sh.np > 0) {
            sh.rtopup();
            if (sh.p.iq == 0) {
                if constexpr (SPLIT) {
                    if (a.P.n_split > 1) sh.photon_split();
                    else sh.photon();
                } else {
                    sh.photon();
                }
                sh.eph = S::kNew;
            } else if (sh.eph >= S::kBrems) {
                sh.electron_interact();
            } else {
                sh.electron_step();
            }
        }
        if (a.grad) sh.grad_sweep();
        sh.end_history();
    }
    add_counters(a.tally, sh, 0);
}

// Wave-level choice of the next action.  Lanes want either the common
// action `common` (photon flight / condensed-history step) or one of the
// rare, expensive ones (interactions, codes lo..hi).  A rare action runs
// when it has at least `batch` lanes waiting, when nobody wants the common
// one, or when a lane has waited `max_wait` turns; otherwise the common
// action runs.  The result is the same for every lane of the wave.
__device__ __forceinline__ int wave_choose(int act, int wait, int common, int lo, int hi,
                                           int batch, int max_wait) {
    unsigned long long starving = __ballot(act != common && act >= 0 && wait >= max_wait);
    if (starving) {
        int src = __ffsll(starving) - 1;
        return __builtin_amdgcn_readfirstlane(__shfl(act, src));
    }
    int ncommon = __popcll(__ballot(act == common));
    int best = common, bestn = 0;
    for (int t = lo; t <= hi; t++) {
        int c = __popcll(__ballot(act == t));
        if (c > bestn) { bestn = c; best = t; }
    }
    if (bestn > 0 && (ncommon == 0 || bestn >= batch)) return best;
    return common;
}

#ifndef EGSGPU_KERMA_WAVES
#define EGSGPU_KERMA_WAVES 2
#endif
// resident waves per SIMD the kerma-mode kernel is compiled for (its
// register budget); the default stream count fills them
constexpr int kKermaWaves = EGSGPU_KERMA_WAVES;

// History schedule in kerma mode (ECUT above every electron energy, so a
// charged particle always takes the ECUT discard of ELECTR and only
// photons are transported).  Each lane runs exactly the histories and the
// random-number sequence of k_transport; what changes is how the lanes of
// a wave share the instruction stream: every turn the wave votes for one
// action -- ksteps voxel steps of the photon flights (the common action),
// one interaction type, the start of new histories or the end-of-history
// flush -- and only the lanes that want it run it, so that flights of
// different lengths and the rare, long samplers do not serialise the
// wave.  Cheap phase changes (photon entry, new energy, the charged
// discard) run at once.
template <typename R, bool LUX>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kKermaWaves)))
k_kerma(KArgs<R> a, int ksteps, int batch, int max_wait) {
    extern __shared__ float rbuf_all[];
    int lane = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = lane < a.L.nlanes;
    using S = Shower<R, false, LUX>;
    S sh(a.T, a.G, a.vox, a.P, a.L, a.tally, live ? lane : 0, rbuf_all + threadIdx.x);
    sh.grad = a.grad;
    const int64_t nh = live ? a.base + (lane < a.rem ? 1 : 0) : 0;
    enum { kNewHist = S::kPPair + 1, kEndHist = S::kPPair + 2 };
    int64_t k = 0;          // histories of this lane started
    bool inhist = false;    // a history's particles are (or were) on the stack
    int wait = 0;
    for (;;) {
        sh.rtopup();
        int act;
        for (;;) {   // cheap steps until the lane wants a voted action
            if (sh.np == 0) {
                act = inhist ? (int)kEndHist : (k < nh ? (int)kNewHist : -1);
                break;
            }
            if (sh.p.iq != 0) {
                sh.c_esteps++;
                sh.ecut_discard(sh.p.iq, sh.p.e);
                sh.pph = S::kPEnter;
                continue;
            }
            if (sh.pph == S::kPEnter) { sh.photon_enter(); continue; }
            if (sh.pph == S::kPNewE) { sh.photon_newe(); continue; }
            act = sh.pph;
            break;
        }
        if (__all(act < 0)) break;
        int chosen = wave_choose(act, wait, S::kPTrans, S::kPRayl, kEndHist, batch, max_wait);
        if (act != chosen) {
            if (act >= 0) wait++;
            continue;
        }
        wait = 0;
        if (chosen == S::kPTrans) {
            sh.photon_trans(ksteps);
        } else if (chosen == kNewHist) {
            Part<R> q;
            const bool hit = sample_primary(a.S, a.G, a.vox, sh, q, a.hist0 + k * a.L.nlanes + lane);
            k++;
            if (hit) {
                sh.p = q;
                sh.np = 1;
                sh.eph = S::kNew;
                sh.pph = S::kPEnter;
                sh.score_extra(S::kIncident, q);
                inhist = true;
            }
        } else if (chosen == kEndHist) {
            if (a.grad) sh.grad_sweep();
            sh.end_history();
            inhist = false;
        } else {
            sh.photon_interact();
        }
    }
    add_counters(a.tally, sh, 0);
}

// ---- event schedule ------------------------------------------------------
template <typename R>
struct EvArgs {
    KArgs<R> k;
    DevSlots slots;
    EvCtl *ctl;
    // the bank this launch drains, the bank particles of the other type go
    // to, and the bank particles deferred at the end of the launch go to
    DevBank<R> cons, push, defr;
    int ic, ip, id;                 // their indices in EvCtl
    unsigned long long free_head;   // first free-ring entry for new histories
    int ninject;                    // new histories of this source launch
    int ksteps;                     // photon voxel steps per turn
    int batch, max_wait;            // interaction batching (wave_choose)
    int defer_below;                // defer when fewer lanes of a wave are busy
    // the simulations of this run (one, or a batch); always in device
    // memory, so that a pointer to the current job's geometry is never a
    // pointer into the kernel arguments (which would move them to scratch)
    const JobDesc<R> *jobs;
    int *slot_job;                  // [nslots] job of the history in each slot
    const long long *job_first;     // [njobs+1] first history index of each job
    int njobs;
    long long hist_base;            // index of this source launch's first history
    int gated;                      // graph-driven: return at once when ctl->done
};

// Point the shower at the simulation of the history in `slot`.
template <typename R, bool LUX>
__device__ __forceinline__ void ev_job(Shower<R, true, LUX> &sh, const EvArgs<R> &a, int slot) {
    const JobDesc<R> &J = a.jobs[a.njobs > 1 ? a.slot_job[slot] : 0];
    sh.Gp = &J.G;
    sh.vox = J.vox;
    sh.tsum = J.sum;
    sh.tsum2 = J.sum2;
}

template <typename R>
__device__ __forceinline__ void bank_store(const DevBank<R> &b, int i, const Part<R> &q, int slot) {
    b.e[i] = q.e; b.x[i] = q.x; b.y[i] = q.y; b.z[i] = q.z;
    b.u[i] = q.u; b.v[i] = q.v; b.w[i] = q.w; b.wt[i] = q.wt; b.dn[i] = q.dnear;
    b.ir[i] = q.ir; b.tag[i] = slot * 4 + q.iq + 1;
}

template <typename R>
__device__ __forceinline__ int bank_load(const DevBank<R> &b, int i, Part<R> &q) {
    q.e = b.e[i]; q.x = b.x[i]; q.y = b.y[i]; q.z = b.z[i];
    q.u = b.u[i]; q.v = b.v[i]; q.w = b.w[i]; q.wt = b.wt[i]; q.dnear = b.dn[i];
    q.ir = b.ir[i];
    int tag = b.tag[i];
    q.iq = (tag & 3) - 1;
    return tag >> 2;
}

// Move the particle on top of the lane's stack into bank b (count *count)
// with loop state ph, s[].  The history's pending count goes up before the
// particle becomes visible.  A full bank absorbs the particle where it is
// (counted as an overflow).
template <typename R, bool LUX>
__device__ __forceinline__ bool bank_try(Shower<R, true, LUX> &sh, const DevBank<R> &b, int *count, int ph,
                         R s0, R s1, R s2, R s3) {
    int i = wave_claim(count);
    if (i >= b.cap) return false;
    bank_store(b, i, sh.p, sh.slot);
    b.ph[i] = ph;
    if (ph) { b.s0[i] = s0; b.s1[i] = s1; b.s2[i] = s2; b.s3[i] = s3; }
    return true;
}
template <typename R, bool LUX>
__device__ __forceinline__ void bank_absorb(Shower<R, true, LUX> &sh) {
    atomicSub(sh.slots->pending + sh.slot, 1);
    sh.c_ovf++;
    const R RM = (R)0.5109989461;
    R edep = sh.p.iq == 0 ? sh.p.e : (sh.p.iq < 0 ? sh.p.e - RM : sh.p.e + RM);
    sh.score(sh.p.ir, edep, sh.p.wt);
}
template <typename R, bool LUX>
__device__ __forceinline__ void bank_push(Shower<R, true, LUX> &sh, const DevBank<R> &b, int *count, int ph = 0,
                          R s0 = 0, R s1 = 0, R s2 = 0, R s3 = 0) {
    atomicAdd(sh.slots->pending + sh.slot, 1);
    if (!bank_try(sh, b, count, ph, s0, s1, s2, s3)) bank_absorb(sh);
}
// A particle of the other type (a photon met by the charged-particle
// kernel, or the reverse) goes to the bank of its type; when that bank is
// full it waits, fresh, in this launch's deferral bank, whose next drain
// passes it on (a drained bank of the other type has room again).
template <typename R, bool LUX>
__device__ __forceinline__ void bank_push_other(Shower<R, true, LUX> &sh, const EvArgs<R> &a) {
    atomicAdd(sh.slots->pending + sh.slot, 1);
    if (bank_try(sh, a.push, &a.ctl->n[a.ip], 0, (R)0, (R)0, (R)0, (R)0)) return;
    if (bank_try(sh, a.defr, &a.ctl->n[a.id], 0, (R)0, (R)0, (R)0, (R)0)) return;
    bank_absorb(sh);
}

template <typename R, bool LUX>
__device__ __forceinline__ void ev_setup(Shower<R, true, LUX> &sh, const EvArgs<R> &a) {
    sh.slots = &a.slots;
    sh.ev_ctl = a.ctl;
    sh.np = 0;
    sh.slot = -1;
}

// New histories into free slots: thread i samples primaries i, i+nlanes, ...
// into bank `push` (photon or charged, by the source).
template <typename R, bool LUX>
__global__ void __launch_bounds__(256) k_ev_source(EvArgs<R> a) {
    extern __shared__ float rbuf_all[];
    int lane = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = lane < a.k.L.nlanes;
    Shower<R, true, LUX> sh(a.k.T, a.jobs[0].G, a.jobs[0].vox, a.k.P, a.k.L, a.k.tally,
                            live ? lane : 0, rbuf_all + threadIdx.x);
    if (a.gated && a.ctl->done) return;
    ev_setup(sh, a);
    const unsigned long long ns = (unsigned long long)a.slots.nslots;
    // a negative count: the plan made on the device by k_ev_plan
    const int ninject = a.ninject >= 0 ? a.ninject : a.ctl->plan_n;
    const unsigned long long free_head = a.ninject >= 0 ? a.free_head : a.ctl->plan_head;
    const long long hist_base = a.ninject >= 0 ? a.hist_base : a.ctl->plan_base;
    for (int i = lane; live && i < ninject; i += a.k.L.nlanes) {
        int slot = a.slots.ring[(free_head + (unsigned long long)i) % ns];
        Part<R> q;
        int lo = 0;
        if (a.njobs > 1) {   // the job of history hist_base + i
            long long hidx = hist_base + i;
            int hi = a.njobs;
            while (hi - lo > 1) {
                int mid = (lo + hi) >> 1;
                if (a.job_first[mid] <= hidx) lo = mid; else hi = mid;
            }
            a.slot_job[slot] = lo;
        }
        const DevSource<R> *src = &a.jobs[lo].S;
        const GeomView<R> *geo = &a.jobs[lo].G;
        // the history's index in its own simulation
        const long long hsim = a.njobs > 1 ? a.jobs[lo].hist0 + hist_base + i - a.job_first[lo]
                                           : a.k.hist0 + hist_base + i;
        if (!sample_primary(*src, *geo, a.jobs[lo].vox, sh, q, hsim)) {
            unsigned long long at = atomicAdd(&a.ctl->free_tail, 1ull);
            a.slots.ring[at % ns] = slot;
            continue;
        }
        atomicExch(a.slots.pending + slot, 1);
        // the primary's energy into the history's table (incident region)
        sh.slot = slot;
        ev_job(sh, a, slot);
        sh.score_extra(Shower<R, true, LUX>::kIncident, q);
        if (sh.cvox >= 0) sh.slot_add(sh.cvox, sh.cval);
        sh.cvox = -1;
        sh.cval = 0;
        sh.slot = -1;
        int j = wave_claim(&a.ctl->n[a.ip]);
        if (j < a.push.cap) {
            bank_store(a.push, j, q, slot);
            a.push.ph[j] = 0;
        } else {   // the host leaves room for every new primary
            sh.c_ovf++;
            atomicExch(a.slots.pending + slot, 0);
            unsigned long long at = atomicAdd(&a.ctl->free_tail, 1ull);
            a.slots.ring[at % ns] = slot;
        }
    }
    add_counters(a.k.tally, sh, 1);
}

// Photon kernel: every lane takes photons from bank `cons` until it is
// empty.  Flights run ksteps voxel steps per turn; the interactions are
// batched by type over the wave; charged secondaries go to bank `push`.
// Once the bank is empty and fewer than defer_below lanes of a wave are
// still busy, they hand their photons (with the loop state) to bank `defr`
// for the next launch instead of holding the wave.
template <typename R, int W, bool LUX>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(W)))
k_ev_photon(EvArgs<R> a) {
    extern __shared__ float rbuf_all[];
    int lane = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = lane < a.k.L.nlanes;
    using S = Shower<R, true, LUX>;
    S sh(a.k.T, a.jobs[0].G, a.jobs[0].vox, a.k.P, a.k.L, a.k.tally, live ? lane : 0,
         rbuf_all + threadIdx.x);
    if (a.gated && a.ctl->done) return;
    ev_setup(sh, a);
    const int n = min(a.ctl->n[a.ic], a.cons.cap);
    // deferral starts only after a lane of this wave found the bank empty,
    // and only in launches with work for at least half of the lanes (small
    // launches run every particle to its end, so everything progresses)
    const int defer_below = n >= (int)(gridDim.x * blockDim.x / 2)
                                ? min(a.defer_below, __popcll(__ballot(live))) : 0;
    int wait = 0;
    for (; live;) {
        if (__popcll(__ballot(1)) < defer_below) {
            // the bank is drained and this wave is nearly idle
            bool top = true;
            while (sh.np > 0) {
                if (sh.p.iq != 0) {
                    bank_push_other(sh, a);
                } else if (top && sh.pph != S::kPEnter) {
                    bank_push(sh, a.defr, &a.ctl->n[a.id], sh.pph, sh.p_dpmfp, sh.p_gle);
                } else {
                    bank_push(sh, a.defr, &a.ctl->n[a.id]);
                }
                sh.pop();
                top = false;
            }
            if (sh.slot >= 0) sh.slot_release();
            break;
        }
        sh.rtopup();
        sh.flush_pend();
        if (sh.np == 0) {
            if (sh.slot >= 0) sh.slot_release();
            int i = wave_claim(&a.ctl->head[a.ic]);
            if (i >= n) break;
            sh.slot = bank_load(a.cons, i, sh.p);
            ev_job(sh, a, sh.slot);
            sh.np = 1;
            sh.hsteps = 0;
            sh.pph = a.cons.ph[i];
            if (sh.pph != S::kPEnter) {   // resume a deferred flight
                sh.p_dpmfp = a.cons.s0[i];
                sh.p_gle = a.cons.s1[i];
                VoxData<R> d = sh.vox[sh.p.ir];
                sh.p_med = d.med;
                sh.p_rhof = d.rho;
                sh.p_cohfac = 1;
                sh.photon_medium();
            }
        }
        if (sh.p.iq != 0) {
            bank_push_other(sh, a);
            sh.pop();
            sh.pph = S::kPEnter;
            continue;
        }
        if (sh.pph == S::kPEnter) {
            sh.photon_enter();
            if (sh.pph == S::kPEnter) continue;
        }
        if (sh.pph == S::kPNewE) {
            sh.photon_newe();
            if (sh.pph != S::kPTrans) continue;
        }
        int act = sh.pph;
        int chosen = wave_choose(act, wait, S::kPTrans, S::kPRayl, S::kPPair, a.batch, a.max_wait);
        if (act != chosen) { wait++; continue; }
        wait = 0;
        if (chosen == S::kPTrans) sh.photon_trans(a.ksteps);
        else sh.photon_interact();
    }
    add_counters(a.k.tally, sh, 2);
}

// Charged-particle kernel: every lane takes electrons/positrons from bank
// `cons` and advances them one condensed-history step per turn, so that
// lanes whose particle ends pick up the next one while the others keep
// stepping; discrete interactions are batched by type over the wave.
// Photons created on the way go to bank `push`; a nearly idle wave after
// the bank is drained defers its particles to bank `defr`.
template <typename R, int W, bool LUX>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(W)))
k_ev_electron(EvArgs<R> a) {
    extern __shared__ float rbuf_all[];
    int lane = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = lane < a.k.L.nlanes;
    using S = Shower<R, true, LUX>;
    S sh(a.k.T, a.jobs[0].G, a.jobs[0].vox, a.k.P, a.k.L, a.k.tally, live ? lane : 0,
         rbuf_all + threadIdx.x);
    if (a.gated && a.ctl->done) return;
    ev_setup(sh, a);
    const int n = min(a.ctl->n[a.ic], a.cons.cap);
    // deferral starts only after a lane of this wave found the bank empty,
    // and only in launches with work for at least half of the lanes (small
    // launches run every particle to its end, so everything progresses)
    const int defer_below = n >= (int)(gridDim.x * blockDim.x / 2)
                                ? min(a.defer_below, __popcll(__ballot(live))) : 0;
    int wait = 0;
    for (; live;) {
        if (__popcll(__ballot(1)) < defer_below) {
            bool top = true;
            while (sh.np > 0) {
                if (sh.p.iq == 0) {
                    bank_push_other(sh, a);
                } else if (top && sh.eph != S::kNew) {
                    bank_push(sh, a.defr, &a.ctl->n[a.id], sh.eph | (sh.e_ctstep ? 8 : 0),
                              sh.e_demfp, sh.e_ttstep, sh.e_sig0, sh.e_elke);
                } else {
                    bank_push(sh, a.defr, &a.ctl->n[a.id]);
                }
                sh.pop();
                top = false;
            }
            if (sh.slot >= 0) sh.slot_release();
            break;
        }
        sh.rtopup();
        sh.flush_pend();
        if (sh.np == 0) {
            if (sh.slot >= 0) sh.slot_release();
            int i = wave_claim(&a.ctl->head[a.ic]);
            if (i >= n) break;
            sh.slot = bank_load(a.cons, i, sh.p);
            ev_job(sh, a, sh.slot);
            sh.np = 1;
            sh.hsteps = 0;
            int ph = a.cons.ph[i];
            sh.eph = ph & 7;
            if (sh.eph != S::kNew) {   // resume a deferred particle
                sh.e_ctstep = (ph & 8) != 0;
                sh.e_demfp = a.cons.s0[i];
                sh.e_ttstep = a.cons.s1[i];
                sh.e_sig0 = a.cons.s2[i];
                sh.e_elke = a.cons.s3[i];
                sh.e_irl = sh.p.ir;
                VoxData<R> d = sh.vox[sh.p.ir];
                sh.e_med = d.med;
                sh.e_rhof = d.rho;
            }
        }
        if (sh.p.iq == 0) {
            bank_push_other(sh, a);
            sh.pop();
            sh.eph = S::kNew;
            continue;
        }
        int act = sh.eph < S::kBrems ? 0 : sh.eph;
        int chosen = wave_choose(act, wait, 0, S::kBrems, S::kAnnih, a.batch, a.max_wait);
        if (act != chosen) { wait++; continue; }
        wait = 0;
        if (chosen == 0) sh.electron_step();
        else sh.electron_interact();
    }
    add_counters(a.k.tally, sh, 3);
}

// Every kernel instantiation the engine launches, as X((argument list),
// kernel with its template arguments), in groups that are compiled in
// separate files.
#define EGSGPU_EVENT_KERNELS(X, R)                  \
    X((EvArgs<R>), k_ev_source<R, false>)           \
    X((EvArgs<R>), k_ev_source<R, true>)            \
    X((EvArgs<R>), k_ev_photon<R, 2, false>)        \
    X((EvArgs<R>), k_ev_photon<R, 3, false>)        \
    X((EvArgs<R>), k_ev_photon<R, 4, false>)        \
    X((EvArgs<R>), k_ev_photon<R, 2, true>)
#define EGSGPU_ELECTRON_KERNELS(X, R)               \
    X((EvArgs<R>), k_ev_electron<R, 2, false>)      \
    X((EvArgs<R>), k_ev_electron<R, 2, true>)
#define EGSGPU_ELECTRON_WAVE_KERNELS(X, R)          \
    X((EvArgs<R>), k_ev_electron<R, 3, false>)      \
    X((EvArgs<R>), k_ev_electron<R, 4, false>)
#define EGSGPU_HISTORY_KERNELS(X, R)                \
    X((KArgs<R>), k_transport<R, false>)            \
    X((KArgs<R>), k_transport<R, true>)             \
    X((KArgs<R>), k_transport<R, false, true>)      \
    X((KArgs<R>), k_transport<R, true, true>)
#define EGSGPU_KERMA_KERNELS(X, R)                  \
    X((KArgs<R>, int, int, int), k_kerma<R, false>) \
    X((KArgs<R>, int, int, int), k_kerma<R, true>)
#define EGSGPU_SPLIT_KERNELS(X, R)                     \
    X((KArgs<R>), k_transport<R, false, false, true>)  \
    X((KArgs<R>), k_transport<R, true, false, true>)   \
    X((KArgs<R>), k_transport<R, false, true, true>)   \
    X((KArgs<R>), k_transport<R, true, true, true>)
#define EGSGPU_ALL_KERNELS(X, R)          \
    EGSGPU_EVENT_KERNELS(X, R)            \
    EGSGPU_ELECTRON_KERNELS(X, R)         \
    EGSGPU_ELECTRON_WAVE_KERNELS(X, R)    \
    EGSGPU_HISTORY_KERNELS(X, R)          \
    EGSGPU_KERMA_KERNELS(X, R)            \
    EGSGPU_SPLIT_KERNELS(X, R)
// explicit instantiation definition / declaration of one kernel
#define EGSGPU_INSTANTIATE(ARGS, ...) template __global__ void __VA_ARGS__ ARGS;
#define EGSGPU_DECLARE(ARGS, ...) extern template __global__ void __VA_ARGS__ ARGS;

}  // namespace egsgpu
