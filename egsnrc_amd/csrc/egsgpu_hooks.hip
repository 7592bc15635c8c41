// egsgpu_hooks.hip -- error handling, device query and the test-hook
// entry points of include/egsgpu.h (RNG streams, ray tracing, hownear).
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>
#include <vector>

#include "../../include/egsgpu.h"
#include "egsgpu_common.h"
#include "egsgpu_geom.h"
#include "egsgpu_rng.h"

namespace egsgpu {

static thread_local std::string g_err;

void set_error(const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
}
const std::string &get_error() { return g_err; }

int require_device() {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n < 1) {
        set_error("no HIP device visible (%s)",
                  e == hipSuccess ? "count 0" : hipGetErrorString(e));
        return EGSGPU_E_NODEVICE;
    }
    return EGSGPU_OK;
}

// ---- RNG fill ----------------------------------------------------------
// Streams drawn through the per-lane ring exactly as the transport draws
// them (ring_topup / ring_next).  The ring is topped up on 16 draws out of
// every 64 and left alone for the other 48, so that some blocks come from
// the top-up of a ring running low and others from a ring that ran dry.
constexpr int kFillBlock = 64;

__device__ __forceinline__ bool fill_topup_turn(int64_t i) { return (i & 63) < 16; }

__global__ void __launch_bounds__(kFillBlock)
k_ranmar_fill(int ixx, int jxx, int nstreams, int64_t n, RanmarRec *recs,
              double *out) {
    __shared__ float ring[kRing * kFillBlock];
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nstreams) return;
    int is, js;
    stream_seeds(ixx, jxx, s, is, js);
    RanmarRec *r = recs + s;
    ranmar_seed(r, is, js);
    float *b = ring + threadIdx.x;
    int pos = 0, avail = 0;
    for (int64_t i = 0; i < n; i++) {
        if (fill_topup_turn(i)) ring_topup<false>(r, b, kFillBlock, pos, avail);
        out[(int64_t)s * n + i] = (double)ring_next<false>(r, b, kFillBlock, pos, avail);
    }
}

__global__ void __launch_bounds__(kFillBlock)
k_ranlux_fill(int lux, int seed, int nstreams, int64_t n, RanluxRec *recs,
              double *out) {
    __shared__ float ring[kRing * kFillBlock];
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nstreams) return;
    RanluxRec *r = recs + s;
    ranlux_seed(r, lux, seed + s);
    float *b = ring + threadIdx.x;
    int pos = 0, avail = 0;
    for (int64_t i = 0; i < n; i++) {
        if (fill_topup_turn(i)) ring_topup<true>(r, b, kFillBlock, pos, avail);
        out[(int64_t)s * n + i] = (double)ring_next<true>(r, b, kFillBlock, pos, avail);
    }
}

// ---- geometry ----------------------------------------------------------
template <typename R>
__global__ void k_trace(GeomView<R> g, int nrays, const double *x,
                        const double *u, int max_steps, int *regions,
                        double *steps, int *nsteps) {
#pragma clang fp contract(off)
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrays) return;
    R px = (R)x[3 * r], py = (R)x[3 * r + 1], pz = (R)x[3 * r + 2];
    R ux = (R)u[3 * r], uy = (R)u[3 * r + 1], uz = (R)u[3 * r + 2];
    int64_t base = (int64_t)r * max_steps;
    int k = 0;
    int ireg = geom_is_where(g, px, py, pz);
    if (ireg < 0) {
        R t = (R)1e30;
        ireg = geom_howfar(g, -1, px, py, pz, ux, uy, uz, t);
        if (ireg >= 0) {
            px = px + ux * t; py = py + uy * t; pz = pz + uz * t;
            regions[base + k] = ireg;
            steps[base + k] = (double)t;
            k++;
        }
    } else {
        regions[base + k] = ireg;
        steps[base + k] = 0;
        k++;
    }
    // inside the grid the ray steps like a photon flight: with the plane
    // indices carried (geom_howfar_idx); every step is also taken by
    // geom_howfar_in, and a difference ends the ray with nsteps = -1
    int iz = 0, iy = 0, ix = 0;
    if (ireg >= 0) {
        iz = ireg / g.nxy;
        iy = (ireg - iz * g.nxy) / g.nx;
        ix = ireg - iz * g.nxy - iy * g.nx;
    }
    while (ireg >= 0 && k < max_steps) {
        R t = (R)1e30, t_in = (R)1e30;
        int inew = geom_howfar_idx(g, ireg, ix, iy, iz, px, py, pz, ux, uy, uz, t);
        int inew_in = geom_howfar_in(g, ireg, px, py, pz, ux, uy, uz, t_in);
        if (inew != inew_in || t != t_in) { nsteps[r] = -1; return; }
        px = px + ux * t; py = py + uy * t; pz = pz + uz * t;
        regions[base + k] = inew;
        steps[base + k] = (double)t;
        k++;
        ireg = inew;
    }
    nsteps[r] = k;
}

template <typename R>
__global__ void k_hownear(GeomView<R> g, int n, const double *x,
                          const int *ireg, double *out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = (double)geom_hownear(g, ireg[i], (R)x[3 * i], (R)x[3 * i + 1],
                                  (R)x[3 * i + 2]);
}

// Device copy of the plane positions in precision R.
template <typename R>
struct DevPlanes {
    R *d = nullptr;
    ~DevPlanes() { if (d) (void)hipFree(d); }
    int upload(const EgsXYZGeometry *g, GeomView<R> &v) {
        int n = (g->nx + 1) + (g->ny + 1) + (g->nz + 1);
        std::vector<R> h(n);
        int k = 0;
        for (int i = 0; i <= g->nx; i++) h[k++] = (R)g->xb[i];
        for (int i = 0; i <= g->ny; i++) h[k++] = (R)g->yb[i];
        for (int i = 0; i <= g->nz; i++) h[k++] = (R)g->zb[i];
        EGS_HIP(hipMalloc(&d, n * sizeof(R)));
        EGS_HIP(hipMemcpy(d, h.data(), n * sizeof(R), hipMemcpyHostToDevice));
        v.nx = g->nx; v.ny = g->ny; v.nz = g->nz; v.nxy = g->nx * g->ny;
        v.xp = d; v.yp = d + g->nx + 1; v.zp = d + g->nx + 1 + g->ny + 1;
        double tol = g->boundary_tolerance;
        if (tol <= 0) tol = sizeof(R) == 4 ? 1.0 / (1 << 21) : 1.0 / 4294967296.0;
        v.btol = (R)tol;
        v.hbtol = v.btol / (R)2;
        return EGSGPU_OK;
    }
};

int check_geom(const EgsXYZGeometry *g) {
    EGS_ARG(g, "geometry is NULL");
    EGS_ARG(g->nx > 0 && g->ny > 0 && g->nz > 0, "grid dimensions must be > 0");
    EGS_ARG(g->xb && g->yb && g->zb, "plane arrays are NULL");
    const double *p[3] = {g->xb, g->yb, g->zb};
    int n[3] = {g->nx, g->ny, g->nz};
    for (int a = 0; a < 3; a++)
        for (int i = 0; i < n[a]; i++)
            EGS_ARG(p[a][i] < p[a][i + 1], "plane positions must increase");
    return EGSGPU_OK;
}

template <typename R>
static int trace_impl(const EgsXYZGeometry *g, int32_t nrays, const double *x,
                      const double *u, int32_t max_steps, int32_t *regions,
                      double *steps, int32_t *nsteps) {
    DevPlanes<R> planes;
    GeomView<R> v;
    int rc = planes.upload(g, v);
    if (rc) return rc;
    double *dx = nullptr, *du = nullptr, *ds = nullptr;
    int *dr = nullptr, *dn = nullptr;
    int64_t tot = (int64_t)nrays * max_steps;
    int err = EGSGPU_OK;
    auto cleanup = [&]() {
        if (dx) (void)hipFree(dx);
        if (du) (void)hipFree(du);
        if (ds) (void)hipFree(ds);
        if (dr) (void)hipFree(dr);
        if (dn) (void)hipFree(dn);
    };
#define T_HIP(call)                                                          \
    do {                                                                     \
        hipError_t e_ = (call);                                              \
        if (e_ != hipSuccess) {                                              \
            set_error("%s failed: %s", #call, hipGetErrorString(e_));        \
            cleanup();                                                       \
            return EGSGPU_E_HIP;                                             \
        }                                                                    \
    } while (0)
    T_HIP(hipMalloc(&dx, 3 * sizeof(double) * nrays));
    T_HIP(hipMalloc(&du, 3 * sizeof(double) * nrays));
    T_HIP(hipMalloc(&ds, sizeof(double) * tot));
    T_HIP(hipMalloc(&dr, sizeof(int) * tot));
    T_HIP(hipMalloc(&dn, sizeof(int) * nrays));
    T_HIP(hipMemset(ds, 0, sizeof(double) * tot));
    T_HIP(hipMemset(dr, 0, sizeof(int) * tot));
    T_HIP(hipMemcpy(dx, x, 3 * sizeof(double) * nrays, hipMemcpyHostToDevice));
    T_HIP(hipMemcpy(du, u, 3 * sizeof(double) * nrays, hipMemcpyHostToDevice));
    int bs = 128;
    hipLaunchKernelGGL(k_trace<R>, dim3((nrays + bs - 1) / bs), dim3(bs), 0, 0,
                       v, nrays, dx, du, max_steps, dr, ds, dn);
    T_HIP(hipGetLastError());
    T_HIP(hipDeviceSynchronize());
    T_HIP(hipMemcpy(regions, dr, sizeof(int) * tot, hipMemcpyDeviceToHost));
    T_HIP(hipMemcpy(steps, ds, sizeof(double) * tot, hipMemcpyDeviceToHost));
    T_HIP(hipMemcpy(nsteps, dn, sizeof(int) * nrays, hipMemcpyDeviceToHost));
#undef T_HIP
    cleanup();
    return err;
}

template <typename R>
static int hownear_impl(const EgsXYZGeometry *g, int32_t n, const double *x,
                        const int32_t *ireg, double *tperp) {
    DevPlanes<R> planes;
    GeomView<R> v;
    int rc = planes.upload(g, v);
    if (rc) return rc;
    double *dx = nullptr, *dout = nullptr;
    int *di = nullptr;
    EGS_HIP(hipMalloc(&dx, 3 * sizeof(double) * n));
    EGS_HIP(hipMalloc(&dout, sizeof(double) * n));
    EGS_HIP(hipMalloc(&di, sizeof(int) * n));
    EGS_HIP(hipMemcpy(dx, x, 3 * sizeof(double) * n, hipMemcpyHostToDevice));
    EGS_HIP(hipMemcpy(di, ireg, sizeof(int) * n, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_hownear<R>, dim3((n + 127) / 128), dim3(128), 0, 0, v,
                       n, dx, di, dout);
    EGS_HIP(hipGetLastError());
    EGS_HIP(hipMemcpy(tperp, dout, sizeof(double) * n, hipMemcpyDeviceToHost));
    EGS_HIP(hipFree(dx));
    EGS_HIP(hipFree(dout));
    EGS_HIP(hipFree(di));
    return EGSGPU_OK;
}

}  // namespace egsgpu

using namespace egsgpu;

extern "C" {

int egsgpu_last_error(char *buf, size_t n) {
    if (!buf || n == 0) return EGSGPU_E_ARG;
    strncpy(buf, g_err.c_str(), n - 1);
    buf[n - 1] = 0;
    return EGSGPU_OK;
}

int egsgpu_device_count(int32_t *n) {
    EGS_ARG(n, "n is NULL");
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    *n = e == hipSuccess ? c : 0;
    return EGSGPU_OK;
}

int egsgpu_ranmar_fill(int32_t ixx, int32_t jxx, int32_t nstreams, int64_t n,
                       double *out) {
    EGS_ARG(nstreams > 0 && n >= 0 && out, "nstreams>0, n>=0, out");
    int rc = require_device();
    if (rc) return rc;
    RanmarRec *recs = nullptr;
    double *dout = nullptr;
    size_t bytes = sizeof(double) * (size_t)nstreams * (size_t)n;
    EGS_HIP(hipMalloc(&recs, sizeof(RanmarRec) * nstreams));
    EGS_HIP(hipMalloc(&dout, bytes ? bytes : 8));
    hipLaunchKernelGGL(k_ranmar_fill, dim3((nstreams + kFillBlock - 1) / kFillBlock),
                       dim3(kFillBlock), 0, 0, ixx, jxx, nstreams, n, recs, dout);
    EGS_HIP(hipGetLastError());
    EGS_HIP(hipMemcpy(out, dout, bytes, hipMemcpyDeviceToHost));
    EGS_HIP(hipFree(recs));
    EGS_HIP(hipFree(dout));
    return EGSGPU_OK;
}

int egsgpu_ranlux_fill(int32_t lux, int32_t seed, int32_t nstreams, int64_t n,
                       double *out) {
    EGS_ARG(nstreams > 0 && n >= 0 && out, "nstreams>0, n>=0, out");
    EGS_ARG(lux >= 0 && lux <= 4, "luxury level 0..4");
    int rc = require_device();
    if (rc) return rc;
    RanluxRec *recs = nullptr;
    double *dout = nullptr;
    size_t bytes = sizeof(double) * (size_t)nstreams * (size_t)n;
    EGS_HIP(hipMalloc(&recs, sizeof(RanluxRec) * nstreams));
    EGS_HIP(hipMalloc(&dout, bytes ? bytes : 8));
    hipLaunchKernelGGL(k_ranlux_fill, dim3((nstreams + kFillBlock - 1) / kFillBlock),
                       dim3(kFillBlock), 0, 0, lux, seed, nstreams, n, recs, dout);
    EGS_HIP(hipGetLastError());
    EGS_HIP(hipMemcpy(out, dout, bytes, hipMemcpyDeviceToHost));
    EGS_HIP(hipFree(recs));
    EGS_HIP(hipFree(dout));
    return EGSGPU_OK;
}

// ---- FP64 atomic rate ---------------------------------------------------
// Throughput of the FP64 atomic adds the tallies and the density-gradient
// sweep are made of.  Every thread issues `per` adds of 1.0 into a table of
// naddr doubles; pattern 0 spreads them over the table by a multiplicative
// hash (the tallies' access pattern: unrelated voxels per lane), pattern 1
// sends all lanes of a wave to one address per add (the worst case of a
// shared counter), pattern 2 gives the lanes of a wave 64 consecutive
// addresses (one coalesced atomic per wave).
__global__ void __launch_bounds__(256) k_atomic_rate(double *t, uint32_t naddr, int per,
                                                     int pattern) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t wave = g >> 6, lane = g & 63;
    for (int i = 0; i < per; i++) {
        uint32_t a;
        if (pattern == 0) a = (g * 2654435761u + (uint32_t)i * 40503u) % naddr;
        else if (pattern == 1) a = (wave * 97u + (uint32_t)i) % naddr;
        else a = ((wave * 64u + (uint32_t)i * 4096u) + lane) % naddr;
        atomicAdd(t + a, 1.0);
    }
}

int egsgpu_atomic_rate(int32_t pattern, int32_t naddr, int64_t nops, double *ms,
                       double *check_sum) {
    EGS_ARG(pattern >= 0 && pattern <= 2, "pattern 0..2");
    EGS_ARG(naddr >= 64 && nops > 0 && ms && check_sum, "naddr >= 64, nops > 0, outputs");
    int rc = require_device();
    if (rc) return rc;
    const int per = 64;
    const int64_t threads = (nops + per - 1) / per;
    EGS_ARG(threads <= ((int64_t)1 << 31) - 256, "nops too large");
    const int blocks = (int)((threads + 255) / 256);
    double *t = nullptr;
    EGS_HIP(hipMalloc(&t, sizeof(double) * (size_t)naddr));
    EGS_HIP(hipMemset(t, 0, sizeof(double) * (size_t)naddr));
    hipEvent_t e0, e1;
    EGS_HIP(hipEventCreate(&e0));
    EGS_HIP(hipEventCreate(&e1));
    // one untimed launch brings the table into the caches
    hipLaunchKernelGGL(k_atomic_rate, dim3(blocks), dim3(256), 0, 0, t, (uint32_t)naddr, per,
                       pattern);
    EGS_HIP(hipGetLastError());
    EGS_HIP(hipEventRecord(e0, 0));
    hipLaunchKernelGGL(k_atomic_rate, dim3(blocks), dim3(256), 0, 0, t, (uint32_t)naddr, per,
                       pattern);
    EGS_HIP(hipGetLastError());
    EGS_HIP(hipEventRecord(e1, 0));
    EGS_HIP(hipEventSynchronize(e1));
    float f = 0;
    EGS_HIP(hipEventElapsedTime(&f, e0, e1));
    *ms = f;
    std::vector<double> h((size_t)naddr);
    EGS_HIP(hipMemcpy(h.data(), t, sizeof(double) * (size_t)naddr, hipMemcpyDeviceToHost));
    double sum = 0;
    for (double v : h) sum += v;
    *check_sum = sum;   // = 2 * blocks * 256 * per: every add landed
    EGS_HIP(hipEventDestroy(e0));
    EGS_HIP(hipEventDestroy(e1));
    EGS_HIP(hipFree(t));
    return EGSGPU_OK;
}

int egsgpu_trace_rays(const EgsXYZGeometry *g, int32_t nrays, const double *x,
                      const double *u, int32_t max_steps, int32_t precision,
                      int32_t *regions, double *steps, int32_t *nsteps) {
    int rc = check_geom(g);
    if (rc) return rc;
    EGS_ARG(nrays > 0 && max_steps > 0 && x && u && regions && steps && nsteps,
            "rays, outputs");
    EGS_ARG(precision == 32 || precision == 64, "precision must be 32 or 64");
    rc = require_device();
    if (rc) return rc;
    if (precision == 32)
        return trace_impl<float>(g, nrays, x, u, max_steps, regions, steps, nsteps);
    return trace_impl<double>(g, nrays, x, u, max_steps, regions, steps, nsteps);
}

int egsgpu_hownear(const EgsXYZGeometry *g, int32_t n, const double *x,
                   const int32_t *ireg, int32_t precision, double *tperp) {
    int rc = check_geom(g);
    if (rc) return rc;
    EGS_ARG(n > 0 && x && ireg && tperp, "points, regions, output");
    EGS_ARG(precision == 32 || precision == 64, "precision must be 32 or 64");
    rc = require_device();
    if (rc) return rc;
    if (precision == 32) return hownear_impl<float>(g, n, x, ireg, tperp);
    return hownear_impl<double>(g, n, x, ireg, tperp);
}

}  // extern "C"
