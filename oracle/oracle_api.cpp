// oracle_api.cpp -- C entry points of the CPU oracle (liboracle.so).
//
// TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg, as the checker of the HIP engine.  The
// signatures mirror include/egsgpu.h so that both are driven by the same
// harness.
#include <cstdint>
#include <cstring>
#include <vector>

#include "oracle_geom.h"
#include "oracle_rng.h"

using namespace oracle;

extern "C" {

int oracle_ranmar_fill(int32_t ixx, int32_t jxx, int32_t nstreams, int64_t n,
                       double *out) {
    if (nstreams <= 0 || n < 0 || !out) return -1;
    for (int s = 0; s < nstreams; s++) {
        int is, js;
        stream_seeds(ixx, jxx, s, is, js);
        Ranmar r;
        r.set_state(is, js);
        for (int64_t i = 0; i < n; i++) out[(int64_t)s * n + i] = r.uniform();
    }
    return 0;
}

int oracle_ranlux_fill(int32_t lux, int32_t seed, int32_t nstreams, int64_t n,
                       double *out) {
    if (nstreams <= 0 || n < 0 || !out) return -1;
    for (int s = 0; s < nstreams; s++) {
        RanluxStream r;
        r.init(lux, seed + s);
        for (int64_t i = 0; i < n; i++) out[(int64_t)s * n + i] = r.uniform();
    }
    return 0;
}

}  // extern "C"

// Same layout as EgsXYZGeometry in include/egsgpu.h.
struct OracleGeomDesc {
    int32_t nx, ny, nz;
    const double *xb, *yb, *zb;
    const int32_t *medium;
    const double *rhor;
    double boundary_tolerance;
};

template <typename R>
static int trace_impl(const OracleGeomDesc *g, int32_t nrays, const double *x,
                      const double *u, int32_t max_steps, int32_t *regions,
                      double *steps, int32_t *nsteps) {
    XYZ<R> geo;
    geo.setup(g->nx, g->ny, g->nz, g->xb, g->yb, g->zb, g->boundary_tolerance);
    for (int r = 0; r < nrays; r++) {
        R px = (R)x[3 * r], py = (R)x[3 * r + 1], pz = (R)x[3 * r + 2];
        R ux = (R)u[3 * r], uy = (R)u[3 * r + 1], uz = (R)u[3 * r + 2];
        int k = 0;
        int ireg = geo.is_where(px, py, pz);
        if (ireg < 0) {
            R t = (R)1e30;
            ireg = geo.howfar(-1, px, py, pz, ux, uy, uz, t);
            if (ireg >= 0) {
                px = px + ux * t; py = py + uy * t; pz = pz + uz * t;
                regions[(int64_t)r * max_steps + k] = ireg;
                steps[(int64_t)r * max_steps + k] = (double)t;
                k++;
            }
        } else {
            regions[(int64_t)r * max_steps + k] = ireg;
            steps[(int64_t)r * max_steps + k] = 0;
            k++;
        }
        while (ireg >= 0 && k < max_steps) {
            R t = (R)1e30;
            int inew = geo.howfar(ireg, px, py, pz, ux, uy, uz, t);
            px = px + ux * t; py = py + uy * t; pz = pz + uz * t;
            regions[(int64_t)r * max_steps + k] = inew;
            steps[(int64_t)r * max_steps + k] = (double)t;
            k++;
            ireg = inew;
        }
        nsteps[r] = k;
    }
    return 0;
}

extern "C" {

int oracle_trace_rays(const OracleGeomDesc *g, int32_t nrays, const double *x,
                      const double *u, int32_t max_steps, int32_t precision,
                      int32_t *regions, double *steps, int32_t *nsteps) {
    if (!g || nrays < 0 || max_steps < 1) return -1;
    if (precision == 32)
        return trace_impl<float>(g, nrays, x, u, max_steps, regions, steps, nsteps);
    return trace_impl<double>(g, nrays, x, u, max_steps, regions, steps, nsteps);
}

int oracle_hownear(const OracleGeomDesc *g, int32_t n, const double *x,
                   const int32_t *ireg, int32_t precision, double *tperp) {
    if (!g || n < 0) return -1;
    if (precision == 32) {
        XYZ<float> geo;
        geo.setup(g->nx, g->ny, g->nz, g->xb, g->yb, g->zb, g->boundary_tolerance);
        for (int i = 0; i < n; i++)
            tperp[i] = geo.hownear(ireg[i], (float)x[3 * i], (float)x[3 * i + 1], (float)x[3 * i + 2]);
    } else {
        XYZ<double> geo;
        geo.setup(g->nx, g->ny, g->nz, g->xb, g->yb, g->zb, g->boundary_tolerance);
        for (int i = 0; i < n; i++)
            tperp[i] = geo.hownear(ireg[i], x[3 * i], x[3 * i + 1], x[3 * i + 2]);
    }
    return 0;
}

int oracle_is_where(const OracleGeomDesc *g, int32_t n, const double *x,
                    int32_t precision, int32_t *ireg) {
    if (!g || n < 0) return -1;
    if (precision == 32) {
        XYZ<float> geo;
        geo.setup(g->nx, g->ny, g->nz, g->xb, g->yb, g->zb, g->boundary_tolerance);
        for (int i = 0; i < n; i++)
            ireg[i] = geo.is_where((float)x[3 * i], (float)x[3 * i + 1], (float)x[3 * i + 2]);
    } else {
        XYZ<double> geo;
        geo.setup(g->nx, g->ny, g->nz, g->xb, g->yb, g->zb, g->boundary_tolerance);
        for (int i = 0; i < n; i++)
            ireg[i] = geo.is_where(x[3 * i], x[3 * i + 1], x[3 * i + 2]);
    }
    return 0;
}

}  // extern "C"
