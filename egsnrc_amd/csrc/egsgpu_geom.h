// egsgpu_geom.h -- device voxel stepping for the XYZ (voxel) geometry.
//
// Restates EGS_XYZGeometry (egs++/geometry/egs_nd_geometry/egs_nd_geometry.h):
//   isWhere :791-804, howfar :963-1104, hownear :1106-1148,
//   howfarFromOut :1281-1381, with EGS_PlanesT::isWhere/hownear
//   (geometry/egs_planes/egs_planes.h:301-311,430-447) and
//   EGS_BaseGeometry::findRegion (egs_base_geometry.h:178-189).
// R = float is the transport build (boundary tolerance 2^-21 as under
// SINGLE, egs_functions.h:80-87); R = double the FP64 build.  Floating
// point contraction is disabled here so that every step length is the
// correctly rounded result of the same operations as the reference code.
#pragma once
#include <hip/hip_runtime.h>

namespace egsgpu {

template <typename R>
struct GeomView {
    int nx, ny, nz, nxy;
    const R *xp, *yp, *zp;       // plane positions (nx+1, ny+1, nz+1)
    R btol, hbtol;
    // equally spaced planes on every axis: 1 / spacing per axis (0 if the
    // planes of that axis are not equally spaced)
    R ixs, iys, izs;
};

template <typename R>
__host__ __device__ inline int planes_find(R x, int n, const R *__restrict__ p) {
    int ml = 0, mu = n;
    while (mu - ml > 1) {
        int mav = (ml + mu) >> 1;
        if (x <= p[mav]) mu = mav; else ml = mav;
    }
    return mu - 1;
}

template <typename R>
__host__ __device__ inline int planes_where(R x, int n, const R *__restrict__ p) {
    if (x < p[0] || x > p[n]) return -1;
    if (n == 1) return 0;
    return planes_find(x, n, p);
}

template <typename R>
__host__ __device__ inline int geom_is_where(const GeomView<R> &g, R x, R y, R z) {
    int ix = planes_where(x, g.nx, g.xp);
    if (ix < 0) return ix;
    int iy = planes_where(y, g.ny, g.yp);
    if (iy < 0) return iy;
    int iz = planes_where(z, g.nz, g.zp);
    if (iz < 0) return iz;
    return ix + iy * g.nx + iz * g.nxy;
}

// howfar for a particle inside voxel ireg; t is the intended step on input
// and the distance to the crossed plane (if closer) on output; returns the
// voxel entered (-1 = leaves the grid) or ireg when no plane is crossed.
template <typename R>
__host__ __device__ inline int geom_howfar_in(const GeomView<R> &g, int ireg,
                                             R x, R y, R z, R u, R v, R w,
                                             R &t) {
#pragma clang fp contract(off)
    int iz = ireg / g.nxy;
    int ir = ireg - iz * g.nxy;
    int iy = ir / g.nx;
    int ix = ir - iy * g.nx;
    int inew = ireg;
    if (u > 0) {
        R d = (g.xp[ix + 1] - x) / u;
        if (d <= t) { t = d <= g.btol ? g.hbtol : d; inew = (++ix) < g.nx ? ireg + 1 : -1; }
    } else if (u < 0) {
        R d = (g.xp[ix] - x) / u;
        if (d <= t) { t = d <= g.btol ? g.hbtol : d; inew = (--ix) >= 0 ? ireg - 1 : -1; }
    }
    if (v > 0) {
        R d = (g.yp[iy + 1] - y) / v;
        if (d <= t) { t = d <= g.btol ? g.hbtol : d; inew = (++iy) < g.ny ? ireg + g.nx : -1; }
    } else if (v < 0) {
        R d = (g.yp[iy] - y) / v;
        if (d <= t) { t = d <= g.btol ? g.hbtol : d; inew = (--iy) >= 0 ? ireg - g.nx : -1; }
    }
    if (w > 0) {
        R d = (g.zp[iz + 1] - z) / w;
        if (d <= t) { t = d <= g.btol ? g.hbtol : d; inew = (++iz) < g.nz ? ireg + g.nxy : -1; }
    } else if (w < 0) {
        R d = (g.zp[iz] - z) / w;
        if (d <= t) { t = d <= g.btol ? g.hbtol : d; inew = (--iz) >= 0 ? ireg - g.nxy : -1; }
    }
    return inew;
}

// geom_howfar_in for a caller that carries the voxel's plane indices
// (ix, iy, iz) beside the region index: the same operands, divisions and
// comparisons in the same order, so t and the returned region are those of
// geom_howfar_in, without the two integer divisions that recover the
// indices from ireg, and with one plane load per axis whatever the sign of
// the direction cosine (lanes of a wave moving either way share the
// instructions).  The indices are advanced only along the axis whose plane
// is crossed (the last axis to pass the test, as in geom_howfar_in).
template <typename R>
__host__ __device__ inline int geom_howfar_idx(const GeomView<R> &g, int ireg,
                                               int &ix, int &iy, int &iz,
                                               R x, R y, R z, R u, R v, R w,
                                               R &t) {
#pragma clang fp contract(off)
    int axis = -1, step = 0;
    if (u != 0) {
        const bool up = u > 0;
        R d = (g.xp[ix + (up ? 1 : 0)] - x) / u;
        if (d <= t) { t = d <= g.btol ? g.hbtol : d; axis = 0; step = up ? 1 : -1; }
    }
    if (v != 0) {
        const bool up = v > 0;
        R d = (g.yp[iy + (up ? 1 : 0)] - y) / v;
        if (d <= t) { t = d <= g.btol ? g.hbtol : d; axis = 1; step = up ? 1 : -1; }
    }
    if (w != 0) {
        const bool up = w > 0;
        R d = (g.zp[iz + (up ? 1 : 0)] - z) / w;
        if (d <= t) { t = d <= g.btol ? g.hbtol : d; axis = 2; step = up ? 1 : -1; }
    }
    if (axis < 0) return ireg;
    if (axis == 0) {
        int n = ix + step;
        if (n < 0 || n >= g.nx) return -1;
        ix = n;
        return ireg + step;
    }
    if (axis == 1) {
        int n = iy + step;
        if (n < 0 || n >= g.ny) return -1;
        iy = n;
        return ireg + step * g.nx;
    }
    int n = iz + step;
    if (n < 0 || n >= g.nz) return -1;
    iz = n;
    return ireg + step * g.nxy;
}

// ---- HOWFARLESS (DOSXYZnrc, dosxyznrc.mortran:3900-4360 with
// ihowfarless): in a homogeneous phantom howfar and hownear see only the
// outer faces of the grid.

// Distance along (u, v, w) from a point inside the grid to its outer
// boundary; a distance within the boundary tolerance becomes half of it,
// as in geom_howfar_in, so that a particle on an outer face moving out
// always leaves.
template <typename R>
__host__ __device__ inline R geom_box_exit(const GeomView<R> &g, R x, R y, R z, R u, R v, R w) {
    R t = (R)1e30;
    if (u > 0) { R d = (g.xp[g.nx] - x) / u; t = d < t ? d : t; }
    else if (u < 0) { R d = (g.xp[0] - x) / u; t = d < t ? d : t; }
    if (v > 0) { R d = (g.yp[g.ny] - y) / v; t = d < t ? d : t; }
    else if (v < 0) { R d = (g.yp[0] - y) / v; t = d < t ? d : t; }
    if (w > 0) { R d = (g.zp[g.nz] - z) / w; t = d < t ? d : t; }
    else if (w < 0) { R d = (g.zp[0] - z) / w; t = d < t ? d : t; }
    return t <= g.btol ? g.hbtol : t;
}

// Perpendicular distance from a point inside the grid to its outer faces.
template <typename R>
__host__ __device__ inline R geom_box_near(const GeomView<R> &g, R x, R y, R z) {
    R t = x - g.xp[0];
    R d = g.xp[g.nx] - x; t = d < t ? d : t;
    d = y - g.yp[0]; t = d < t ? d : t;
    d = g.yp[g.ny] - y; t = d < t ? d : t;
    d = z - g.zp[0]; t = d < t ? d : t;
    d = g.zp[g.nz] - z; t = d < t ? d : t;
    return t > 0 ? t : (R)0;
}

// Voxel of a point inside the grid (the binary searches of DOSXYZnrc's
// ibsearch4, clamped to the grid).
template <typename R>
__host__ __device__ inline int geom_locate(const GeomView<R> &g, R x, R y, R z) {
    int ix = planes_find(x, g.nx, g.xp), iy = planes_find(y, g.ny, g.yp),
        iz = planes_find(z, g.nz, g.zp);
    ix = ix < 0 ? 0 : (ix >= g.nx ? g.nx - 1 : ix);
    iy = iy < 0 ? 0 : (iy >= g.ny ? g.ny - 1 : iy);
    iz = iz < 0 ? 0 : (iz >= g.nz ? g.nz - 1 : iz);
    return ix + iy * g.nx + iz * g.nxy;
}

// geom_locate with a first guess from the spacing when the planes of an
// axis are equally spaced: the guess is moved by at most one plane so that
// the result is the binary search's (x <= p[i+1] belongs to voxel i).
template <typename R>
__host__ __device__ inline int axis_find(R x, int n, const R *__restrict__ p, R inv) {
    if (inv > 0) {
        int i = (int)((x - p[0]) * inv);
        i = i < 0 ? 0 : (i >= n ? n - 1 : i);
        if (i > 0 && x <= p[i]) i--;
        else if (i < n - 1 && x > p[i + 1]) i++;
        return i;
    }
    int i = planes_find(x, n, p);
    return i < 0 ? 0 : (i >= n ? n - 1 : i);
}
template <typename R>
__host__ __device__ inline int geom_locate_fast(const GeomView<R> &g, R x, R y, R z) {
    return axis_find(x, g.nx, g.xp, g.ixs) + axis_find(y, g.ny, g.yp, g.iys) * g.nx +
           axis_find(z, g.nz, g.zp, g.izs) * g.nxy;
}

// howfarFromOut: entry into the grid from outside (t overwritten only on a
// hit; the tolerance re-check overwrites t1 through howfar before t = t1,
// as the reference does).
template <typename R>
__host__ __device__ inline int geom_howfar_out(const GeomView<R> &g, R x, R y,
                                              R z, R u, R v, R w, R &t) {
#pragma clang fp contract(off)
    R t1 = 0;
    int ix = -1, iy, iz;
    if (x <= g.xp[0] && u > 0) { t1 = (g.xp[0] - x) / u; ix = 0; }
    else if (x >= g.xp[g.nx] && u < 0) { t1 = (g.xp[g.nx] - x) / u; ix = g.nx - 1; }
    if (ix >= 0 && t1 <= t) {
        R tx = x + u * t1, ty = y + v * t1, tz = z + w * t1;
        iy = planes_where(ty, g.ny, g.yp);
        if (iy >= 0) {
            iz = planes_where(tz, g.nz, g.zp);
            if (iz >= 0) {
                int inew = ix + iy * g.nx + iz * g.nxy;
                if (t1 < g.btol) {
                    int itest = geom_howfar_in(g, inew, tx, ty, tz, u, v, w, t1);
                    if (itest == -1 && t1 < g.btol) return -1;
                }
                t = t1;
                return inew;
            }
        }
    }
    iy = -1;
    if (y <= g.yp[0] && v > 0) { t1 = (g.yp[0] - y) / v; iy = 0; }
    else if (y >= g.yp[g.ny] && v < 0) { t1 = (g.yp[g.ny] - y) / v; iy = g.ny - 1; }
    if (iy >= 0 && t1 <= t) {
        R tx = x + u * t1, ty = y + v * t1, tz = z + w * t1;
        ix = planes_where(tx, g.nx, g.xp);
        if (ix >= 0) {
            iz = planes_where(tz, g.nz, g.zp);
            if (iz >= 0) {
                int inew = ix + iy * g.nx + iz * g.nxy;
                if (t1 < g.btol) {
                    int itest = geom_howfar_in(g, inew, tx, ty, tz, u, v, w, t1);
                    if (itest == -1 && t1 < g.btol) return -1;
                }
                t = t1;
                return inew;
            }
        }
    }
    iz = -1;
    if (z <= g.zp[0] && w > 0) { t1 = (g.zp[0] - z) / w; iz = 0; }
    else if (z >= g.zp[g.nz] && w < 0) { t1 = (g.zp[g.nz] - z) / w; iz = g.nz - 1; }
    if (iz >= 0 && t1 <= t) {
        R tx = x + u * t1, ty = y + v * t1, tz = z + w * t1;
        ix = planes_where(tx, g.nx, g.xp);
        if (ix >= 0) {
            iy = planes_where(ty, g.ny, g.yp);
            if (iy >= 0) {
                int inew = ix + iy * g.nx + iz * g.nxy;
                if (t1 < g.btol) {
                    int itest = geom_howfar_in(g, inew, tx, ty, tz, u, v, w, t1);
                    if (itest == -1 && t1 < g.btol) return -1;
                }
                t = t1;
                return inew;
            }
        }
    }
    return -1;
}

template <typename R>
__host__ __device__ inline int geom_howfar(const GeomView<R> &g, int ireg, R x,
                                          R y, R z, R u, R v, R w, R &t) {
    return ireg >= 0 ? geom_howfar_in(g, ireg, x, y, z, u, v, w, t)
                     : geom_howfar_out(g, x, y, z, u, v, w, t);
}

template <typename R>
__host__ __device__ inline R planes_near(int i, R x, int n, const R *__restrict__ p) {
#pragma clang fp contract(off)
    R t = x - p[i];
    if (i + 1 <= n) {
        R t2 = p[i + 1] - x;
        if (t2 < t) t = t2;
    }
    return t;
}

template <typename R>
__host__ __device__ inline R planes_near_out(R x, int n, const R *__restrict__ p) {
    if (x <= p[0]) return p[0] - x;
    if (x >= p[n]) return x - p[n];
    return 0;
}

template <typename R>
__host__ __device__ inline R geom_hownear(const GeomView<R> &g, int ireg, R x,
                                         R y, R z) {
#pragma clang fp contract(off)
    if (ireg >= 0) {
        int iz = ireg / g.nxy;
        int ir = ireg - iz * g.nxy;
        int iy = ir / g.nx;
        int ix = ir - iy * g.nx;
        R t = planes_near(ix, x, g.nx, g.xp);
        R t1 = planes_near(iy, y, g.ny, g.yp);
        if (t1 < t) t = t1;
        t1 = planes_near(iz, z, g.nz, g.zp);
        if (t1 < t) t = t1;
        return t;
    }
    int nc = 0;
    R s1 = 0, s2 = 0;
    if (x < g.xp[0] || x > g.xp[g.nx]) { R t = planes_near_out(x, g.nx, g.xp); nc++; s1 += t; s2 += t * t; }
    if (y < g.yp[0] || y > g.yp[g.ny]) { R t = planes_near_out(y, g.ny, g.yp); nc++; s1 += t; s2 += t * t; }
    if (z < g.zp[0] || z > g.zp[g.nz]) { R t = planes_near_out(z, g.nz, g.zp); nc++; s1 += t; s2 += t * t; }
    if (nc == 1) return s1;
    return sqrt(s2);
}

}  // namespace egsgpu
