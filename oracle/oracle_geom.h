// oracle_geom.h -- CPU restatement of EGS_XYZGeometry voxel stepping.
//
// TEST INFRASTRUCTURE ONLY (checker for the HIP engine; see oracle_rng.h).
//
// Follows egs++/geometry/egs_nd_geometry/egs_nd_geometry.h:
//   isWhere      :791-804   (EGS_PlanesT::isWhere egs_planes.h:301-311 and
//                            EGS_BaseGeometry::findRegion egs_base_geometry.h:178-189)
//   howfar       :963-1104
//   hownear      :1106-1148 (EGS_PlanesT::hownear egs_planes.h:430-447)
//   howfarFromOut:1281-1381
// R = double is the reference build; R = float is EGS_Float under SINGLE
// (egs_config1.h), where boundaryTolerance defaults to 2^-21 instead of
// 2^-32 (egs_functions.h:80-87).
#pragma once
#include <cmath>
#include <vector>

namespace oracle {

template <typename R>
struct XYZ {
    int nx, ny, nz, nxy;
    std::vector<R> xp, yp, zp;   // nx+1, ny+1, nz+1 plane positions
    R btol, hbtol;

    void setup(int NX, int NY, int NZ, const double *xb, const double *yb,
               const double *zb, double tol) {
        nx = NX; ny = NY; nz = NZ; nxy = nx * ny;
        xp.assign(xb, xb + nx + 1);
        yp.assign(yb, yb + ny + 1);
        zp.assign(zb, zb + nz + 1);
        if (tol <= 0) tol = sizeof(R) == 4 ? 1.0 / (1 << 21) : 1.0 / 4294967296.0;
        btol = (R)tol;
        hbtol = btol / (R)2;
    }
    // EGS_BaseGeometry::findRegion
    static int find_region(R x, int np, const R *p) {
        int ml = 0, mu = np;
        while (mu - ml > 1) {
            int mav = (ml + mu) / 2;
            if (x <= p[mav]) mu = mav; else ml = mav;
        }
        return mu - 1;
    }
    // EGS_PlanesT::isWhere with nreg = n planes-1
    static int planes_where(R x, int n, const std::vector<R> &p) {
        if (x < p[0] || x > p[n]) return -1;
        if (n == 1) return 0;
        return find_region(x, n, p.data());
    }
    int is_where(R x, R y, R z) const {
        int ix = planes_where(x, nx, xp);
        if (ix < 0) return ix;
        int iy = planes_where(y, ny, yp);
        if (iy < 0) return iy;
        int iz = planes_where(z, nz, zp);
        if (iz < 0) return iz;
        return ix + iy * nx + iz * nxy;
    }
    // egs_nd_geometry.h:963-1094 (inside part)
    int howfar_in(int ireg, R x, R y, R z, R u, R v, R w, R &t) const {
        int iz = ireg / nxy;
        int ir = ireg - iz * nxy;
        int iy = ir / nx;
        int ix = ir - iy * nx;
        int inew = ireg;
        if (u > 0) {
            R d = (xp[ix + 1] - x) / u;
            if (d <= t) {
                t = d <= btol ? hbtol : d;
                inew = (++ix) < nx ? ireg + 1 : -1;
            }
        } else if (u < 0) {
            R d = (xp[ix] - x) / u;
            if (d <= t) {
                t = d <= btol ? hbtol : d;
                inew = (--ix) >= 0 ? ireg - 1 : -1;
            }
        }
        if (v > 0) {
            R d = (yp[iy + 1] - y) / v;
            if (d <= t) {
                t = d <= btol ? hbtol : d;
                inew = (++iy) < ny ? ireg + nx : -1;
            }
        } else if (v < 0) {
            R d = (yp[iy] - y) / v;
            if (d <= t) {
                t = d <= btol ? hbtol : d;
                inew = (--iy) >= 0 ? ireg - nx : -1;
            }
        }
        if (w > 0) {
            R d = (zp[iz + 1] - z) / w;
            if (d <= t) {
                t = d <= btol ? hbtol : d;
                inew = (++iz) < nz ? ireg + nxy : -1;
            }
        } else if (w < 0) {
            R d = (zp[iz] - z) / w;
            if (d <= t) {
                t = d <= btol ? hbtol : d;
                inew = (--iz) >= 0 ? ireg - nxy : -1;
            }
        }
        return inew;
    }
    // egs_nd_geometry.h:1281-1378.  Note that the tolerance re-check calls
    // howfar(inew, tmp, u, t1), which overwrites t1 before t = t1.
    int howfar_out(R x, R y, R z, R u, R v, R w, R &t) const {
        R t1 = 0;
        int ix = -1, iy, iz;
        if (x <= xp[0] && u > 0) { t1 = (xp[0] - x) / u; ix = 0; }
        else if (x >= xp[nx] && u < 0) { t1 = (xp[nx] - x) / u; ix = nx - 1; }
        if (ix >= 0 && t1 <= t) {
            R ty = y + v * t1, tz = z + w * t1, tx = x + u * t1;
            iy = planes_where(ty, ny, yp);
            if (iy >= 0) {
                iz = planes_where(tz, nz, zp);
                if (iz >= 0) {
                    int inew = ix + iy * nx + iz * nxy;
                    if (t1 < btol) {
                        int itest = howfar_in(inew, tx, ty, tz, u, v, w, t1);
                        if (itest == -1 && t1 < btol) return -1;
                    }
                    t = t1;
                    return inew;
                }
            }
        }
        iy = -1;
        if (y <= yp[0] && v > 0) { t1 = (yp[0] - y) / v; iy = 0; }
        else if (y >= yp[ny] && v < 0) { t1 = (yp[ny] - y) / v; iy = ny - 1; }
        if (iy >= 0 && t1 <= t) {
            R tx = x + u * t1, tz = z + w * t1, ty = y + v * t1;
            ix = planes_where(tx, nx, xp);
            if (ix >= 0) {
                iz = planes_where(tz, nz, zp);
                if (iz >= 0) {
                    int inew = ix + iy * nx + iz * nxy;
                    if (t1 < btol) {
                        int itest = howfar_in(inew, tx, ty, tz, u, v, w, t1);
                        if (itest == -1 && t1 < btol) return -1;
                    }
                    t = t1;
                    return inew;
                }
            }
        }
        iz = -1;
        if (z <= zp[0] && w > 0) { t1 = (zp[0] - z) / w; iz = 0; }
        else if (z >= zp[nz] && w < 0) { t1 = (zp[nz] - z) / w; iz = nz - 1; }
        if (iz >= 0 && t1 <= t) {
            R tx = x + u * t1, ty = y + v * t1, tz = z + w * t1;
            ix = planes_where(tx, nx, xp);
            if (ix >= 0) {
                iy = planes_where(ty, ny, yp);
                if (iy >= 0) {
                    int inew = ix + iy * nx + iz * nxy;
                    if (t1 < btol) {
                        int itest = howfar_in(inew, tx, ty, tz, u, v, w, t1);
                        if (itest == -1 && t1 < btol) return -1;
                    }
                    t = t1;
                    return inew;
                }
            }
        }
        return -1;
    }
    int howfar(int ireg, R x, R y, R z, R u, R v, R w, R &t) const {
        return ireg >= 0 ? howfar_in(ireg, x, y, z, u, v, w, t)
                         : howfar_out(x, y, z, u, v, w, t);
    }
    // egs_nd_geometry.h:1106-1148 + egs_planes.h:430-447
    static R planes_near(int i, R x, int n, const std::vector<R> &p) {
        R t = x - p[i];
        if (i + 1 <= n) {
            R t2 = p[i + 1] - x;
            if (t2 < t) t = t2;
        }
        return t;
    }
    static bool planes_inside(R x, int n, const std::vector<R> &p) {
        return !(x < p[0] || x > p[n]);
    }
    static R planes_near_out(R x, int n, const std::vector<R> &p) {
        if (x <= p[0]) return p[0] - x;
        if (x >= p[n]) return x - p[n];
        return 0;
    }
    R hownear(int ireg, R x, R y, R z) const {
        if (ireg >= 0) {
            int iz = ireg / nxy;
            int ir = ireg - iz * nxy;
            int iy = ir / nx;
            int ix = ir - iy * nx;
            R t = planes_near(ix, x, nx, xp);
            R t1 = planes_near(iy, y, ny, yp);
            if (t1 < t) t = t1;
            t1 = planes_near(iz, z, nz, zp);
            if (t1 < t) t = t1;
            return t;
        }
        int nc = 0;
        R s1 = 0, s2 = 0;
        if (!planes_inside(x, nx, xp)) { R t = planes_near_out(x, nx, xp); nc++; s1 += t; s2 += t * t; }
        if (!planes_inside(y, ny, yp)) { R t = planes_near_out(y, ny, yp); nc++; s1 += t; s2 += t * t; }
        if (!planes_inside(z, nz, zp)) { R t = planes_near_out(z, nz, zp); nc++; s1 += t; s2 += t * t; }
        if (nc == 1) return s1;
        return std::sqrt(s2);
    }
};

}  // namespace oracle
