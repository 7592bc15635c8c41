"""Voxel stepping (EGS_XYZGeometry howfar/hownear/isWhere/howfarFromOut,
egs++/geometry/egs_nd_geometry/egs_nd_geometry.h:791-1381).

CPU: the oracle is pinned against exact rational ray casting (brute force
over all plane crossings with fractions.Fraction) and geometric invariants.
GPU: device voxel and step sequences are bit-exact with the oracle in both
the FP32 transport build and the FP64 build."""
from fractions import Fraction

import numpy as np
import pytest

from egsnrc_amd._lib import make_geometry
from tests import oracle_lib


def small_grid():
    # non-uniform planes, all dyadic rationals
    xb = np.array([-2.0, -1.25, -0.5, 0.0, 0.75, 2.0])
    yb = np.array([-1.0, -0.5, 0.0, 0.5, 1.0])
    zb = np.array([0.0, 0.5, 1.0, 2.0, 3.5])
    return xb, yb, zb


def ray_corpus(xb, yb, zb, n, seed):
    """Random rays from inside and outside plus special cases."""
    rng = np.random.default_rng(seed)
    lo = np.array([xb[0], yb[0], zb[0]])
    hi = np.array([xb[-1], yb[-1], zb[-1]])
    ext = hi - lo
    x = lo - 0.3 * ext + rng.random((n, 3)) * 1.6 * ext
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    # aim half of the outside rays at the box centre so that they enter
    c = 0.5 * (lo + hi) + (rng.random((n, 3)) - 0.5) * ext
    aim = rng.random(n) < 0.5
    d = c - x
    d /= np.linalg.norm(d, axis=1)[:, None]
    u[aim] = d[aim]
    return x, u


def special_rays(xb, yb, zb):
    xs, us = [], []
    cx, cy = 0.5 * (xb[0] + xb[-1]), 0.5 * (yb[0] + yb[-1])
    # axis aligned, inside and from outside, both senses
    for ax in range(3):
        for sgn in (1, -1):
            u = [0.0, 0.0, 0.0]; u[ax] = float(sgn)
            xs.append([cx + 0.1, cy + 0.1, 0.25]); us.append(u)
            p = [cx + 0.1, cy + 0.1, 0.25]
            p[ax] = ([xb, yb, zb][ax][0] - 1.0) if sgn > 0 else ([xb, yb, zb][ax][-1] + 1.0)
            xs.append(p); us.append(u)
    # starting exactly on a plane, moving across and along it
    xs += [[0.0, 0.25, 0.75], [0.0, 0.25, 0.75], [-0.5, 0.0, 1.0]]
    us += [[1.0, 0.0, 0.0], [0.0, 0.0, 1.0], [0.0, 0.0, -1.0]]
    # through plane intersections (edges/corners): 45 degree rays
    s = np.sqrt(0.5)
    xs += [[-0.5, -0.5, 0.25], [-2.5, -1.5, 0.3], [0.0, 0.0, -0.5]]
    us += [[s, s, 0.0], [s, s, 0.0], [0.0, s, s]]
    # grazing: tiny direction component
    g = 1e-7
    v = np.array([1.0, g, g]); v /= np.linalg.norm(v)
    xs.append([-3.0, 0.1, 0.6]); us.append(list(v))
    # entering at a face exactly, on an edge of the box
    xs.append([xb[0] - 1.0, yb[0], zb[0] + 0.25]); us.append([1.0, 0.0, 0.0])
    # a ray missing the box
    xs.append([xb[-1] + 1.0, 0.0, 0.5]); us.append([1.0, 0.0, 0.0])
    return np.array(xs, float), np.array(us, float)


def exact_sequence(xb, yb, zb, x, u):
    """Regions crossed by the ray x + s u (s>=0), by exact arithmetic.
    Returns None when two crossings coincide (a tie)."""
    X = [Fraction(v) for v in x]
    U = [Fraction(v) for v in u]
    planes = [[Fraction(v) for v in p] for p in (xb, yb, zb)]
    n = [len(p) - 1 for p in planes]
    events = []
    for a in range(3):
        if U[a] == 0:
            continue
        for p in planes[a]:
            s = (p - X[a]) / U[a]
            if s > 0:
                events.append((s, a))
    events.sort()
    for i in range(1, len(events)):
        if events[i][0] - events[i - 1][0] < Fraction(1, 10**6):
            return None

    def region(s):
        idx = []
        for a in range(3):
            c = X[a] + U[a] * s
            p = planes[a]
            if c < p[0] or c > p[-1]:
                return -1
            k = 0
            while k + 1 < n[a] and c > p[k + 1]:
                k += 1
            idx.append(k)
        return idx[0] + idx[1] * n[0] + idx[2] * n[0] * n[1]

    seq = []
    cur = region(Fraction(0))
    if cur >= 0:
        seq.append(cur)
    prev_s = Fraction(0)
    bounds = [s for s, _ in events] + [None]
    for k, (s, a) in enumerate(events):
        nxt = events[k + 1][0] if k + 1 < len(events) else s + 1
        r = region((s + nxt) / 2)
        if r != cur:
            if r >= 0 or cur >= 0:
                seq.append(r)
            cur = r
        if cur < 0 and seq:
            break
    return seq


def test_oracle_matches_exact_ray_casting():
    xb, yb, zb = small_grid()
    g, keep = make_geometry(xb, yb, zb)
    x, u = ray_corpus(xb, yb, zb, 400, seed=11)
    reg, st, ns = oracle_lib.trace_rays(g, x, u, 200, 64)
    checked = 0
    for r in range(len(x)):
        ex = exact_sequence(xb, yb, zb, x[r], u[r])
        if ex is None:
            continue
        got = list(reg[r, :ns[r]])
        assert got == ex, (r, got, ex)
        checked += 1
    assert checked > 300


def test_oracle_step_lengths_sum_to_chord():
    xb, yb, zb = small_grid()
    g, keep = make_geometry(xb, yb, zb)
    x, u = ray_corpus(xb, yb, zb, 300, seed=5)
    reg, st, ns = oracle_lib.trace_rays(g, x, u, 200, 64)
    lo = np.array([xb[0], yb[0], zb[0]]); hi = np.array([xb[-1], yb[-1], zb[-1]])
    for r in range(len(x)):
        if ns[r] == 0:
            continue
        with np.errstate(divide="ignore", invalid="ignore"):
            t1 = (lo - x[r]) / u[r]; t2 = (hi - x[r]) / u[r]
        tmin = np.nanmax(np.where(np.isfinite(np.minimum(t1, t2)), np.minimum(t1, t2), -np.inf))
        tmax = np.nanmin(np.where(np.isfinite(np.maximum(t1, t2)), np.maximum(t1, t2), np.inf))
        inside0 = reg[r, 0] >= 0 and st[r, 0] == 0
        entry = 0.0 if inside0 else st[r, 0]
        chord = tmax - max(tmin, 0.0)
        total = st[r, 1:ns[r]].sum()
        assert abs(total - chord) < 1e-9 * max(1, chord), (r, total, chord)
        if not inside0:
            assert abs(entry - tmin) < 1e-9


def test_oracle_sequence_is_face_connected():
    xb, yb, zb = small_grid()
    nx, ny = len(xb) - 1, len(yb) - 1
    g, keep = make_geometry(xb, yb, zb)
    x, u = ray_corpus(xb, yb, zb, 300, seed=9)
    reg, st, ns = oracle_lib.trace_rays(g, x, u, 200, 64)
    for r in range(len(x)):
        seq = [v for v in reg[r, :ns[r]] if v >= 0]
        for a, b in zip(seq, seq[1:]):
            assert abs(a - b) in (1, nx, nx * ny), (a, b)


def test_oracle_hownear_closed_form():
    xb, yb, zb = small_grid()
    g, keep = make_geometry(xb, yb, zb)
    rng = np.random.default_rng(3)
    pts = np.column_stack([rng.uniform(xb[0], xb[-1], 500),
                           rng.uniform(yb[0], yb[-1], 500),
                           rng.uniform(zb[0], zb[-1], 500)])
    ireg = oracle_lib.is_where(g, pts, 64)
    t = oracle_lib.hownear(g, pts, ireg, 64)
    for p, tt in zip(pts, t):
        d = min(np.min(np.abs(pl - c)) for pl, c in zip((xb, yb, zb), p))
        assert abs(d - tt) < 1e-12
    # outside points: distance to the box (1 axis) or Euclidean (2-3 axes)
    out = np.array([[-3.0, 0.0, 1.0], [-3.0, -2.0, 1.0], [3.0, 2.0, -1.0]])
    t = oracle_lib.hownear(g, out, np.full(3, -1), 64)
    assert np.allclose(t, [1.0, np.hypot(1.0, 1.0), np.sqrt(3.0)])


def test_oracle_is_where_boundaries():
    xb, yb, zb = small_grid()
    g, keep = make_geometry(xb, yb, zb)
    nx, ny = len(xb) - 1, len(yb) - 1
    # on an interior plane the lower voxel wins (findRegion uses x <= p[m])
    pts = np.array([[0.0, 0.25, 0.25], [xb[0], yb[0], zb[0]], [xb[-1], yb[-1], zb[-1]],
                    [xb[-1] + 1e-9, 0, 0.5]])
    r = oracle_lib.is_where(g, pts, 64)
    assert r[0] == 2 + 2 * nx
    assert r[1] == 0
    assert r[2] == (nx - 1) + (ny - 1) * nx + (len(zb) - 2) * nx * ny
    assert r[3] == -1


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [32, 64])
def test_device_trace_bit_exact(engine, precision):
    cases = [small_grid(),
             (np.linspace(-15, 15, 31), np.linspace(-15, 15, 31), np.linspace(0, 30, 31)),
             (np.linspace(-16, 16, 129), np.linspace(-16, 16, 129), np.linspace(0, 32, 129))]
    for ci, (xb, yb, zb) in enumerate(cases):
        g, keep = make_geometry(xb, yb, zb)
        x, u = ray_corpus(xb, yb, zb, 3000, seed=100 + ci)
        xs, us = special_rays(xb, yb, zb)
        x = np.vstack([x, xs]); u = np.vstack([u, us])
        ms = 3 * (len(xb) + len(yb) + len(zb)) + 8
        ro, so, no = oracle_lib.trace_rays(g, x, u, ms, precision)
        rd, sd, nd = engine.trace_rays(g, x, u, ms, precision)
        assert np.array_equal(no, nd), ci
        for r in range(len(x)):
            k = no[r]
            assert np.array_equal(ro[r, :k], rd[r, :k]), (ci, r)
            assert np.array_equal(so[r, :k], sd[r, :k]), (ci, r)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [32, 64])
def test_device_hownear_bit_exact(engine, precision):
    xb, yb, zb = small_grid()
    g, keep = make_geometry(xb, yb, zb)
    rng = np.random.default_rng(17)
    pts = np.column_stack([rng.uniform(xb[0] - 1, xb[-1] + 1, 4000),
                           rng.uniform(yb[0] - 1, yb[-1] + 1, 4000),
                           rng.uniform(zb[0] - 1, zb[-1] + 1, 4000)])
    ireg = oracle_lib.is_where(g, pts, precision)
    a = oracle_lib.hownear(g, pts, ireg, precision)
    b = engine.hownear(g, pts, ireg, precision)
    assert np.array_equal(a, b)
