"""Pins of the oracle's RefMapping boundary tracking (ParticleBCTracking,
particle_reftracking.f90:417-694, and the face intersections of
particle_intersection.f90) by what the geometry fixes:

  * the Bezier control points of every boundary side reproduce the side's
    HOPR nodes when the Bernstein patch is evaluated at the equidistant
    parameters, and the patch normal points out of the element;
  * SideType: the annulus walls are CURVED, its z faces PLANAR_CURVED, the
    faces of a Cartesian box PLANAR_RECT;
  * the intersection of a segment with a curved wall equals an independent
    Newton solve of X(xi, eta) = last + alpha traj on the same Bernstein patch
    (to the Bezier clipping tolerance), a segment that stops short has none,
    and a chord that crosses the inner wall twice reports the first crossing;
    on a planar z face alpha is the closed form (z_face - z_last) / traj_z;
  * every element's list of boundary sides in range is sorted by
    |bary - origin| - ElemRadiusNGeo - side radius and holds its own sides;
  * a particle whose straight path leaves through the inner wall and comes
    back into the annulus within one step is removed (the first crossing is
    an open wall), as the reference's tracking does.
"""
import itertools
from math import comb

import numpy as np
import pytest

from piclas_amd import mesh as M

SHAPE = (3, 12, 3)
ST_PLANAR_RECT, ST_PLANAR_CURVED, ST_CURVED = 0, 3, 4


def bernstein(N, t):
    return np.array([comb(N, p) * ((1 + t) / 2) ** p * ((1 - t) / 2) ** (N - p) for p in range(N + 1)])


def patch(P, xi, eta):
    N = P.shape[0] - 1
    return np.einsum("p,q,pqc->c", bernstein(N, xi), bernstein(N, eta), P)


def patch_grad(P, xi, eta, h=1e-7):
    return ((patch(P, xi + h, eta) - patch(P, xi - h, eta)) / (2 * h),
            (patch(P, xi, eta + h) - patch(P, xi, eta - h)) / (2 * h))


def solve_line_patch(P, last, x):
    """All (alpha, xi, eta) with X(xi,eta) = last + alpha traj, xi, eta in [-1,1],
    0 <= alpha <= |x - last|, by Newton from a grid of starts."""
    d = x - last
    L = np.linalg.norm(d)
    t = d / L
    roots = []
    for s0, e0 in itertools.product(np.linspace(-1, 1, 7), repeat=2):
        X = np.array([s0, e0, np.dot(patch(P, s0, e0) - last, t)])
        for _ in range(60):
            F = patch(P, X[0], X[1]) - last - X[2] * t
            gx, ge = patch_grad(P, X[0], X[1])
            J = np.stack([gx, ge, -t], 1)
            try:
                dX = np.linalg.solve(J, F)
            except np.linalg.LinAlgError:
                break
            X = X - dX
            if np.abs(dX).max() < 1e-15:
                break
        F = patch(P, X[0], X[1]) - last - X[2] * t
        if np.abs(F).max() < 1e-12 and abs(X[0]) <= 1 + 1e-9 and abs(X[1]) <= 1 + 1e-9 and -1e-12 <= X[2] <= L:
            if not any(abs(X[2] - r[0]) < 1e-9 for r in roots):
                roots.append(tuple(X[[2, 0, 1]]))
    return sorted(roots)


def annulus_oracle(oracle_mod, shape=SHAPE):
    m = M.annulus(shape)
    o = oracle_mod.Oracle(m, 3, tracking=oracle_mod.TRACK_REFMAPPING, refmapping_guess=2, charge=(0.0,),
                          mass=(1.0,), mpf=(1.0,))
    return m, o


def face_nodes(m, side_row):
    """HOPR nodes of the face of a side row (elements store (NGeo+1)^3 nodes, i fastest)."""
    g1 = m.NGeo + 1
    e = np.searchsorted(m.ElemInfo[:, 3], side_row, side="right")
    ls = side_row - m.ElemInfo[e, 2] + 1
    X = m.NodeCoords[m.ElemInfo[e, 4]:m.ElemInfo[e, 5]].reshape(g1, g1, g1, 3)   # [k][j][i]
    ng = g1 - 1
    sl = {5: X[:, :, 0], 3: X[:, :, ng], 2: X[:, 0, :], 4: X[:, ng, :], 1: X[0], 6: X[ng]}[ls]
    return e, ls, sl.reshape(-1, 3)


def test_bezier_points_interpolate_face_nodes_and_normals_point_out(oracle_mod):
    m, o = annulus_oracle(oracle_mod)
    side, typ, bez, org, rad = o.bc_sides()
    assert len(side) == 2 * SHAPE[1] * SHAPE[2] + 2 * SHAPE[0] * SHAPE[1]
    ts = np.array([-1.0, 0.0, 1.0])
    for b in range(len(side)):
        e, ls, nodes = face_nodes(m, side[b])
        ev = np.array([patch(bez[b], a, c) for a in ts for c in ts])
        # the same point set (the side orientation permutes them)
        dmat = np.linalg.norm(ev[:, None, :] - nodes[None, :, :], axis=2)
        assert dmat.min(1).max() < 1e-14 and dmat.min(0).max() < 1e-14
        gx, ge = patch_grad(bez[b], 0.1, -0.2)
        n = np.cross(gx, ge)
        c = patch(bez[b], 0.1, -0.2)
        r = np.hypot(c[0], c[1])
        if typ[b] == ST_CURVED:   # walls: outward is -r on the inner, +r on the outer cylinder
            radial = np.array([c[0], c[1], 0.0]) / r
            assert np.sign(np.dot(n, radial)) == (1 if r > 0.75 else -1)
        else:
            assert typ[b] == ST_PLANAR_CURVED
            assert np.sign(n[2]) == (1 if c[2] > 0.5 else -1)
        assert np.allclose(org[b], patch(bez[b], 0.0, 0.0), atol=1e-15)
        assert np.isclose(rad[b], np.linalg.norm(bez[b].reshape(-1, 3) - org[b], axis=1).max())


def test_cartesian_box_faces_are_planar_rect(oracle_mod):
    m = M.cartesian_box((3, 3, 3), periodic=(False, False, False), wall_bc=M.PART_BC_OPEN)
    o = oracle_mod.Oracle(m, 3, tracking=oracle_mod.TRACK_REFMAPPING, charge=(0.0,), mass=(1.0,), mpf=(1.0,))
    side, typ, bez, org, rad = o.bc_sides()
    assert len(side) == 6 * 9 and (typ == ST_PLANAR_RECT).all()


def test_curved_wall_intersection_matches_independent_solve(oracle_mod):
    m, o = annulus_oracle(oracle_mod)
    side, typ, bez, org, rad = o.bc_sides()
    rng = np.random.default_rng(3)
    walls = np.where(typ == ST_CURVED)[0]
    checked = hits = 0
    for b in rng.choice(walls, 24, replace=False):
        c = patch(bez[b], rng.uniform(-0.8, 0.8), rng.uniform(-0.8, 0.8))
        radial = np.array([c[0], c[1], 0.0]) / np.hypot(c[0], c[1])
        outer = np.hypot(c[0], c[1]) > 0.75
        into = radial if outer else -radial          # out of the domain through this wall
        for reach in (1.4, 0.6):                      # crossing, stopping short
            tang = np.cross(radial, [0, 0, 1.0]) * rng.uniform(-0.2, 0.2) + np.array([0, 0, rng.uniform(-.2, .2)])
            last = c - 0.01 * into + 0.01 * tang
            x = last + (0.01 * into - 0.01 * tang) * reach + 0.002 * tang
            a, xi, eta = o.bc_side_intersection(b, last, x)
            ref = solve_line_patch(bez[b], last, x)
            if ref:
                hits += 1
                assert abs(a - ref[0][0]) <= 1e-7 * rad[b], (a, ref)
                assert np.allclose(patch(bez[b], xi, eta), last + a * (x - last) / np.linalg.norm(x - last),
                                   atol=1e-7 * rad[b])
            else:
                assert a == -1.0
            checked += 1
    assert hits >= 20 and checked == 48


def test_chord_through_inner_hole_reports_first_crossing(oracle_mod):
    m, o = annulus_oracle(oracle_mod)
    side, typ, bez, org, rad = o.bc_sides()
    # a chord at x = 0.49 from inside the annulus (r = 0.529) through the hole
    # (r = 0.49 at y = 0) back into the annulus
    last, x = np.array([0.49, -0.2, 0.5]), np.array([0.49, 0.2, 0.5])
    crossings = []
    for b in np.where(typ == ST_CURVED)[0]:
        a, xi, eta = o.bc_side_intersection(b, last, x)
        if a > -1:
            crossings.append(a)
            ref = solve_line_patch(bez[b], last, x)
            assert ref and abs(a - ref[0][0]) <= 1e-7 * rad[b]
    assert len(crossings) >= 2
    first = min(crossings)
    # the circle r = 0.5 is met at y = -sqrt(0.25 - 0.49^2); the NGeo = 2 wall is within 1e-3 of it
    assert abs((last[1] + first) - (-np.sqrt(0.25 - 0.49 ** 2))) < 2e-3


def test_planar_z_face_closed_form(oracle_mod):
    m, o = annulus_oracle(oracle_mod)
    side, typ, bez, org, rad = o.bc_sides()
    rng = np.random.default_rng(4)
    for b in np.where(typ == ST_PLANAR_CURVED)[0][:12]:
        c = patch(bez[b], rng.uniform(-0.7, 0.7), rng.uniform(-0.7, 0.7))
        zf = c[2]
        sgn = 1.0 if zf > 0.5 else -1.0
        last = c - np.array([0.003, -0.002, 0.01 * sgn])
        x = last + np.array([0.001, 0.001, 0.02 * sgn])
        a, xi, eta = o.bc_side_intersection(b, last, x)
        t = (x - last) / np.linalg.norm(x - last)
        # BezierNewton stops at |dxi|^2 <= 1e-8: the point it returns lies on
        # the plane, off the crossing by what is left of the Newton step
        assert abs(a - (zf - last[2]) / t[2]) <= 1e-9


def test_element_lists_sorted_with_own_sides(oracle_mod):
    m, o = annulus_oracle(oracle_mod)
    side, typ, bez, org, rad = o.bc_sides()
    off, bs, dist = o.bc_elem_lists()
    bary, radius = o.elem_bary()
    elem_of_side = np.searchsorted(m.ElemInfo[:, 3], side, side="right")
    for e in range(m.nElems):
        lst, d = bs[off[e]:off[e + 1]], dist[off[e]:off[e + 1]]
        assert (np.diff(d) >= 0).all()
        own = np.where(elem_of_side == e)[0]
        assert set(own) <= set(lst)
        ref = np.linalg.norm(bary[e] - org[lst], axis=1) - radius[e] - rad[lst]
        assert np.allclose(d, ref, rtol=0, atol=1e-15)


def test_leave_and_reenter_is_removed(oracle_mod):
    m, o = annulus_oracle(oracle_mod)
    o.set_field(np.zeros((m.nElems, 64, 3)))
    # a straight path at x = 0.49 from y = -0.2 to y = +0.2: starts and ends in
    # the annulus (r = 0.529) and passes through the hole (r = 0.49 at y = 0)
    ps = np.array([[0.49, -0.2, 0.5, 0.0, 1.0, 0.0],
                   [0.60, -0.2, 0.5, 0.0, 1.0, 0.0]])   # r >= 0.6: stays inside
    o.set_particles(ps, elem=np.zeros(2, np.int32))
    assert o.locate() == 0
    o.time_step(0.4, 1)
    out, el, ins = o.particles()
    assert list(ins) == [False, True]
    assert np.allclose(out[1, :3], [0.60, 0.2, 0.5])
    assert o.lost() == 0


def side_normal_slab(P):
    """Unit normal of the side's normal slab (the cross product of the xi and
    eta edge directions, as SideSlabNormals) and the interval of n . X over
    the control points."""
    N = P.shape[0] - 1
    t = P[N, 0] - P[0, 0] + P[N, N] - P[0, N]
    t /= np.linalg.norm(t)
    u = P[0, N] - P[0, 0] + P[N, N] - P[N, 0]
    n = np.cross(t, u)
    n /= np.linalg.norm(n)
    proj = P.reshape(-1, 3) @ n
    return n, proj.min(), proj.max()


def test_steps_outside_normal_slab_hit_nothing(oracle_mod):
    """The device takes as boundary candidates only particles whose step
    reaches the normal slab of a listed side (the control points' interval of
    n . X, widened by 1e-3 of the side's extent).  Steps that stay outside
    it, on either side and in any direction within the slab's plane region,
    get no intersection from the oracle's routines, on the curved walls and
    the planar z faces; steps that cross the patch do reach the slab."""
    m, o = annulus_oracle(oracle_mod)
    side, typ, bez, org, rad = o.bc_sides()
    rng = np.random.default_rng(11)
    checked = 0
    for b in rng.choice(len(side), 40, replace=False):
        n, lo, hi = side_normal_slab(bez[b])
        ext = 2 * rad[b]
        for _ in range(6):
            above = rng.random() < 0.5
            gap = ext * rng.uniform(1e-3, 2e-2)
            c0 = bez[b].reshape(-1, 3).mean(0)
            c0 = c0 + (((hi + gap) if above else (lo - gap)) - c0 @ n) * n
            tang = rng.normal(size=3)
            tang -= (tang @ n) * n
            # steps that stay on one side of the slab: normal coordinate moves
            # away from it or parallel to it
            away = (1.0 if above else -1.0) * rng.uniform(0.0, 0.5) * ext
            last = c0 + rng.uniform(-0.5, 0.5) * ext * tang / np.linalg.norm(tang)
            x = last + rng.uniform(0.2, 1.5) * ext * tang / np.linalg.norm(tang) + away * n
            a, xi, eta = o.bc_side_intersection(b, last, x)
            assert a == -1.0, (b, typ[b], a, xi, eta)
            checked += 1
        # a step through the middle of the patch crosses it and reaches the slab
        c = patch(bez[b], 0.1, -0.1)
        last, x = c - 0.01 * ext * n, c + 0.01 * ext * n
        assert max(last @ n, x @ n) >= lo and min(last @ n, x @ n) <= hi
    assert checked == 240
