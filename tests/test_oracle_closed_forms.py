"""Pins of the CPU oracle by what the mathematics fixes.

Each check is chosen so that a plausible slip in the restatement (a dropped
term, a wrong sign, an index or operand transposed) fails it: nodes and
weights against numpy's Gauss-Legendre rule, Lagrange cardinality and
polynomial exactness, Newton round trips on affine and twisted hexahedra,
Boris invariants (|u| under pure B, the rotation angle, the exact momentum
update under pure E, second-order convergence), ownership against the
closed-form Cartesian owner, the TriaTracking tie-break keeping the side
crossed first (with a planted mistake that the check must catch), charge
conservation of cell_volweight_mean, a
uniform lattice giving uniform density, and the shape-function deposit
against a brute-force sum over all DOFs and periodic images.
"""
import numpy as np
import pytest

from piclas_amd import mesh as M
from piclas_amd import workloads as W

Q, MASS = -1.602176634e-19, 9.1093826e-31
C_LIGHT = 299792458.0


@pytest.mark.parametrize("N", [1, 2, 3, 4, 5, 7])
def test_gauss_nodes_weights(oracle_mod, N):
    x, w = oracle_mod.legendre_gauss_nodes(N)
    xr, wr = np.polynomial.legendre.leggauss(N + 1)
    assert np.allclose(x, xr, rtol=0, atol=4e-16)
    assert np.allclose(w, wr, rtol=0, atol=4e-15)


def test_cgl_nodes_and_bary_weights(oracle_mod):
    x = oracle_mod.cheby_gauss_lobatto_nodes(4)
    assert np.allclose(x, -np.cos(np.arange(5) / 4 * np.pi), atol=0)
    wb = oracle_mod.barycentric_weights(x)
    ref = np.array([1.0 / np.prod([x[j] - x[k] for k in range(5) if k != j]) for j in range(5)])
    assert np.allclose(wb, ref, rtol=1e-14)


@pytest.mark.parametrize("N", [1, 3, 5])
def test_lagrange_cardinal_and_exact(oracle_mod, N):
    x, _ = oracle_mod.legendre_gauss_nodes(N)
    wb = oracle_mod.barycentric_weights(x)
    for i in range(N + 1):
        L = oracle_mod.lagrange_polys(x[i], x, wb)
        assert np.array_equal(L, np.eye(N + 1)[i])
    for xv in (-0.93, -0.2, 0.31, 0.77, 1.0, -1.0):
        L = oracle_mod.lagrange_polys(xv, x, wb)
        assert abs(L.sum() - 1) < 1e-14
        assert abs(L @ x ** N - xv ** N) < 1e-13


def _twisted_cell():
    """One non-affine (twisted) hexahedron, periodic-free, open walls."""
    m = M.cartesian_box((1, 1, 1), periodic=(False, False, False))
    nc = m.NodeCoords.copy()
    nc[7] += np.array([0.12, -0.07, 0.09])   # node (1,1,1)
    nc[2] += np.array([-0.05, 0.04, 0.0])    # node (0,1,0)
    nc[5] += np.array([0.0, 0.06, -0.08])    # node (1,0,1)
    m.NodeCoords = nc
    return m


def _trilinear(nc, xi):
    a = (np.asarray(xi) + 1) / 2
    X = np.zeros(3)
    for t in range(8):
        i, j, k = t & 1, (t >> 1) & 1, (t >> 2) & 1
        X += nc[t] * (a[0] if i else 1 - a[0]) * (a[1] if j else 1 - a[1]) * (a[2] if k else 1 - a[2])
    return X


def test_newton_round_trip_twisted(oracle_mod):
    m = _twisted_cell()
    o = oracle_mod.Oracle(m, 3, refmapping_eps=1e-24)
    rng = np.random.default_rng(3)
    xi0 = rng.uniform(-0.95, 0.95, (200, 3))
    x = np.array([_trilinear(m.NodeCoords, q) for q in xi0])
    xi, ok = o.position_in_ref_elem(x, np.ones(len(x), np.int32))
    assert ok.all()
    assert np.abs(xi - xi0).max() < 1e-13


def test_newton_affine_exact(oracle_mod):
    m = M.cartesian_box((3, 2, 2), hi=(1.5, 1.0, 0.7))
    o = oracle_mod.Oracle(m, 3)
    rng = np.random.default_rng(4)
    x = rng.uniform(0, 1, (500, 3)) * np.array([1.5, 1.0, 0.7])
    el = m.element_of(x)
    xi, ok = o.position_in_ref_elem(x, el)
    h = np.array([0.5, 0.5, 0.35])
    e0 = el - 1
    ijk = np.stack([e0 % 3, (e0 // 3) % 2, e0 // 6], 1)
    ref = 2 * (x - ijk * h) / h - 1
    assert ok.all() and np.abs(xi - ref).max() < 1e-14


def test_field_evaluation_polynomial_exact(oracle_mod):
    N = 3
    m = M.cartesian_box((2, 2, 2), hi=(2.0, 2.0, 2.0))
    o = oracle_mod.Oracle(m, N)
    xg = W.dof_positions(m, N)
    f = lambda p: np.stack([p[..., 0] ** 3 - 2 * p[..., 1] * p[..., 2], p[..., 1] ** 2 * p[..., 0],
                            p[..., 2] ** 3 + p[..., 0] * p[..., 1] * p[..., 2]], -1)
    o.set_field(f(xg))
    rng = np.random.default_rng(5)
    x = rng.uniform(0, 2, (300, 3))
    el = m.element_of(x)
    xi, _ = o.position_in_ref_elem(x, el)
    U = o.eval_field_at_ref_pos(xi, el)
    assert np.abs(U - f(x)).max() < 1e-12


def _one_particle_oracle(O, v0, E=(0, 0, 0), B=(0, 0, 0), L=1e3):
    m = M.cartesian_box((2, 2, 2), lo=(-L, -L, -L), hi=(L, L, L))
    o = O.Oracle(m, 1, external_field=tuple(E) + tuple(B), charge=(Q,), mass=(MASS,))
    st = np.array([[1e-3, 2e-3, -1e-3, *v0]])
    o.set_particles(st, elem=m.element_of(st[:, :3]))
    return o


def test_boris_pure_b_rotation(oracle_mod):
    v0 = np.array([3e5, -1e5, 2e5])
    B = np.array([0.0, 0.0, 0.01])
    o = _one_particle_oracle(oracle_mod, v0, B=B)
    dt = 1e-10
    o.time_step(dt, 1)
    v1 = o.particles()[0][0, 3:]
    g = 1 / np.sqrt(1 - v0 @ v0 / C_LIGHT ** 2)
    assert abs(np.linalg.norm(v1) - np.linalg.norm(v0)) < 1e-15 * np.linalg.norm(v0) * 10
    # rotation about z by theta = q|B|dt/(gamma m) (sign from q<0: counter-clockwise)
    th = -Q * np.linalg.norm(B) * dt / (g * MASS)
    ang = np.arctan2(v1[1], v1[0]) - np.arctan2(v0[1], v0[0])
    assert abs(np.remainder(ang - th + np.pi, 2 * np.pi) - np.pi) < 1e-12
    assert abs(v1[2] - v0[2]) < 1e-9


def test_boris_pure_e_momentum(oracle_mod):
    v0 = np.array([1e6, 2e5, -3e5])
    E = np.array([1e3, -2e3, 5e2])
    o = _one_particle_oracle(oracle_mod, v0, E=E)
    dt = 2e-10
    o.time_step(dt, 1)
    v1 = o.particles()[0][0, 3:]
    u0 = v0 / np.sqrt(1 - v0 @ v0 / C_LIGHT ** 2)
    u1 = v1 / np.sqrt(1 - v1 @ v1 / C_LIGHT ** 2)
    assert np.abs(u1 - (u0 + Q / MASS * dt * E)).max() < 1e-9 * np.abs(u0).max()


def test_boris_second_order(oracle_mod):
    """Gyration in uniform B: the tan() rotation of the 508 push turns v by
    exactly omega*dt per step, so the orbit is a polygon; after a quarter
    period its distance to the exact circle falls as dt^2."""
    B = 0.01
    v0 = np.array([2e5, 0.0, 0.0])
    g = 1 / np.sqrt(1 - v0 @ v0 / C_LIGHT ** 2)
    wc = -Q * B / (g * MASS)
    T = 2 * np.pi / wc
    rL = np.linalg.norm(v0) / wc
    x0 = np.array([1e-3, 2e-3, -1e-3])
    errs = []
    for nstep in (256, 512):
        o = _one_particle_oracle(oracle_mod, v0, B=(0, 0, B))
        o.time_step(T / nstep, nstep // 4)
        x = o.particles()[0][0, :3]
        # v0 is v(-dt/2); the electron (q<0) gyrates counter-clockwise about z with
        # phase phi(s) = wc (s + dt/2); exact x(T/4) = x0 + integral of v over [0, T/4]
        dt = T / nstep
        p0, p1 = wc * dt / 2, wc * (T / 4 + dt / 2)
        exact = x0 + rL * np.array([np.sin(p1) - np.sin(p0), np.cos(p0) - np.cos(p1), 0.0])
        errs.append(np.linalg.norm(x - exact))
    assert 3.5 < errs[0] / errs[1] < 4.5


def test_tracking_closed_form_ownership(oracle_mod):
    m = M.cartesian_box((5, 4, 3), hi=(1.0, 0.8, 0.6))
    o = oracle_mod.Oracle(m, 1)
    rng = np.random.default_rng(6)
    n = 3000
    x = rng.uniform(0, 1, (n, 3)) * np.array([1.0, 0.8, 0.6])
    v = rng.normal(size=(n, 3)) * 4e7     # several elements per step
    o.set_particles(np.hstack([x, v]), elem=m.element_of(x))
    for _ in range(5):
        o.time_step(1e-9, 1)
        ps, el, ins = o.particles()
        assert ins.all()
        assert (ps[:, :3] >= -1e-12).all() and (ps[:, :3] <= np.array([1.0, 0.8, 0.6]) + 1e-12).all()
        assert np.array_equal(el, m.element_of(ps[:, :3]))


def test_tracking_open_boundary_removes(oracle_mod):
    m = M.cartesian_box((3, 3, 3), periodic=(False, False, False))
    o = oracle_mod.Oracle(m, 1)
    rng = np.random.default_rng(7)
    n = 2000
    x = rng.uniform(0.05, 0.95, (n, 3))
    v = rng.normal(size=(n, 3)) * 2e7     # below c: the push is relativistic
    dt = 1e-8
    o.set_particles(np.hstack([x, v]), elem=m.element_of(x))
    o.time_step(dt, 1)
    ps, el, ins = o.particles()
    x1 = x + v * dt
    exits = ((x1 < 0) | (x1 > 1)).any(1)
    assert np.array_equal(~ins, exits)
    assert o.lost() == 0          # removed at an open boundary, not lost
    assert np.array_equal(el[ins], m.element_of(ps[ins, :3]))


def _deformed_box(shape, amp):
    """Generated periodic unit box with its nodes displaced by a smooth
    periodic field (bilinear, partly concave faces)."""
    m = M.cartesian_box(shape)
    X = m.NodeCoords
    x, y, z = X[:, 0].copy(), X[:, 1].copy(), X[:, 2].copy()
    a = amp / shape[0]
    s, c = (lambda u: np.sin(2 * np.pi * u)), (lambda u: np.cos(2 * np.pi * u))
    X[:, 0] = x + a * s(x) * c(y) * c(z)
    X[:, 1] = y + a * c(x) * s(y) * c(z)
    X[:, 2] = z + a * c(x) * c(y) * s(z)
    return m


def _first_crossing_violations(o, m, log, tol=1e-12):
    """For each recorded multi-crossing event of the TriaTracking walk, the
    parameter s in (0, 1] at which the segment LastPartPos -> PartState
    crosses the plane of each candidate triangle, computed here from the node
    coordinates.  The reference keeps the candidate with the smallest
    det(PartState)/det(LastPartPos) (particle_triatracking.f90:309-405); as
    det is proportional to the signed distance from the triangle's plane,
    that ratio is r = -d_new/d_last and s = 1/(1 - r), so the kept side must
    be the one crossed first.  Returns (events with two or more crossed
    candidates whose s differ by more than 1e-9, events whose kept side is
    not the first crossed)."""
    sid = o.elem_side_node_id()
    X = m.NodeCoords
    distinct = bad = 0
    for r in log:
        if r[3] == 0:
            continue
        e, n, last, x = int(r[1]) - 1, int(r[2]), r[5:8], r[8:11]
        cand = []
        for k in range(n):
            ls, tn, flags = int(r[11 + 3 * k]), int(r[12 + 3 * k]), int(r[13 + 3 * k])
            if flags != 3:
                continue
            ids = sid[e, ls - 1]
            p1, p2, p3 = X[ids[0]], X[ids[tn]], X[ids[tn + 1]]
            nrm = np.cross(p2 - p1, p3 - p1)
            dl, dx = np.dot(last - p1, nrm), np.dot(x - p1, nrm)
            if dl != dx:
                cand.append((dl / (dl - dx), ls, tn))
        if len(cand) < 2:
            continue
        s_all = [c[0] for c in cand]
        s_kept = [c[0] for c in cand if (c[1], c[2]) == (int(r[3]), int(r[4]))][0]
        distinct += max(s_all) - min(s_all) > 1e-9
        bad += s_kept > min(s_all) + tol
    return distinct, bad


def test_tracking_tiebreak_first_crossing(oracle_mod):
    """The multi-crossing branch of TriaTracking keeps the side that the
    trajectory crosses first.  On a strongly deformed periodic box (faces
    split into two triangles, partly concave), moves of about 0.75 element
    widths produce walk steps where several sides register a crossing at
    different points of the trajectory; every kept side must be the first
    crossed, and a planted mistake (keep the side crossed last) must be
    caught on every such step."""
    m = _deformed_box((5, 5, 5), 0.3)
    rng = np.random.default_rng(3)
    n = 30000
    x = rng.uniform(0, 1, (n, 3))
    v = rng.normal(size=(n, 3)) * 1.5e5
    res = {}
    for plant in (0, 1):
        o = oracle_mod.Oracle(m, 1)
        o.plant_mistake(plant)
        o.tiebreak_log(200000)
        o.set_particles(np.hstack([x, v]), elem=np.zeros(n, np.int32))
        o.locate()
        for _ in range(3):
            o.time_step(1e-6, 1)
        log = o.tiebreak_log()
        assert len(log) == o.multi_crossings()
        res[plant] = _first_crossing_violations(o, m, log) + (o.lost(),)
    distinct, bad, lost = res[0]
    assert distinct >= 50, res
    assert bad == 0 and lost == 0, res
    assert res[1][1] == res[1][0] >= 50, res      # the planted mistake fails every distinct step


def test_cvwm_charge_conservation(oracle_mod):
    for shape, hi in [((5, 3, 4), (2.0, 1.0, 1.3)), ((3, 3, 3), (1, 1, 1))]:
        m = M.cartesian_box(shape, hi=hi)
        o = oracle_mod.Oracle(m, 3, deposition=oracle_mod.DEPO_CVWM, charge=(1.0,), mass=(1.0,), mpf=(3.0,))
        rng = np.random.default_rng(8)
        x = rng.uniform(0, 1, (4000, 3)) * np.array(hi)
        o.set_particles(np.hstack([x, rng.normal(size=(4000, 3))]), elem=m.element_of(x))
        o.deposition()
        assert abs(o.deposited_charge() - 12000.0) < 1e-12 * 12000.0
        PS = o.part_source()
        # current density integral = sum q v
        sj = o.sj()
        xg, wg = oracle_mod.legendre_gauss_nodes(3)
        wq = np.einsum("i,j,k->kji", wg, wg, wg).ravel()
        J = (PS[..., :3] * (wq / sj)[..., None]).sum((0, 1))
        assert np.allclose(J, 3.0 * o.particles()[0][:, 3:].sum(0), rtol=1e-12, atol=1e-9)


def test_cvwm_uniform_lattice(oracle_mod):
    m = M.cartesian_box((4, 4, 4))
    o = oracle_mod.Oracle(m, 2, deposition=oracle_mod.DEPO_CVWM, charge=(1.0,), mass=(1.0,), mpf=(1.0,))
    c = (np.arange(4) + 0.5) / 4
    x = np.stack(np.meshgrid(c, c, c, indexing="ij"), -1).reshape(-1, 3)
    o.set_particles(np.hstack([x, np.zeros_like(x)]), elem=m.element_of(x))
    o.deposition()
    rho = o.part_source()[..., 3]
    assert np.allclose(rho, 64.0, rtol=1e-13)


def _brute_sf(xp, vp, q, dofs, R, alpha, L):
    """(J_x, J_y, J_z, rho) at each DOF: sum over particles and 27 periodic
    images of w_sf q (v, 1) (1-r^2/R^2)^alpha."""
    from math import gamma, pi
    B = gamma(1.5) * gamma(alpha + 1) / gamma(alpha + 2.5)
    w = 1 / (2 * B * pi * R ** 3)
    qv1 = q * np.hstack([vp, np.ones((len(xp), 1))])            # (n, 4)
    src = np.zeros((len(dofs), 4))
    for s in np.array(np.meshgrid([-1, 0, 1], [-1, 0, 1], [-1, 0, 1], indexing="ij")).reshape(3, -1).T:
        d = dofs[:, None, :] - (xp[None, :, :] + s * L)
        r2 = (d ** 2).sum(-1)
        S = np.where(r2 <= R * R, (1 - r2 / R ** 2) ** alpha, 0.0)
        src += w * (S @ qv1)
    return src


@pytest.mark.parametrize("alpha", [2, 3])
def test_sf_brute_force(oracle_mod, alpha):
    """Charge and all three current components: a velocity component taken
    from the wrong column, or a dropped factor, fails one of them."""
    m = M.cartesian_box((4, 4, 4))
    R = 1.5 / 4
    o = oracle_mod.Oracle(m, 3, deposition=oracle_mod.DEPO_SF, sf_radius=R, sf_alpha=alpha, charge=(2.0,),
                          mass=(1.0,), mpf=(1.5,))
    rng = np.random.default_rng(9)
    x = rng.uniform(0, 1, (60, 3))
    v = rng.normal(size=(60, 3)) * np.array([1.0, 3.0, -7.0])
    o.set_particles(np.hstack([x, v]), elem=m.element_of(x))
    o.deposition()
    src = o.part_source().reshape(-1, 4)
    dofs = o.elem_xgp().reshape(-1, 3)
    ref = _brute_sf(x, v, 3.0, dofs, R, alpha, np.ones(3))
    for c in range(4):
        assert np.abs(src[:, c] - ref[:, c]).max() < 1e-12 * np.abs(ref[:, c]).max(), c


def test_sf_weight_normalises_kernel():
    """w_sf * integral of (1-r^2/R^2)^alpha over the ball = 1 (pic_depo_shapefunction_tools.f90:1243-1246)."""
    from math import gamma, pi
    for alpha in (2, 3, 5):
        R = 0.7
        B = gamma(1.5) * gamma(alpha + 1) / gamma(alpha + 2.5)
        w = 1 / (2. * B * alpha + 2 * B) * (alpha + 1.) / (pi * R ** 3)
        r = np.linspace(0, R, 200001)
        integral = np.trapezoid(4 * pi * r ** 2 * (1 - r ** 2 / R ** 2) ** alpha, r)
        assert abs(w * integral - 1) < 1e-9


def test_field_energy_constant_field(oracle_mod):
    m = M.cartesian_box((3, 2, 2), hi=(1.5, 1.0, 0.5))
    o = oracle_mod.Oracle(m, 2)
    E0 = np.array([3.0, -4.0, 12.0])
    o.set_field(np.broadcast_to(E0, (12, 27, 3)).copy())
    assert abs(o.field_energy() - 0.5 * 8.8541878176e-12 * 169.0 * 0.75) < 1e-12 * o.field_energy()
