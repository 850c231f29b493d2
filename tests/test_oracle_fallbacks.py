"""Pins of the oracle's fallback branches and of its self-consistent field
path, by what geometry and the closed forms fix.

  * GetEMFieldDW (pic_interpolation_tools.f90:458-572), reached when
    RefElemNewton gives up for a particle far outside its element with
    cell_volweight_mean deposition (:248-249):
      - a constant field comes back exactly;
      - in general the field is the inverse-distance mean over the Gauss
        points, with the distances taken from the closed-form Gauss-point
        positions of a Cartesian box;
      - on a Gauss point the EXIT leaves only the innermost loop (DO k; DO l;
        DO m): for N = 1 the expected value is worked out by hand from which
        Gauss points the loop visits after the hit.  Leaving all three loops,
        not leaving at all, or swapping the loop order each changes it.
  * The inverse-distance fallback of cell_volweight_mean
    (pic_depo_method.f90:512-539): the eight corner charges, recovered from
    the N = 1 source of one particle, sum to q*MPF and are proportional to
    the inverse distances to the corners.  (Its zero-distance branch needs a
    particle on a corner of the element Newton failed for; Newton solves a
    corner exactly, so that branch is not reachable.)
  * The self-consistent field path om_sc_field_from_source: a per-element
    charge mode cos(k x) gives the Gauss-point field of the closed-form node
    field rho0/(eps0 k) sin(k x_node), trilinearly interpolated.
"""
import numpy as np
import pytest

from piclas_amd import adjoint as A
from piclas_amd import mesh as M
from piclas_amd import workloads as W
from piclas_amd.stage import EPS0


def _far_elem(m, x, shift=(2, 2, 2)):
    """1-based id of an element `shift` cells away (periodic wrap) from the owner of x."""
    nx, ny, nz = m.shape
    own = m.element_of(x) - 1
    i, j, k = own % nx, (own // nx) % ny, own // (nx * ny)
    i, j, k = (i + shift[0]) % nx, (j + shift[1]) % ny, (k + shift[2]) % nz
    return (i + nx * (j + ny * k) + 1).astype(np.int32)


def _cvwm_oracle(oracle_mod, m, N, ext=(0, 0, 0, 0, 0, 0)):
    return oracle_mod.Oracle(m, N, deposition=oracle_mod.DEPO_CVWM, charge=(2.0,), mass=(1.0,), mpf=(1.5,),
                             external_field=ext)


def test_dw_fallback_constant_field(oracle_mod):
    m = M.cartesian_box((5, 5, 5))
    o = _cvwm_oracle(oracle_mod, m, 3, ext=(1.0, -2.0, 0.5, 0, 0, 0.01))
    E0 = np.array([3.0, -4.0, 12.0])
    o.set_field(np.broadcast_to(E0, (m.nElems, 64, 3)).copy())
    rng = np.random.default_rng(21)
    x = rng.uniform(0, 1, (200, 3))
    far = _far_elem(m, x)
    _, ok = o.position_in_ref_elem(x, far)
    assert not ok.any()                      # Newton gives up for every particle
    o.set_particles(np.hstack([x, np.zeros_like(x)]), elem=far)
    o.interpolate()
    F = o.field_at_particle()
    assert np.abs(F[:, :3] - (E0 + [1.0, -2.0, 0.5])).max() <= 4e-15 * 13
    assert np.array_equal(F[:, 3:], np.broadcast_to([0, 0, 0.01], (200, 3)))


def test_dw_fallback_inverse_distance_mean(oracle_mod):
    m = M.cartesian_box((4, 4, 4))
    o = _cvwm_oracle(oracle_mod, m, 3)
    rng = np.random.default_rng(22)
    E = rng.normal(size=(m.nElems, 64, 3))
    o.set_field(E)
    xg = W.dof_positions(m, 3)              # closed-form Gauss points, r = (k*4+j)*4+i
    x = rng.uniform(0, 1, (100, 3))
    far = _far_elem(m, x)
    o.set_particles(np.hstack([x, np.zeros_like(x)]), elem=far)
    o.interpolate()
    F = o.field_at_particle()[:, :3]
    for p in range(len(x)):
        e = far[p] - 1
        w = 1.0 / np.linalg.norm(xg[e] - x[p], axis=1)
        ref = (w[:, None] * E[e]).sum(0) / w.sum()
        assert np.abs(F[p] - ref).max() <= 1e-14 * np.abs(E[e]).max()
    # particles inside their element take the Lagrange interpolation instead
    own = m.element_of(x)
    o.set_particles(np.hstack([x, np.zeros_like(x)]), elem=own)
    o.interpolate()
    xi, ok = o.position_in_ref_elem(x, own)
    assert ok.all()
    assert np.abs(o.field_at_particle()[:, :3] - o.eval_field_at_ref_pos(xi, own)).max() == 0.0


def test_dw_gauss_point_exit_leaves_innermost_loop(oracle_mod):
    """N = 1, one unit-cube element: Gauss points g-/g+ = 1/2 -+ 1/(2 sqrt 3),
    spacing s = 1/sqrt 3.  Point (k,l,m) = (x, y, z) index."""
    m = M.cartesian_box((1, 1, 1))
    o = _cvwm_oracle(oracle_mod, m, 1)
    U = (1.0 + np.arange(8))[:, None] * np.array([1.0, 10.0, 100.0])   # U at r = (m*2+l)*2+k
    o.set_field(U[None])
    s = 1.0 / np.sqrt(3)

    def Ur(k, l, mm):
        return U[(mm * 2 + l) * 2 + k]

    xgp = o.elem_xgp()[0]                    # the positions the oracle holds, so the hit is exact
    assert np.abs(xgp - W.dof_positions(m, 1)[0]).max() <= 1e-15

    def at(k, l, mm):
        return xgp[(mm * 2 + l) * 2 + k][None, :]

    one = np.ones(1, np.int32)
    # hit (1,0,0): (1,0,1) is skipped by the EXIT; (1,1,0) at s and (1,1,1) at s*sqrt2 follow
    got = o.em_field_dw(at(1, 0, 0), one)[0]
    ref = (Ur(1, 0, 0) + Ur(1, 1, 0) / s + Ur(1, 1, 1) / (s * np.sqrt(2))) / (1 + 1 / s + 1 / (s * np.sqrt(2)))
    assert np.abs(got - ref).max() <= 1e-14 * np.abs(ref).max()
    # hit (1,1,0): the last (k,l) row, nothing follows
    assert np.array_equal(o.em_field_dw(at(1, 1, 0), one)[0], Ur(1, 1, 0))
    # hit (0,0,1): the rows (0,1), (1,0), (1,1) follow in full
    later = {(0, 1, 0): s * np.sqrt(2), (0, 1, 1): s, (1, 0, 0): s * np.sqrt(2), (1, 0, 1): s,
             (1, 1, 0): s * np.sqrt(3), (1, 1, 1): s * np.sqrt(2)}
    num = Ur(0, 0, 1) + sum(Ur(*p) / d for p, d in later.items())
    den = 1.0 + sum(1 / d for d in later.values())
    got = o.em_field_dw(at(0, 0, 1), one)[0]
    assert np.abs(got - num / den).max() <= 1e-14 * np.abs(num / den).max()
    # leaving all three loops at the hit would return the hit's value alone
    alt_all_loops = Ur(0, 0, 1)
    assert np.abs(got - alt_all_loops).max() > 1e-3 * np.abs(got).max()


def _corner_charges_from_n1_source(PS_e):
    """Invert the N = 1 node -> Gauss point trilinear map of one element:
    PS_e [8][4] at r = (k*2+j)*2+i -> node values [8][4] in CGNS corner order."""
    gm, gp = 0.5 - 0.5 / np.sqrt(3), 0.5 + 0.5 / np.sqrt(3)
    a = (gm, gp)
    cg = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    T = np.zeros((8, 8))
    for k in range(2):
        for j in range(2):
            for i in range(2):
                r = (k * 2 + j) * 2 + i
                for c, (ci, cj, ck) in enumerate(cg):
                    T[r, c] = (a[i] if ci else 1 - a[i]) * (a[j] if cj else 1 - a[j]) * (a[k] if ck else 1 - a[k])
    return np.linalg.solve(T, PS_e), cg


def test_cvwm_inverse_distance_fallback(oracle_mod):
    """One particle in element (0,0,0) of an open 4^3 box of width 1, given
    element (2,2,2): RefElemNewton gives up, the charge goes to the eight
    corners of element (2,2,2) by inverse distance."""
    n, h = 4, 0.25
    m = M.cartesian_box((n, n, n), periodic=(False, False, False))
    o = _cvwm_oracle(oracle_mod, m, 1)
    x = np.array([[0.11, 0.07, 0.16]])
    v = np.array([[3.0, -1.0, 2.0]])
    e = 2 + n * (2 + n * 2)
    _, ok = o.position_in_ref_elem(x, np.array([e + 1], np.int32))
    assert not ok[0]
    o.set_particles(np.hstack([x, v]), elem=np.array([e + 1], np.int32))
    o.deposition()
    PS = o.part_source()
    vals, cg = _corner_charges_from_n1_source(PS[e])
    # NodeVolume of an interior node of a box of linear hexahedra: 8 elements x h^3/8
    nodes = np.array([[(2 + ci) * h, (2 + cj) * h, (2 + ck) * h] for ci, cj, ck in cg])
    q = vals * h ** 3
    Q = 2.0 * 1.5
    w = 1.0 / np.linalg.norm(nodes - x[0], axis=1)
    ref = (w / w.sum())[:, None] * np.array([v[0, 0] * Q, v[0, 1] * Q, v[0, 2] * Q, Q])[None, :]
    assert abs(q[:, 3].sum() - Q) <= 1e-12 * Q
    assert np.abs(q - ref).max() <= 1e-12 * np.abs(ref).max()
    assert abs(o.deposited_charge() - Q) <= 1e-12 * Q


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_sc_field_from_source_single_mode(oracle_mod, axis):
    """Element-wise constant charge rho0 cos(k c_axis) (c = element centre):
    element means are exact, the operator gives rho0/(eps0 k) sin(k x) at
    the corner nodes, and each Gauss point takes the linear interpolation
    between its two node planes along the axis."""
    n, N = 8, 3
    m = M.cartesian_box((n, n, n))
    o = oracle_mod.Oracle(m, N, deposition=oracle_mod.DEPO_SF, sf_radius=1.5 / n, sf_alpha=3)
    o.sc_set_operator(A.periodic_poisson_operator(m))
    e = np.arange(n ** 3)
    idx = [e % n, (e // n) % n, e // (n * n)][axis]
    rho0, k, h = 1e-9, 2 * np.pi, 1.0 / n
    rho_e = rho0 * np.cos(k * (idx + 0.5) * h)
    PS = np.zeros((n ** 3, (N + 1) ** 3, 4))
    PS[..., 3] = rho_e[:, None]
    PS[..., :3] = 7.0                        # current components play no part
    o.set_part_source(PS)
    o.sc_field_from_source()
    E = o.field()
    xg = W.dof_positions(m, N)[..., axis]    # closed-form Gauss-point coordinate along the axis
    lo = idx[:, None] * h
    a = (xg - lo) / h
    ref = rho0 / (EPS0 * k) * ((1 - a) * np.sin(k * lo) + a * np.sin(k * (lo + h)))
    scale = np.abs(ref).max()
    assert np.abs(E[..., axis] - ref).max() <= 1e-12 * scale
    others = [c for c in range(3) if c != axis]
    assert np.abs(E[..., others]).max() <= 1e-12 * scale
