"""Parity of the HIP particle stage with the CPU oracle (and with the
reference's own golden deposition), through the C ABI.

Bars:
  element ownership   bit-exact, except particles within 1e-12 of a face
  positions           |dx| <= 1e-12 * domain size          (FP64)
  velocities          |dv| <= 1e-12 * max|v|               (FP64)
  deposited source    |drho| <= 1e-12 * max|rho|            (FP64; the sums
                      run in another order, so bit equality is not expected)
"""
import os

import numpy as np
import pytest

from piclas_amd import mesh as M
from piclas_amd import stage, workloads as W

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def run_pair(O, w, steps, deposition=None, staged=False, **over):
    kw = W.stage_kwargs(w)
    if deposition is not None:
        kw["deposition"] = deposition
    kw.update(over)
    print("case %s steps=%d depo=%d n=%d staged=%s" % (w.name, steps, kw["deposition"], len(w.PartState), staged),
          flush=True)
    s = stage.PICStage(w.mesh, w.N, max_particles=len(w.PartState), **kw)
    s.set_field(w.E)
    s.set_particles(w.PartState, elem=w.elem, species=w.species)
    o = O.Oracle(w.mesh, w.N, deposition=kw["deposition"], sf_radius=kw["sf_radius"], sf_alpha=kw["sf_alpha"],
                 external_field=kw["external_field"], charge=kw["charge"], mass=kw["mass"], mpf=kw["mpf"],
                 tracking=kw["tracking"], refmapping_guess=kw["refmapping_guess"])
    o.set_field(w.E)
    o.set_particles(w.PartState, elem=w.elem, species=w.species)
    if (np.asarray(w.elem) == 0).any():
        o.locate()
    if staged:
        for _ in range(steps):
            s.deposition(); o.deposition()
            s.interpolate_field_to_particle(); o.interpolate()
            s.push_boris_leapfrog(w.dt); o.push(w.dt)
            s.perform_tracking(); o.track()
    else:
        s.time_step(w.dt, steps)
        o.time_step(w.dt, steps)
    s.deposition()
    o.deposition()
    return s, o


def compare(s, o, L, near_face_tol=1e-12, mesh=None):
    ps, el, ins = s.particles()
    ops, oel, oins = o.particles()
    assert np.array_equal(ins, oins), "inside flags differ: %d" % (ins != oins).sum()
    # removed particles have no owner (RemoveParticle only clears ParticleInside)
    diff = (el != oel) & ins
    if diff.any():
        # allowed only for particles within 1e-12 of a face (Cartesian boxes)
        assert mesh is not None and mesh.shape is not None
        h = (np.asarray(mesh.hi) - np.asarray(mesh.lo)) / np.asarray(mesh.shape)
        r = np.mod(ops[diff, :3] - np.asarray(mesh.lo), h) / h
        dist = np.minimum(r, 1 - r).min(1)
        assert (dist <= near_face_tol).all(), "ownership differs away from faces"
    keep = ~diff & ins
    dx = np.abs(ps[keep, :3] - ops[keep, :3]).max() if keep.any() else 0.0
    vs = max(np.abs(ops[:, 3:]).max(), 1e-300)
    dv = np.abs(ps[keep, 3:] - ops[keep, 3:]).max() / vs if keep.any() else 0.0
    PS, oPS = s.part_source(), o.part_source()
    rs = max(np.abs(oPS).max(), 1e-300)
    drho = np.abs(PS - oPS).max() / rs
    print("near-face ownership differences %d, |dx|/L %.2e, |dv| %.2e, |drho| %.2e" % (diff.sum(), dx / L, dv, drho))
    assert dx <= 1e-12 * L
    assert dv <= 1e-12
    assert drho <= 1e-12
    return diff.sum()


def test_cvwm_golden_reference_dg_source(oracle_mod):
    """GPU cell_volweight_mean on the reference's HOPR mesh and particles vs its DG_Source."""
    d = np.load(os.path.join(HERE, "golden", "cvwm_plasma_ball.npz"))
    names = [str(x) for x in d["BCNames"]]
    m = M.from_hopr_arrays(d["ElemInfo"], d["SideInfo"], d["NodeCoords"], d["GlobalNodeIDs"], d["BCType"],
                           {nm: M.PART_BC_PERIODIC for nm in names}, names, d["PeriodicVectors"])
    pd = d["PartData"]
    print("case golden plasma ball CVWM, n=%d" % len(pd), flush=True)
    with stage.PICStage(m, int(d["N"]), deposition=stage.DEPO_CVWM, charge=tuple(d["ChargeIC"]),
                        mass=tuple(d["MassIC"]), mpf=tuple(d["MacroParticleFactor"]), max_particles=len(pd)) as s:
        s.set_particles(pd[:, :6], elem=None, species=pd[:, 6].astype(np.int32))
        _, el, ins = s.particles()
        pi = d["PartInt"]
        ref = np.zeros(len(pd), np.int32)
        for e in range(pi.shape[1]):
            ref[pi[0, e]:pi[1, e]] = e + 1
        assert ins.all() and np.array_equal(el, ref)
        s.deposition()
        PS = s.part_source()
    ref = d["DG_Source"].reshape(PS.shape)
    assert np.abs(PS - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("depo", [stage.DEPO_SF, stage.DEPO_CVWM])
def test_time_step_parity_small_cube(oracle_mod, depo):
    w = W.thermal_cube(nx=6, ppc=24, deposition=depo, steps=10)
    s, o = run_pair(oracle_mod, w, 10, sf_radius=w.sf_radius if depo == stage.DEPO_SF else 0.0)
    compare(s, o, 1.0, mesh=w.mesh)
    s.close()


def test_time_step_parity_plasma_wave_box(oracle_mod):
    w = W.plasma_wave_box(ppc=16)
    s, o = run_pair(oracle_mod, w, 40)
    compare(s, o, 4 * np.pi, mesh=w.mesh)
    # CalcDepositedCharge: node-volume weighting conserves the charge to round-off
    dep, abs_err, rel_err = s.deposited_charge()
    assert abs(dep - o.deposited_charge()) <= 1e-12 * abs(dep)
    assert abs(abs_err - abs(o.particle_charge() - dep)) <= 1e-12 * abs(dep)
    assert rel_err <= 1e-12
    # CalcPotentialEnergy of the frozen field
    assert abs(s.field_energy() - o.field_energy()) <= 1e-12 * o.field_energy()
    s.close()


def test_plasma_wave_box_stated_size(oracle_mod):
    """The plain FP64 electrostatic case at its stated size: 16x4x4 linear
    hexes, N = 3, 64 electrons per cell (16,384), 200 steps, TriaTracking,
    cell_volweight_mean."""
    w = W.plasma_wave_box(ppc=64)
    assert len(w.PartState) == 16384 and w.steps == 200
    s, o = run_pair(oracle_mod, w, w.steps)
    compare(s, o, 4 * np.pi, mesh=w.mesh)
    dep, abs_err, rel_err = s.deposited_charge()
    assert abs(dep - o.deposited_charge()) <= 1e-12 * abs(dep)
    assert rel_err <= 1e-12
    s.close()


def test_staged_calls_match_oracle(oracle_mod):
    """Deposition / InterpolateFieldToParticle / push / PerformTracking one by one."""
    w = W.thermal_cube(nx=5, ppc=30, steps=1)
    kw = W.stage_kwargs(w)
    with stage.PICStage(w.mesh, w.N, max_particles=len(w.PartState), **kw) as s:
        s.set_field(w.E)
        s.set_particles(w.PartState, elem=w.elem, species=w.species)
        o = oracle_mod.Oracle(w.mesh, w.N, deposition=kw["deposition"], sf_radius=kw["sf_radius"],
                              sf_alpha=kw["sf_alpha"], external_field=kw["external_field"], charge=kw["charge"],
                              mass=kw["mass"], mpf=kw["mpf"])
        o.set_field(w.E)
        o.set_particles(w.PartState, elem=w.elem, species=w.species)
        for step in range(3):
            s.deposition(); o.deposition()
            PS, oPS = s.part_source(), o.part_source()
            assert np.abs(PS - oPS).max() <= 1e-12 * np.abs(oPS).max()
            s.interpolate_field_to_particle(); o.interpolate()
            F, oF = s.field_at_particle(), o.field_at_particle()
            assert np.abs(F - oF).max() <= 1e-12 * np.abs(oF).max()
            before, _, _ = s.particles()
            s.push_boris_leapfrog(w.dt); o.push(w.dt)
            # LastPartPos is the position before the push, exactly; the
            # positions themselves agree with the oracle's to round-off
            assert np.array_equal(s.last_part_pos(), before[:, :3])
            assert np.abs(s.last_part_pos() - o.last_part_pos()).max() <= 1e-12
            ps, _, _ = s.particles()
            ops, _, _ = o.particles()
            assert np.abs(ps[:, :3] - ops[:, :3]).max() <= 1e-12
            assert np.abs(ps[:, 3:] - ops[:, 3:]).max() <= 1e-12 * np.abs(ops[:, 3:]).max()
            s.perform_tracking(); o.track()
            ps, el, ins = s.particles()
            ops, oel, oins = o.particles()
            assert ins.all() and oins.all()
            assert np.array_equal(el, oel)
            assert np.array_equal(el, w.mesh.element_of(np.mod(ps[:, :3], 1.0)))
            assert np.abs(ps[:, :3] - ops[:, :3]).max() <= 1e-12


def test_staged_time_step_equals_fused(oracle_mod):
    w = W.thermal_cube(nx=6, ppc=20, steps=3)
    s, o = run_pair(oracle_mod, w, 3, staged=True)
    compare(s, o, 1.0, mesh=w.mesh)
    s.close()


def test_open_boundary_removal(oracle_mod):
    m = M.cartesian_box((4, 4, 4), periodic=(False, False, False))
    rng = np.random.default_rng(11)
    n = 4000
    x = rng.uniform(0.02, 0.98, (n, 3))
    v = rng.normal(size=(n, 3)) * 2e7
    w = W.Workload("open_box4", m, 2, np.hstack([x, v]), m.element_of(x), np.ones(n, np.int32),
                   np.zeros((64, 27, 3)), (0, 0, 0, 0, 0, 0), stage.DEPO_CVWM, 0.0, 2, 1e-8, 3,
                   (stage.ELEMENTARY_CHARGE * -1,), (stage.ELECTRON_MASS,), (1.0,))
    s, o = run_pair(oracle_mod, w, 3)
    ps, el, ins = s.particles()
    assert (~ins).sum() > 100
    compare(s, o, 1.0, mesh=m)
    assert s.stats()["n_lost"] == 0
    s.close()


def test_deterministic_repeat(oracle_mod):
    w = W.thermal_cube(nx=8, ppc=40, steps=4)
    out = []
    for _ in range(2):
        with stage.PICStage(w.mesh, w.N, max_particles=len(w.PartState), **W.stage_kwargs(w)) as s:
            s.set_field(w.E)
            s.set_particles(w.PartState, elem=w.elem, species=w.species)
            s.time_step(w.dt, 4)
            s.deposition()
            out.append((s.particles()[0], s.part_source()))
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])


def test_thermal_cube48_full_size_properties():
    """48^3 elements, 22.1M particles: ownership equals the closed-form owner
    after two steps, no particle lost, and the shape-function source is
    reproducible bit for bit."""
    w = W.thermal_cube()
    print("case thermal_cube48 full size n=%d" % len(w.PartState), flush=True)
    with stage.PICStage(w.mesh, w.N, max_particles=len(w.PartState), **W.stage_kwargs(w)) as s:
        s.set_field(w.E)
        s.set_particles(w.PartState, elem=w.elem, species=w.species)
        s.time_step(w.dt, 2)
        ps, el, ins = s.particles()
        assert ins.all()
        own = w.mesh.element_of(ps[:, :3])
        bad = own != el
        if bad.any():
            h = 1.0 / 48
            r = np.mod(ps[bad, :3], h) / h
            assert (np.minimum(r, 1 - r).min(1) <= 1e-12).all()
        s.deposition()
        a = s.part_source()
        s.deposition()
        b = s.part_source()
        assert np.array_equal(a, b)
        assert s.stats()["n_lost"] == 0
        # CalcDepositedCharge against the Gauss quadrature of the same source
        # (affine cube: 1/sJ = (h/2)^3 at every DOF), summed exactly
        import math
        xg, wg = np.polynomial.legendre.leggauss(w.N + 1)
        wijk = (wg[None, None, :] * wg[None, :, None] * wg[:, None, None]).reshape(-1)
        J = (0.5 / 48) ** 3
        terms = (wijk[None, :] * a[:, :, 3] * J).reshape(-1)
        dep = s.deposited_charge()[0]
        assert abs(dep - math.fsum(terms)) <= 1e-12 * np.abs(terms).sum()
        # CalcPotentialEnergy of the frozen field, the same way
        e2 = (w.E ** 2).sum(-1)
        ref = 0.5 * stage.EPS0 * math.fsum((wijk[None, :] * J * e2).reshape(-1))
        assert abs(s.field_energy() - ref) <= 1e-12 * ref


@pytest.mark.parametrize("staged", [False, True])
def test_refmapping_annulus_parity(oracle_mod, staged):
    """RefMapping on curved (NGeo=2) hexahedra: device localization, 20 steps
    of interpolation at PartPosRef, push, ParticleRefTracking with periodic z
    and open walls, and shape-function deposition, against the oracle."""
    w = W.annulus_plasma(shape=(4, 16, 4), n_total=4000, dt=2e-9)
    s, o = run_pair(oracle_mod, w, 20, staged=staged)
    ps, el, ins = s.particles()
    ops, oel, oins = o.particles()
    assert np.array_equal(ins, oins), "inside flags differ: %d" % (ins != oins).sum()
    assert np.array_equal(el[ins], oel[ins]), "owners differ: %d" % (el[ins] != oel[ins]).sum()
    assert (~ins).sum() > 0          # some particles left through the walls
    assert s.stats()["n_lost"] == 0 and o.lost() == 0
    assert np.abs(ps[ins, :3] - ops[ins, :3]).max() <= 1e-12
    vmax = np.abs(ops[ins, 3:]).max()
    assert np.abs(ps[ins, 3:] - ops[ins, 3:]).max() <= 1e-12 * vmax
    assert np.abs(s.part_pos_ref()[ins] - o.part_pos_ref()[ins]).max() <= 1e-12
    PS, oPS = s.part_source(), o.part_source()
    assert np.abs(PS - oPS).max() <= 1e-12 * np.abs(oPS).max()
    s.close()


def test_refmapping_linear_box_matches_triatracking(oracle_mod):
    """On a linear periodic box RefMapping and TriaTracking own every particle
    alike (away from faces), both against the oracle."""
    w = W.thermal_cube(nx=5, ppc=30, steps=5)
    w.tracking, w.refmapping_guess = stage.TRACK_REFMAPPING, 1
    s, o = run_pair(oracle_mod, w, 5)
    compare(s, o, 1.0, mesh=w.mesh)
    s.close()


@pytest.mark.parametrize("mesh_kind,alpha", [("cube", 2), ("deformed", 2), ("cube", 3), ("deformed", 3)])
def test_mixed_precision_error_bounds(oracle_mod, mesh_kind, alpha):
    """Mixed precision (FP32 Lagrange basis, field interpolation and shape
    weights with FP32 partial sums, FP64 positions, Newton and face tests)
    against the FP64 oracle: deposited source within 1e-6 relative (L2 and
    max), positions within 1e-9 of the domain after 10 steps.  The Cartesian
    cube runs the tensor-grid gather, the deformed box the general one."""
    w = W.thermal_cube(nx=8, ppc=40, steps=10)
    w.sf_alpha = alpha
    if mesh_kind == "deformed":
        w = deformed_box(w)
    print("case mixed precision on %s alpha=%d" % (mesh_kind, alpha), flush=True)
    kw = W.stage_kwargs(w)
    with stage.PICStage(w.mesh, w.N, max_particles=len(w.PartState), precision=stage.PREC_MIXED, **kw) as s:
        s.set_field(w.E)
        s.set_particles(w.PartState, elem=w.elem, species=w.species)
        s.deposition()
        assert s.stats()["deposit_kernel"] == stage.KERNEL_SF_MIXED
        PS0 = s.part_source()
        s.time_step(w.dt, 10)
        s.deposition()
        ps, el, ins = s.particles()
        PS = s.part_source()
    o = oracle_mod.Oracle(w.mesh, w.N, deposition=kw["deposition"], sf_radius=kw["sf_radius"], sf_alpha=kw["sf_alpha"],
                          external_field=kw["external_field"], charge=kw["charge"], mass=kw["mass"], mpf=kw["mpf"])
    o.set_field(w.E)
    o.set_particles(w.PartState, elem=w.elem, species=w.species)
    if mesh_kind == "deformed":
        o.locate()
    o.deposition()
    oPS0 = o.part_source()
    o.time_step(w.dt, 10)
    o.deposition()
    ops, oel, oins = o.particles()
    oPS = o.part_source()
    for a, b in ((PS0, oPS0), (PS, oPS)):
        l2 = np.linalg.norm(a - b) / np.linalg.norm(b)
        linf = np.abs(a - b).max() / np.abs(b).max()
        print("mixed vs FP64: rho L2 %.2e  Linf %.2e" % (l2, linf))
        assert l2 <= 1e-6 and linf <= 1e-6
    assert ins.all() and oins.all()
    dx = np.abs(ps[:, :3] - ops[:, :3]).max()
    print("mixed vs FP64: max |dx| %.2e" % dx)
    assert dx <= 1e-9
    # owners agree except for particles within 1e-9 of a face
    diff = el != oel
    if mesh_kind == "deformed":
        assert diff.sum() <= 2
    elif diff.any():
        h = 1.0 / 8
        r = np.mod(ops[diff, :3], h) / h
        assert (np.minimum(r, 1 - r).min(1) <= 1e-9).all()


@pytest.mark.parametrize("case", ["cube_sf", "open_box", "annulus"])
def test_graph_replay_matches_eager(case):
    """pic_time_step replaying the captured two-step graph gives, bit for bit,
    what the eager step gives; each replay is checked, with eager work (a
    deposition and read-backs) on the same stream between replays, and with
    particles leaving through open walls so the count shrinks under the graph."""
    if case == "cube_sf":
        w = W.thermal_cube(nx=6, ppc=20, steps=2)
    elif case == "open_box":
        w = W.thermal_cube(nx=6, ppc=20, steps=2)
        w.mesh = M.cartesian_box((6, 6, 6), periodic=(False, True, True))
        w.dt *= 40.0
    else:
        w = W.annulus_plasma(shape=(4, 16, 4), n_total=4000, dt=2e-9)
    kw = W.stage_kwargs(w)
    if case == "open_box":
        kw["sf_radius"] = w.sf_radius
    print("case graph replay %s n=%d" % (case, len(w.PartState)), flush=True)
    sg = stage.PICStage(w.mesh, w.N, max_particles=len(w.PartState), **kw)
    se = stage.PICStage(w.mesh, w.N, max_particles=len(w.PartState), **kw)
    sg.set_tuning(use_graph=True)
    for s in (sg, se):
        s.set_field(w.E)
        s.set_particles(w.PartState, elem=w.elem, species=w.species)
    counts = []
    for rep in range(4):
        sg.time_step(w.dt, 2)
        se.time_step(w.dt, 2)
        sg.deposition()
        se.deposition()
        a, b = sg.particles(), se.particles()
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2]), rep
        assert np.array_equal(sg.part_source(), se.part_source()), rep
        counts.append(sg.stats()["n_particles"])
    if case == "open_box":
        assert counts[-1] < len(w.PartState)
    sg.close()
    se.close()


def test_tuning_change_rerecords_graph(oracle_mod):
    """A graph recorded with one shape-function kernel is not replayed after
    pic_set_tuning chose another: graph on with the matrix-core gather, then
    the VALU gather for two single eager steps (the sort then writes no
    records), then two replayed steps.  The source each step leaves behind
    matches an eager stage that runs the same kernel sequence, and the oracle."""
    w = W.thermal_cube(nx=6, ppc=20, steps=6)
    kw = W.stage_kwargs(w)
    print("case tuning switch under graph replay n=%d" % len(w.PartState), flush=True)
    sg = stage.PICStage(w.mesh, w.N, max_particles=len(w.PartState), **kw)
    se = stage.PICStage(w.mesh, w.N, max_particles=len(w.PartState), **kw)
    o = oracle_mod.Oracle(w.mesh, w.N, deposition=kw["deposition"], sf_radius=kw["sf_radius"],
                          sf_alpha=kw["sf_alpha"], external_field=kw["external_field"], charge=kw["charge"],
                          mass=kw["mass"], mpf=kw["mpf"])
    for s in (sg, se, o):
        s.set_field(w.E)
        s.set_particles(w.PartState, elem=w.elem, species=w.species)
    sg.set_tuning(use_graph=True, sf_kernel=2)
    se.set_tuning(use_graph=False, sf_kernel=2)
    plan = [(2, 2), (1, 1), (1, 1), (1, 2), (2, 2)]   # (sf_kernel, steps)
    for kern, steps in plan:
        sg.set_tuning(use_graph=True, sf_kernel=kern)
        se.set_tuning(use_graph=False, sf_kernel=kern)
        sg.time_step(w.dt, steps)
        se.time_step(w.dt, steps)
        o.time_step(w.dt, steps)
        a, b = sg.particles(), se.particles()
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
        PSg, PSe, oPS = sg.part_source(), se.part_source(), o.part_source()
        rs = np.abs(oPS).max()
        assert np.abs(PSg - PSe).max() <= 1e-12 * rs, (kern, steps)
        assert np.abs(PSg - oPS).max() <= 1e-12 * rs, (kern, steps)
    sg.close()
    se.close()


@pytest.mark.parametrize("kern", [1, 2])
@pytest.mark.parametrize("alpha", [2, 3])
def test_sf_alpha_and_kernel_parity(oracle_mod, alpha, kern):
    """Shape-function exponent 2 (squared-weight kernels) and 3 (the
    run-time exponent kernels) on the VALU gather and the matrix-core
    tensor-grid gather, 5 steps against the oracle."""
    w = W.thermal_cube(nx=6, ppc=24, steps=5)
    w.sf_alpha = alpha
    kw = W.stage_kwargs(w)
    print("case sf alpha=%d kernel=%d" % (alpha, kern), flush=True)
    s = stage.PICStage(w.mesh, w.N, max_particles=len(w.PartState), **kw)
    s.set_tuning(sf_kernel=kern)
    s.set_field(w.E)
    s.set_particles(w.PartState, elem=w.elem, species=w.species)
    o = oracle_mod.Oracle(w.mesh, w.N, deposition=kw["deposition"], sf_radius=kw["sf_radius"],
                          sf_alpha=kw["sf_alpha"], external_field=kw["external_field"], charge=kw["charge"],
                          mass=kw["mass"], mpf=kw["mpf"])
    o.set_field(w.E)
    o.set_particles(w.PartState, elem=w.elem, species=w.species)
    s.time_step(w.dt, 5)
    o.time_step(w.dt, 5)
    s.deposition()
    o.deposition()
    want = stage.KERNEL_SF_VALU if kern == 1 else stage.KERNEL_SF_CARTESIAN
    assert s.stats()["deposit_kernel"] == want
    compare(s, o, 1.0, mesh=w.mesh)
    s.close()


@pytest.mark.parametrize("env", ["PIC_SF_CARTESIAN", "PIC_AFFINE"])
def test_general_kernels_on_cartesian_mesh(oracle_mod, env, monkeypatch):
    """The general shape-function gather and the Newton interpolation, forced
    on a Cartesian mesh (where the tensor-grid gather and the affine
    reference coordinates would run), against the oracle."""
    monkeypatch.setenv(env, "0")
    w = W.thermal_cube(nx=6, ppc=24, steps=10)
    s, o = run_pair(oracle_mod, w, 10)
    if env == "PIC_SF_CARTESIAN":
        assert s.stats()["deposit_kernel"] == stage.KERNEL_SF_MATRIX
    compare(s, o, 1.0, mesh=w.mesh)
    s.close()


def deformed_box(w, amp=0.12):
    """Displace the nodes of a generated periodic unit box by a smooth
    periodic field that keeps the boundary planes (faces become bilinear, no
    element is affine, no Gauss-point grid is a tensor grid)."""
    X = w.mesh.NodeCoords
    x, y, z = X[:, 0].copy(), X[:, 1].copy(), X[:, 2].copy()
    a = amp / w.mesh.shape[0]
    s, c = (lambda u: np.sin(2 * np.pi * u)), (lambda u: np.cos(2 * np.pi * u))
    X[:, 0] = x + a * s(x) * c(y) * c(z)
    X[:, 1] = y + a * c(x) * s(y) * c(z)
    X[:, 2] = z + a * c(x) * c(y) * s(z)
    w.elem = np.zeros(len(w.PartState), np.int32)   # located by the stage and the oracle
    return w


@pytest.mark.parametrize("depo,alpha", [(stage.DEPO_SF, 2), (stage.DEPO_SF, 3), (stage.DEPO_CVWM, 2)])
def test_deformed_box_parity(oracle_mod, depo, alpha):
    """Linear hexahedra with bilinear faces (ConcaveElemSide, Newton on the
    trilinear map, the general gather): device localization and 10 steps
    against the oracle, ownership bit-exact."""
    w = deformed_box(W.thermal_cube(nx=5, ppc=30, deposition=depo, steps=10))
    w.sf_alpha = alpha
    s, o = run_pair(oracle_mod, w, 10, sf_radius=w.sf_radius if depo == stage.DEPO_SF else 0.0)
    if depo == stage.DEPO_SF:
        assert s.stats()["deposit_kernel"] == stage.KERNEL_SF_MATRIX
    assert compare(s, o, 1.0) == 0
    s.close()


def test_tiebreak_deformed_box_parity(oracle_mod):
    """TriaTracking's multi-crossing branch (several sides register a
    crossing; the side crossed first is kept) on a strongly deformed periodic
    box with moves of about 0.75 element widths per step and no field: 3
    steps against the oracle, ownership bit-exact.  The oracle's own check
    that the kept side is the first crossed is in the closed-form tests."""
    w = deformed_box(W.thermal_cube(nx=5, ppc=40, deposition=stage.DEPO_CVWM, steps=3), amp=0.3)
    rng = np.random.default_rng(3)
    w.PartState[:, 3:] = rng.normal(size=(len(w.PartState), 3)) * 1.5e5
    w.E[:] = 0.0
    w.dt = 1e-6
    s, o = run_pair(oracle_mod, w, 3, sf_radius=0.0, external_field=np.zeros(6))
    print("multi-crossing walk steps in the oracle: %d" % o.multi_crossings())
    assert o.multi_crossings() >= 20
    assert compare(s, o, 1.0) == 0
    s.close()


def test_device_resident_field_and_source():
    """pic_set_field_device / pic_part_source_device: the field handed over
    from device memory gives the same steps as the host hand-over
    (bit-identical), and the source read at the device address equals the
    host copy."""
    w = W.thermal_cube(nx=5, ppc=20, steps=3)
    kw = W.stage_kwargs(w)
    E_dev = stage.DeviceBuffer.from_array(w.E)
    res = []
    for device_path in (False, True):
        s = stage.PICStage(w.mesh, w.N, max_particles=len(w.PartState), **kw)
        if device_path:
            s.set_field_device(E_dev.ptr)
        else:
            s.set_field(w.E)
        s.set_particles(w.PartState, elem=w.elem, species=w.species)
        s.time_step(w.dt, 3)
        s.deposition()
        ps, el, ins = s.particles()
        src = s.part_source()
        if device_path:
            dev = stage.device_to_host(s.part_source_device(), src.shape)
            assert np.array_equal(dev, src)
        res.append((ps, el, src))
        s.close()
    E_dev.free()
    assert np.array_equal(res[0][0], res[1][0]) and np.array_equal(res[0][1], res[1][1])
    assert np.array_equal(res[0][2], res[1][2])
