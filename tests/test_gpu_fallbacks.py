"""The Newton-failure fallbacks on the device, forced from the host.

Particles are handed over with an element far from the one that holds them
(pic_set_particles trusts a given element id), so RefElemNewton gives up
for them as it does in the reference for a particle far outside its
element.  With cell_volweight_mean deposition the stage then takes:
  * GetEMFieldDW for the field (pic_interpolation_tools.f90:248-249,458-572);
  * the inverse-distance corner weights of the deposition
    (pic_depo_method.f90:512-539).
Both are compared with the oracle at the usual bars, on the affine cube
(where the device forms xi directly and must report the failure itself) and
on a deformed box (Newton on the trilinear map)."""
import numpy as np
import pytest

from piclas_amd import stage, workloads as W

from test_gpu_parity import deformed_box

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mesh_kind", ["cube", "deformed"])
def test_newton_failure_fallbacks_match_oracle(oracle_mod, mesh_kind):
    w = W.thermal_cube(nx=6, ppc=20, deposition=stage.DEPO_CVWM, steps=1)
    if mesh_kind == "deformed":
        w = deformed_box(w)
        w.elem = None
    nx = 6
    rng = np.random.default_rng(31)
    n = len(w.PartState)
    kw = W.stage_kwargs(w)
    kw["sf_radius"] = 0.0
    o = oracle_mod.Oracle(w.mesh, w.N, deposition=kw["deposition"], external_field=kw["external_field"],
                          charge=kw["charge"], mass=kw["mass"], mpf=kw["mpf"])
    o.set_field(w.E)
    # owners: closed form on the cube, located by the oracle on the deformed box
    if mesh_kind == "cube":
        own = w.mesh.element_of(w.PartState[:, :3])
    else:
        o.set_particles(w.PartState, elem=np.zeros(n, np.int32))
        o.locate()
        own = o.particles()[1]
        assert (own > 0).all()
    # a third of the particles get an element three cells away in x and y
    far = rng.uniform(size=n) < 1 / 3
    e0 = own - 1
    i, j, k = e0 % nx, (e0 // nx) % nx, e0 // (nx * nx)
    moved = ((i + 3) % nx + nx * ((j + 3) % nx + nx * k) + 1).astype(np.int32)
    elem = np.where(far, moved, own).astype(np.int32)
    _, ok = o.position_in_ref_elem(w.PartState[:, :3], elem)
    assert not ok[far].any() and ok[~far].all()
    print("case Newton failure %s: %d of %d particles far from their element" % (mesh_kind, far.sum(), n),
          flush=True)
    o.set_particles(w.PartState, elem=elem, species=w.species)
    with stage.PICStage(w.mesh, w.N, max_particles=n, **kw) as s:
        s.set_field(w.E)
        s.set_particles(w.PartState, elem=elem, species=w.species)
        s.deposition()
        o.deposition()
        PS, oPS = s.part_source(), o.part_source()
        assert np.abs(PS - oPS).max() <= 1e-12 * np.abs(oPS).max()
        dep = s.deposited_charge()[0]
        assert abs(dep - o.deposited_charge()) <= 1e-12 * abs(dep)
        s.interpolate_field_to_particle()
        o.interpolate()
        F, oF = s.field_at_particle(), o.field_at_particle()
        assert np.abs(F - oF).max() <= 1e-12 * np.abs(oF).max()
        # the fallback field is a weighted mean of the element's nodal field
        # (never an extrapolation): bounded by the element's extreme values
        Ee = w.E[elem[far] - 1]
        ext = np.asarray(kw["external_field"][:3])
        Ff = F[far, :3] - ext
        assert (Ff <= Ee.max(1) + 1e-12).all() and (Ff >= Ee.min(1) - 1e-12).all()
        s.push_boris_leapfrog(w.dt)
        o.push(w.dt)
        ps, _, _ = s.particles()
        ops, _, _ = o.particles()
        assert np.abs(ps[:, :3] - ops[:, :3]).max() <= 1e-12
        assert np.abs(ps[:, 3:] - ops[:, 3:]).max() <= 1e-12 * np.abs(ops[:, 3:]).max()
