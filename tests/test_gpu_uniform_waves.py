"""Dense elements: most waves hold particles of a single element, so the push
kernels take their element-uniform path (element tables by scalar loads).
The same bars as the other parity tests, against the oracle:

  * TriaTracking on affine boxes (both depositions) and on deformed boxes
    (Newton on the trilinear map);
  * RefMapping on the curved annulus (start guess 2, Newton on the 27-point map);
  * mixed precision against the FP64 oracle (1e-6 on the source, 1e-9 of the
    domain on positions).
"""
import numpy as np
import pytest

from piclas_amd import stage, workloads as W
from test_gpu_parity import compare, deformed_box, run_pair

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("depo", [stage.DEPO_SF, stage.DEPO_CVWM])
def test_dense_affine_box(oracle_mod, depo):
    w = W.thermal_cube(nx=3, ppc=640, deposition=depo, steps=5)
    s, o = run_pair(oracle_mod, w, 5, sf_radius=w.sf_radius if depo == stage.DEPO_SF else 0.0)
    compare(s, o, 1.0, mesh=w.mesh)
    s.close()


def test_dense_deformed_box(oracle_mod):
    w = deformed_box(W.thermal_cube(nx=3, ppc=640, steps=5))
    s, o = run_pair(oracle_mod, w, 5, sf_radius=w.sf_radius)
    assert compare(s, o, 1.0) == 0
    s.close()


def test_dense_refmapping_annulus(oracle_mod):
    w = W.annulus_plasma(shape=(2, 8, 2), n_total=12000, dt=2e-9)
    s, o = run_pair(oracle_mod, w, 5)
    ps, el, ins = s.particles()
    ops, oel, oins = o.particles()
    assert np.array_equal(ins, oins)
    assert np.array_equal(el[ins], oel[ins])
    assert np.abs(ps[ins, :3] - ops[ins, :3]).max() <= 1e-12
    vmax = np.abs(ops[ins, 3:]).max()
    assert np.abs(ps[ins, 3:] - ops[ins, 3:]).max() <= 1e-12 * vmax
    assert np.abs(s.part_pos_ref()[ins] - o.part_pos_ref()[ins]).max() <= 1e-12
    PS, oPS = s.part_source(), o.part_source()
    assert np.abs(PS - oPS).max() <= 1e-12 * np.abs(oPS).max()
    s.close()


def test_dense_mixed_precision(oracle_mod):
    w = W.thermal_cube(nx=3, ppc=640, steps=5)
    kw = W.stage_kwargs(w)
    print("case dense mixed n=%d" % len(w.PartState), flush=True)
    with stage.PICStage(w.mesh, w.N, max_particles=len(w.PartState), precision=stage.PREC_MIXED, **kw) as s:
        s.set_field(w.E)
        s.set_particles(w.PartState, elem=w.elem, species=w.species)
        s.time_step(w.dt, 5)
        s.deposition()
        ps, el, ins = s.particles()
        PS = s.part_source()
    o = oracle_mod.Oracle(w.mesh, w.N, deposition=kw["deposition"], sf_radius=kw["sf_radius"], sf_alpha=kw["sf_alpha"],
                          external_field=kw["external_field"], charge=kw["charge"], mass=kw["mass"], mpf=kw["mpf"])
    o.set_field(w.E)
    o.set_particles(w.PartState, elem=w.elem, species=w.species)
    o.time_step(w.dt, 5)
    o.deposition()
    ops, oel, oins = o.particles()
    oPS = o.part_source()
    assert ins.all() and oins.all()
    dx = np.abs(ps[:, :3] - ops[:, :3]).max()
    dv = np.abs(ps[:, 3:] - ops[:, 3:]).max() / np.abs(ops[:, 3:]).max()
    l2 = np.linalg.norm(PS - oPS) / np.linalg.norm(oPS)
    print("dense mixed vs FP64: |dx| %.2e |dv| %.2e rho L2 %.2e" % (dx, dv, l2))
    assert dx <= 1e-9 and dv <= 1e-6 and l2 <= 1e-6
