"""Edge cases of the particle stage on the GPU, through the C ABI.

  * no particles at all: every call succeeds and the source is zero;
  * a neutral species beside electrons (isDepositParticle: no charge, no
    field force) and all particles in one element of the mesh (every other
    target element of the shape-function gather empty), against the oracle;
  * every particle leaving through open walls: the following steps run with
    an empty stage.
"""
import numpy as np
import pytest

from piclas_amd import mesh as M
from piclas_amd import stage, workloads as W
from test_gpu_parity import compare, run_pair

pytestmark = pytest.mark.gpu

AR_MASS = 6.63e-26


def test_zero_particles():
    w = W.thermal_cube(nx=4, ppc=1, steps=1)
    kw = W.stage_kwargs(w)
    print("case zero particles", flush=True)
    with stage.PICStage(w.mesh, w.N, max_particles=16, **kw) as s:
        s.set_field(w.E)
        s.set_particles(np.zeros((0, 6)), elem=np.zeros(0, np.int32), species=np.zeros(0, np.int32))
        s.time_step(w.dt, 3)
        s.deposition()
        assert not s.part_source().any()
        assert not any(s.deposited_charge())
        ps, el, ins = s.particles()
        assert ps.shape == (0, 6) and el.shape == (0,)
        pd, pi = s.particle_data()
        assert pd.shape == (0, 7) and not pi.any()
        assert s.restart_particles(np.zeros((0, 7)), np.zeros((s.nElems, 2), np.int64)) == (0, 0)
        s.time_step(w.dt, 2)


@pytest.mark.parametrize("depo", [stage.DEPO_SF, stage.DEPO_CVWM])
def test_neutral_species_in_one_element(oracle_mod, depo):
    w = W.thermal_cube(nx=4, ppc=1, steps=5, deposition=depo)
    rng = np.random.Generator(np.random.PCG64(7))
    n = 400
    h = 0.25
    pos = h * (1.0 + rng.uniform(0.05, 0.95, size=(n, 3)))          # element (1,1,1)
    vel = np.vstack([W.maxwellian(rng, n // 2, 0.05, W.ELECTRON_MASS), W.maxwellian(rng, n - n // 2, 0.05, AR_MASS)])
    w.PartState = np.hstack([pos, vel])
    w.elem = w.mesh.element_of(pos)
    assert (w.elem == w.elem[0]).all()
    w.species = np.r_[np.ones(n // 2, np.int32), np.full(n - n // 2, 2, np.int32)]
    q, m, mpf = w.charge[0], w.mass[0], w.mpf[0]
    s, o = run_pair(oracle_mod, w, 5, deposition=depo, charge=(q, 0.0), mass=(m, AR_MASS), mpf=(mpf, mpf),
                    sf_radius=w.sf_radius if depo == stage.DEPO_SF else 0.0)
    compare(s, o, 1.0, mesh=w.mesh)
    ps, _, _ = s.particles()
    # neutral particles keep their velocity exactly (no charge, no force)
    neutral = w.species == 2
    assert np.array_equal(ps[neutral, 3:], w.PartState[neutral, 3:])
    s.close()


def test_all_particles_leave_through_open_walls():
    mesh = M.cartesian_box((3, 3, 3), lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), periodic=(False, False, False))
    rng = np.random.Generator(np.random.PCG64(11))
    n = 300
    pos = rng.uniform(0.2, 0.8, size=(n, 3))
    vel = np.tile([3.0, 0.0, 0.0], (n, 1))                           # all leave through x = 1 in the first step
    E = np.zeros((mesh.nElems, 64, 3))
    print("case all particles leave, n=%d" % n, flush=True)
    with stage.PICStage(mesh, 3, max_particles=n, deposition=stage.DEPO_CVWM, charge=(-1.6e-19,),
                        mass=(9.1e-31,), mpf=(1.0,)) as s:
        s.set_field(E)
        s.set_particles(np.hstack([pos, vel]), elem=None, species=np.ones(n, np.int32))
        s.time_step(0.5, 2)
        ps, el, ins = s.particles()
        assert not ins.any()
        s.time_step(0.5, 3)
        s.deposition()
        assert not s.part_source().any()
        pd, pi = s.particle_data()
        assert pd.shape[0] == 0 and not pi.any()
