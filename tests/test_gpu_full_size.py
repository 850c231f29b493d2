"""The scaled FP64 throughput case at its full stated size against the
oracle: periodic unit cube of 48^3 linear hexes, N = 3, 200 electrons per
element (22,118,400), Maxwellian 2 eV, frozen smooth E and B = 0.01 T,
shape_function R = 1.5h, alpha = 2, TriaTracking.  One time step, then the
deposition of the moved particles, on the device and in the oracle (OpenMP on
the host's thread share); the bars are those of the other parity tests
(ownership bit-exact except within 1e-12 of a face, positions and velocities
to 1e-12, the deposited source to 1e-12 of its maximum)."""
import os

import pytest

from piclas_amd import stage, workloads as W
from test_gpu_parity import compare

pytestmark = pytest.mark.gpu


def host_threads():
    cores = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS", "")
    return min(cores, int(omp)) if omp.isdigit() and int(omp) > 0 else cores


def test_throughput_cube_full_size_step(oracle_mod):
    w = W.thermal_cube(nx=48, ppc=200, steps=1)
    assert len(w.PartState) == 22_118_400
    kw = W.stage_kwargs(w)
    print("case %s one step, n=%d, oracle threads %d" % (w.name, len(w.PartState), host_threads()), flush=True)
    s = stage.PICStage(w.mesh, w.N, max_particles=len(w.PartState), **kw)
    s.set_field(w.E)
    s.set_particles(w.PartState, elem=w.elem, species=w.species)
    s.time_step(w.dt, 1)
    s.deposition()
    o = oracle_mod.Oracle(w.mesh, w.N, deposition=kw["deposition"], sf_radius=kw["sf_radius"], sf_alpha=kw["sf_alpha"],
                          external_field=kw["external_field"], charge=kw["charge"], mass=kw["mass"], mpf=kw["mpf"])
    o.set_threads(host_threads())
    o.set_field(w.E)
    o.set_particles(w.PartState, elem=w.elem, species=w.species)
    o.time_step(w.dt, 1)
    o.deposition()
    compare(s, o, 1.0, mesh=w.mesh)
    s.close()


@pytest.mark.parametrize("dt_scale,steps", [(1.0, 5), (10.0, 3)])
def test_annulus_full_mesh_parity(oracle_mod, dt_scale, steps):
    """The uneven-occupancy case on its full stated mesh: the 24 x 96 x 24
    curved NGeo = 2 annulus, RefMapping, open radial walls (ParticleBCTracking
    through the Bezier patches), periodic z sides, shape_function R = 0.03 m;
    300,000 of the workload's particles (the oracle finishes them in seconds),
    at the stated dt and at ten times it (many wall and z-side crossings).
    Bars: inside flags and owners bit-exact, positions and PartPosRef to
    1e-12, the deposited source to 1e-12 of its maximum."""
    import numpy as np
    w = W.annulus_plasma(n_total=300_000)
    assert w.mesh.nElems == 24 * 96 * 24
    dt = w.dt * dt_scale
    kw = W.stage_kwargs(w)
    print("case %s full mesh, n=%d, dt=%g, %d steps, oracle threads %d" % (w.name, len(w.PartState), dt, steps,
                                                                           host_threads()), flush=True)
    s = stage.PICStage(w.mesh, w.N, max_particles=len(w.PartState), **kw)
    s.set_field(w.E)
    s.set_particles(w.PartState, elem=None, species=w.species)
    o = oracle_mod.Oracle(w.mesh, w.N, deposition=kw["deposition"], sf_radius=kw["sf_radius"], sf_alpha=kw["sf_alpha"],
                          external_field=kw["external_field"], charge=kw["charge"], mass=kw["mass"], mpf=kw["mpf"],
                          tracking=kw["tracking"], refmapping_guess=kw["refmapping_guess"])
    o.set_threads(host_threads())
    o.set_field(w.E)
    o.set_particles(w.PartState, elem=np.zeros(len(w.PartState), np.int32), species=w.species)
    assert o.locate() == 0
    s.time_step(dt, steps)
    o.time_step(dt, steps)
    ps, el, ins = s.particles()
    ops, oel, oins = o.particles()
    assert np.array_equal(ins, oins), "inside flags differ: %d" % (ins != oins).sum()
    assert np.array_equal(el[ins], oel[ins]), "owners differ: %d" % (el[ins] != oel[ins]).sum()
    assert s.stats()["n_lost"] == 0 and o.lost() == 0
    assert np.abs(ps[ins, :3] - ops[ins, :3]).max() <= 1e-12
    assert np.abs(s.part_pos_ref()[ins] - o.part_pos_ref()[ins]).max() <= 1e-12
    removed = int((~oins).sum())
    print("removed through the walls: %d" % removed, flush=True)
    if dt_scale > 1:
        assert removed > 100
    s.deposition()
    o.deposition()
    PS, oPS = s.part_source(), o.part_source()
    assert np.abs(PS - oPS).max() <= 1e-12 * np.abs(oPS).max()
    s.close()
