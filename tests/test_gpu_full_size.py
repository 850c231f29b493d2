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
