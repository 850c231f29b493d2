"""Execution tuning leaves the results alone: the persistent matrix-core
shape-function waves (targets taken from per-XCD counters), the gather whose
workgroups share the record loads of four targets (on the deformed box and
the curved annulus; the cube's tensor-grid gather ignores it) and graph replay give, bit for bit, the
particles and the source of the default stage, whose results the parity
tests check against the oracle; and the autotuner's choice is one of its
candidates, applied to the same bits.

Cases: the 6^3 periodic cube at 20 particles per cell (tensor-grid gather),
the same box with its nodes displaced (general gather on linear elements;
216 targets in 54 workgroups on a grid of 56, so two lie past the mesh), the
4 x 16 x 4 NGeo = 2 annulus with 6000 particles (general matrix-core
gather, boundary tracking), 4 steps, a deposition after each pair."""
import numpy as np
import pytest

from piclas_amd import autotune, stage
from piclas_amd import workloads as W
from test_gpu_parity import deformed_box

pytestmark = pytest.mark.gpu


def _workload(case):
    if case == "cube":
        return W.thermal_cube(nx=6, ppc=20, steps=4)
    if case == "deformed":
        return deformed_box(W.thermal_cube(nx=6, ppc=20, steps=4))
    return W.annulus_plasma(shape=(4, 16, 4), n_total=6000, dt=2e-9)


@pytest.mark.parametrize("knob", ["sf_persistent", "sf_shared_loads"])
@pytest.mark.parametrize("case", ["cube", "deformed", "annulus"])
@pytest.mark.parametrize("graph", [False, True])
def test_gather_knobs_bit_identical(case, graph, knob):
    w = _workload(case)
    kw = W.stage_kwargs(w)
    print("case gather knob %s %s graph=%s n=%d" % (knob, case, graph, len(w.PartState)), flush=True)
    sd = stage.PICStage(w.mesh, w.N, max_particles=len(w.PartState), **kw)
    sp = stage.PICStage(w.mesh, w.N, max_particles=len(w.PartState), **kw)
    sp.set_tuning(use_graph=graph, sf_kernel=2, **{knob: True})
    for s in (sd, sp):
        s.set_field(w.E)
        s.set_particles(w.PartState, elem=w.elem, species=w.species)
    for rep in range(2):
        sd.time_step(w.dt, 2)
        sp.time_step(w.dt, 2)
        sd.deposition()
        sp.deposition()
        a, b = sd.particles(), sp.particles()
        assert all(np.array_equal(u, v) for u, v in zip(a, b)), rep
        assert np.array_equal(sd.part_source(), sp.part_source()), rep
    assert sp.stats()["deposit_kernel"] == sd.stats()["deposit_kernel"]
    sd.close()
    sp.close()


def test_autotune_choice_applies_bit_identically():
    w = W.annulus_plasma(shape=(4, 16, 4), n_total=6000, dt=2e-9)
    print("case autotune annulus n=%d" % len(w.PartState), flush=True)
    best, table = autotune.autotune(w, steps=2, warmup=1)
    assert best in [c for c, _ in table] and len(table) == len(autotune.candidates(w.N, w.deposition))
    kw = W.stage_kwargs(w)
    sd = stage.PICStage(w.mesh, w.N, max_particles=len(w.PartState), **kw)
    st = stage.PICStage(w.mesh, w.N, max_particles=len(w.PartState), **kw)
    st.set_tuning(**best)
    sd.set_tuning(sf_kernel=best["sf_kernel"])   # the VALU and matrix sums differ in order only
    for s in (sd, st):
        s.set_field(w.E)
        s.set_particles(w.PartState, elem=w.elem, species=w.species)
        s.time_step(w.dt, 3)
        s.deposition()
    a, b = sd.particles(), st.particles()
    assert all(np.array_equal(u, v) for u, v in zip(a, b))
    assert np.array_equal(sd.part_source(), st.part_source())
    sd.close()
    st.close()
