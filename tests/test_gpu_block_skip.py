"""The general-element shape-function gather skips the blocks of 64 records
whose bounding box cannot reach its target; the gather whose workgroups
share the record loads of four targets reads every record.  Both stage the
records they keep in record order, so they agree bit for bit exactly when no
skipped block held a record the cull keeps.

Cases chosen so the boxes see what the culls see: a neutral species (records
with f = 0, left out of the boxes and rejected by every cull) on every third
particle, a band of elements emptied of particles (empty runs, blocks that
span several elements), and a particle count that leaves the last block
partly filled.  Meshes: the 6^3 periodic box with displaced nodes (linear
elements, TriaTracking, 20 per cell before the cut) and the 4 x 16 x 4
NGeo = 2 annulus with 6000 particles (RefMapping, barycentre cull), 4 steps,
a deposition after each pair."""
import numpy as np
import pytest

from piclas_amd import stage
from piclas_amd import workloads as W
from test_gpu_parity import deformed_box

pytestmark = pytest.mark.gpu


def _neutral_and_gap(w, gap):
    """Every third particle becomes a neutral species; particles in `gap` are dropped."""
    keep = ~gap(w.PartState[:, :3])
    w.PartState = np.ascontiguousarray(w.PartState[keep])
    w.elem = np.ascontiguousarray(w.elem[keep])
    sp = np.ascontiguousarray(w.species[keep]).copy()
    neutral = len(w.charge) + 1
    sp[::3] = neutral
    w.species = sp
    w.charge = tuple(w.charge) + (0.0,)
    w.mass = tuple(w.mass) + (w.mass[-1],)
    w.mpf = tuple(w.mpf) + (w.mpf[-1],)
    return w


def _workload(case):
    if case == "deformed":
        w = deformed_box(W.thermal_cube(nx=6, ppc=20, steps=4))
        return _neutral_and_gap(w, lambda x: (x[:, 0] > 0.12) & (x[:, 0] < 0.55) & (x[:, 1] < 0.5))
    w = W.annulus_plasma(shape=(4, 16, 4), n_total=6000, dt=2e-9)
    th = lambda x: np.mod(np.arctan2(x[:, 1], x[:, 0]), 2 * np.pi)
    return _neutral_and_gap(w, lambda x: (th(x) > 1.0) & (th(x) < 1.9))


@pytest.mark.parametrize("case", ["deformed", "annulus"])
def test_block_skip_matches_full_read(case):
    w = _workload(case)
    n = len(w.PartState)
    assert n % 64 != 0
    kw = W.stage_kwargs(w)
    print("case block skip %s n=%d species=%d" % (case, n, len(w.charge)), flush=True)
    sd = stage.PICStage(w.mesh, w.N, max_particles=n, **kw)
    sc = stage.PICStage(w.mesh, w.N, max_particles=n, **kw)
    sd.set_tuning(sf_kernel=2)
    sc.set_tuning(sf_kernel=2, sf_shared_loads=True)
    for s in (sd, sc):
        s.set_field(w.E)
        s.set_particles(w.PartState, elem=w.elem, species=w.species)
    for rep in range(2):
        sd.time_step(w.dt, 2)
        sc.time_step(w.dt, 2)
        sd.deposition()
        sc.deposition()
        a, b = sd.particles(), sc.particles()
        assert all(np.array_equal(u, v) for u, v in zip(a, b)), rep
        ps = sd.part_source()
        assert np.array_equal(ps, sc.part_source()), rep
        assert np.abs(ps).max() > 0.0
    assert sd.stats()["deposit_kernel"] == stage.KERNEL_SF_MATRIX
    sd.close()
    sc.close()
