"""ParticleBCTracking on the device against the oracle (RefMapping on the
curved annulus: open inner and outer walls, periodic z sides).

  * hand-placed moves: out through the outer wall, out through the inner
    wall, a chord that leaves through the inner wall and comes back into the
    annulus within the step (removed: its first crossing is an open wall),
    through the upper and the lower z side (wrapped by PeriodicBoundary), and
    moves along the walls that stay inside;
  * the annulus plasma with steps of about half an element, so that many
    particles of the boundary elements meet a wall or a z side: owners,
    inside flags, positions and PartPosRef after 10 steps, fused and staged.

Bars: ownership and inside flags bit-exact, positions 1e-12 of the domain,
PartPosRef 1e-12."""
import numpy as np
import pytest

from piclas_amd import mesh as M
from piclas_amd import stage, workloads as W
from test_gpu_parity import run_pair

pytestmark = pytest.mark.gpu


def test_bc_tracking_hand_placed_moves(oracle_mod):
    m = M.annulus((3, 12, 3))
    ps = np.array([[0.97, 0.01, 0.5, 1.0, 0.0, 0.0],      # out through r = 1
                   [0.52, 0.02, 0.5, -1.0, 0.0, 0.0],     # out through r = 0.5
                   [0.49, -0.2, 0.5, 0.0, 1.0, 0.0],      # chord through the hole: removed
                   [0.2, 0.75, 0.98, 0.0, 0.0, 1.0],      # through z = 1: wrapped
                   [-0.8, 0.05, 0.01, 0.0, 0.0, -1.0],    # through z = 0: wrapped
                   [0.60, -0.2, 0.5, 0.0, 1.0, 0.0],      # stays inside
                   [0.0, -0.95, 0.3, 0.1, 0.05, 0.02],    # near the outer wall, stays
                   [-0.51, 0.1, 0.7, 0.0, -0.1, 0.0]])    # near the inner wall, stays
    dts = [0.05, 0.05, 0.4, 0.05, 0.05, 0.4, 0.4, 0.4]
    # one step each (different dt per particle: one stage per particle group)
    for dt in sorted(set(dts)):
        sel = [k for k, d in enumerate(dts) if d == dt]
        p = ps[sel]
        print("case bc tracking hand-placed dt=%g particles %s" % (dt, sel), flush=True)
        kw = dict(tracking=stage.TRACK_REFMAPPING, refmapping_guess=2, charge=(0.0,), mass=(1.0,), mpf=(1.0,))
        with stage.PICStage(m, 3, max_particles=len(p), **kw) as s:
            s.set_field(np.zeros((m.nElems, 64, 3)))
            s.set_particles(p, elem=None, species=np.ones(len(p), np.int32))
            s.time_step(dt, 1)
            ps_d, el_d, in_d = s.particles()
        o = oracle_mod.Oracle(m, 3, tracking=oracle_mod.TRACK_REFMAPPING, refmapping_guess=2, charge=(0.0,),
                              mass=(1.0,), mpf=(1.0,))
        o.set_field(np.zeros((m.nElems, 64, 3)))
        o.set_particles(p, elem=np.zeros(len(p), np.int32))
        assert o.locate() == 0
        o.time_step(dt, 1)
        ps_o, el_o, in_o = o.particles()
        assert np.array_equal(in_d, in_o)
        assert np.array_equal(el_d[in_o], el_o[in_o])
        assert np.abs(ps_d[in_o, :3] - ps_o[in_o, :3]).max(initial=0) <= 1e-12
        expect = {0: False, 1: False, 2: False, 3: True, 4: True, 5: True, 6: True, 7: True}
        assert list(in_o) == [expect[k] for k in sel]


@pytest.mark.parametrize("staged", [False, True])
def test_bc_tracking_large_steps_parity(oracle_mod, staged):
    w = W.annulus_plasma(shape=(4, 16, 4), n_total=6000, dt=1e-8, seed=20261207)
    s, o = run_pair(oracle_mod, w, 10, staged=staged)
    ps, el, ins = s.particles()
    ops, oel, oins = o.particles()
    assert np.array_equal(ins, oins), "inside flags differ: %d" % (ins != oins).sum()
    assert np.array_equal(el[ins], oel[ins]), "owners differ: %d" % (el[ins] != oel[ins]).sum()
    assert s.stats()["n_lost"] == 0 and o.lost() == 0
    assert np.abs(ps[ins, :3] - ops[ins, :3]).max() <= 1e-12
    assert np.abs(s.part_pos_ref()[ins] - o.part_pos_ref()[ins]).max() <= 1e-12
    # removed particles whose final position is back inside the annulus left
    # and re-entered within one step
    r = np.hypot(ops[~oins, 0], ops[~oins, 1])
    reentered = int(((r > 0.5) & (r < 1.0)).sum())
    print("removed %d, of which back inside the annulus %d" % ((~oins).sum(), reentered))
    assert (~oins).sum() > 100
    s.close()
