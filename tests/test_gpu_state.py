"""State-file particle data on the GPU: ParticleRestart and FillParticleData
through the C ABI (pic_restart_particles, pic_get_particle_data).

Pins:
  * the reference's own restart data (plasma-ball fixture, PartData / PartInt
    written by the reference): every particle passes the in-element check
    in its stored element, and the deposition reproduces the reference's
    DG_Source (|d rho| <= 1e-12 max|rho|);
  * particles listed under the wrong element are located again, into the
    element the reference's PartInt gives them, and counted as missing;
  * the element-sorted output equals the stage's particles grouped by element
    in caller order, and a write -> restart round trip through an HDF5 state
    file restores them bit-exactly (TriaTracking and RefMapping).
"""
import os

import numpy as np
import pytest

from piclas_amd import mesh as M
from piclas_amd import stage, state_file as SF, workloads as W

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def plasma_ball():
    d = np.load(os.path.join(HERE, "golden", "cvwm_plasma_ball.npz"))
    names = [str(x) for x in d["BCNames"]]
    m = M.from_hopr_arrays(d["ElemInfo"], d["SideInfo"], d["NodeCoords"], d["GlobalNodeIDs"], d["BCType"],
                           {nm: M.PART_BC_PERIODIC for nm in names}, names, d["PeriodicVectors"])
    assert float(d["File_Version"]) < 2.8          # PartInt stored as (2, nElems)
    return d, m, np.ascontiguousarray(d["PartInt"].T, np.int64)


def ball_stage(d, m, n):
    return stage.PICStage(m, int(d["N"]), deposition=stage.DEPO_CVWM, charge=tuple(d["ChargeIC"]),
                          mass=tuple(d["MassIC"]), mpf=tuple(d["MacroParticleFactor"]), max_particles=n)


def elem_of_rows(pi, n):
    el = np.zeros(n, np.int32)
    for e in range(pi.shape[0]):
        el[pi[e, 0]:pi[e, 1]] = e + 1
    return el


def test_restart_from_reference_restart_data():
    d, m, pi = plasma_ball()
    pd = d["PartData"]
    print("case plasma ball restart, n=%d" % len(pd), flush=True)
    with ball_stage(d, m, len(pd)) as s:
        miss, lost = s.restart_particles(pd, pi)
        assert (miss, lost) == (0, 0)
        ps, el, ins = s.particles()
        assert ins.all() and np.array_equal(el, elem_of_rows(pi, len(pd)))
        assert np.array_equal(ps, pd[:, :6])
        s.deposition()
        PS = s.part_source()
        ref = d["DG_Source"].reshape(PS.shape)
        assert np.abs(PS - ref).max() <= 1e-12 * np.abs(ref).max()
        # FillParticleData gives the file's arrays back
        opd, opi = s.particle_data()
        assert np.array_equal(opi, pi) and np.array_equal(opd, pd[:, :7])


def test_restart_locates_particles_listed_under_wrong_elements():
    d, m, pi = plasma_ball()
    pd = d["PartData"][::-1].copy()                 # rows reversed, PartInt kept
    true_el = elem_of_rows(pi, len(pd))[::-1]
    listed = elem_of_rows(pi, len(pd))
    print("case plasma ball restart with permuted rows, n=%d" % len(pd), flush=True)
    with ball_stage(d, m, len(pd)) as s:
        miss, lost = s.restart_particles(pd, pi)
        assert lost == 0 and miss == int((listed != true_el).sum()) and miss > 0
        ps, el, ins = s.particles()
        assert ins.all() and np.array_equal(el, true_el)
        assert np.array_equal(ps, pd[:, :6])


def test_restart_rejects_corrupt_data():
    d, m, pi = plasma_ball()
    pd = d["PartData"].copy()
    with ball_stage(d, m, len(pd)) as s:
        bad = pd.copy()
        bad[5, 6] = 0.0                             # species ID zero: the reference aborts
        with pytest.raises(stage.StageError):
            s.restart_particles(bad, pi)
        gap = pi.copy()
        gap[3, 1] += 1                              # ranges no longer tile the rows
        with pytest.raises(stage.StageError):
            s.restart_particles(pd, gap)


def grouped_by_element(el, ins):
    keep = np.flatnonzero(ins)
    order = keep[np.argsort(el[keep], kind="stable")]
    return order


def test_particle_data_after_steps_and_file_round_trip(tmp_path):
    w = W.thermal_cube(nx=6, ppc=24, steps=10)
    kw = W.stage_kwargs(w)
    print("case particle data thermal cube n=%d" % len(w.PartState), flush=True)
    with stage.PICStage(w.mesh, w.N, max_particles=len(w.PartState), **kw) as s:
        s.set_field(w.E)
        s.set_particles(w.PartState, elem=w.elem, species=w.species)
        s.time_step(w.dt, 10)
        ps, el, ins = s.particles()
        pd, pi = s.particle_data()
        order = grouped_by_element(el, ins)
        assert pd.shape == (len(order), 7)
        assert np.array_equal(pd[:, :6], ps[order])
        assert np.array_equal(pd[:, 6], np.asarray(w.species, np.float64)[order])
        cnt = np.bincount(el[ins] - 1, minlength=s.nElems)
        assert np.array_equal(pi[:, 1] - pi[:, 0], cnt) and pi[0, 0] == 0 and pi[-1, 1] == len(order)
        s.deposition()
        PS = s.part_source()
        if not os.path.exists(SF.HDF5_LIB):
            pytest.skip("HDF5 C library absent")
        path = str(tmp_path / "cube_State.h5")
        SF.write_stage_state(path, s, time=10 * w.dt, mesh_file="cube")
    with stage.PICStage(w.mesh, w.N, max_particles=len(w.PartState), **kw) as r:
        r.set_field(w.E)
        assert SF.restart_stage(r, path) == (0, 0)
        rpd, rpi = r.particle_data()
        assert np.array_equal(rpd, pd) and np.array_equal(rpi, pi)
        r.deposition()
        rPS = r.part_source()
        assert np.abs(rPS - PS).max() <= 1e-12 * np.abs(PS).max()
        st = SF.read_state(path)
        assert np.array_equal(st["DG_Source"], PS)


def test_refmapping_restart_round_trip():
    w = W.annulus_plasma(shape=(4, 16, 4), n_total=4000, dt=2e-9)
    kw = W.stage_kwargs(w)
    print("case refmapping restart annulus n=%d" % len(w.PartState), flush=True)
    with stage.PICStage(w.mesh, w.N, max_particles=len(w.PartState), **kw) as s:
        s.set_field(w.E)
        s.set_particles(w.PartState, elem=w.elem, species=w.species)
        s.time_step(w.dt, 5)
        pd, pi = s.particle_data()
        ps, el, ins = s.particles()
        order = grouped_by_element(el, ins)
        ref = s.part_pos_ref()[order]
    with stage.PICStage(w.mesh, w.N, max_particles=len(w.PartState), **kw) as r:
        r.set_field(w.E)
        assert r.restart_particles(pd, pi) == (0, 0)
        rps, rel, rins = r.particles()
        assert rins.all() and np.array_equal(rps, pd[:, :6])
        assert np.array_equal(rel, elem_of_rows(pi, len(pd)))
        assert np.abs(r.part_pos_ref() - ref).max() <= 1e-12
