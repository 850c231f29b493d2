"""A Fortran (flang) host driving the stage through ISO_C_BINDING
(piclas_amd/fortran/mod_pic_stage_c.f90): the 508 loop written as the
reference writes it, checked against the CPU oracle."""
import os
import subprocess

import numpy as np
import pytest

from piclas_amd import stage, workloads as W

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "piclas_amd", "fortran", "pic_stage_driver")


def write_case(path, w, nsteps, mode, deposition):
    m = w.mesh
    n = len(w.PartState)
    hdr = np.zeros(16, np.int32)
    hdr[:15] = [m.nElems, m.SideInfo.shape[0], m.NodeCoords.shape[0], m.NGeo, w.N, m.BCType.shape[0],
                m.PeriodicVectors.shape[0], n, nsteps, mode, deposition, w.sf_alpha, len(w.charge), w.tracking,
                w.refmapping_guess]
    with open(path, "wb") as f:
        f.write(hdr.tobytes())
        for a, t in ((m.ElemInfo, np.int32), (m.SideInfo, np.int32), (m.NodeCoords, np.float64),
                     (m.GlobalNodeIDs, np.int32), (m.BCType, np.int32), (m.PartBoundCond, np.int32)):
            f.write(np.ascontiguousarray(a, t).tobytes())
        if m.PeriodicVectors.shape[0]:
            f.write(np.ascontiguousarray(m.PeriodicVectors, np.float64).tobytes())
        f.write(np.array([w.dt, w.sf_radius, 1e-4, *w.external_field], np.float64).tobytes())
        f.write(np.array([[w.charge[s], w.mass[s], w.mpf[s]] for s in range(len(w.charge))], np.float64).tobytes())
        f.write(np.ascontiguousarray(w.E, np.float64).tobytes())
        f.write(np.ascontiguousarray(w.PartState, np.float64).tobytes())
        f.write(np.ascontiguousarray(w.elem, np.int32).tobytes())
        f.write(np.ascontiguousarray(w.species, np.int32).tobytes())


def read_result(path, n, nElems, nDof):
    raw = open(path, "rb").read()
    o = 0
    ps = np.frombuffer(raw, np.float64, n * 6, o).reshape(n, 6); o += n * 48
    el = np.frombuffer(raw, np.int32, n, o); o += n * 4
    ins = np.frombuffer(raw, np.int8, n, o).astype(bool); o += n
    PS = np.frombuffer(raw, np.float64, nElems * nDof * 4, o).reshape(nElems, nDof, 4); o += nElems * nDof * 32
    tail = np.frombuffer(raw, np.float64, 4, o)
    return ps, el, ins, PS, tail[:3], tail[3]


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("depo", [stage.DEPO_SF, stage.DEPO_CVWM])
def test_fortran_host_matches_oracle(oracle_mod, tmp_path, mode, depo):
    assert os.path.exists(DRIVER), "build the Fortran driver with __graft_entry__.build()"
    w = W.thermal_cube(nx=5, ppc=20, deposition=depo, steps=4)
    if depo == stage.DEPO_CVWM:
        w.sf_radius = 0.0
    case, res = str(tmp_path / "case.bin"), str(tmp_path / "res.bin")
    write_case(case, w, 4, mode, depo)
    print("case fortran host mode=%d depo=%d" % (mode, depo), flush=True)
    out = subprocess.run([DRIVER, case, res], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    ps, el, ins, PS, part_charge, wel = read_result(res, len(w.PartState), w.mesh.nElems, (w.N + 1) ** 3)
    o = oracle_mod.Oracle(w.mesh, w.N, deposition=depo, sf_radius=w.sf_radius, sf_alpha=w.sf_alpha,
                          external_field=w.external_field, charge=w.charge, mass=w.mass, mpf=w.mpf)
    o.set_field(w.E)
    o.set_particles(w.PartState, elem=w.elem, species=w.species)
    o.time_step(w.dt, 4)
    o.deposition()
    ops, oel, oins = o.particles()
    assert np.array_equal(ins, oins) and np.array_equal(el, oel)
    assert np.abs(ps[:, :3] - ops[:, :3]).max() <= 1e-12
    assert np.abs(ps[:, 3:] - ops[:, 3:]).max() <= 1e-12 * np.abs(ops[:, 3:]).max()
    oPS = o.part_source()
    assert np.abs(PS - oPS).max() <= 1e-12 * np.abs(oPS).max()
    # CalcDepositedCharge / CalcPotentialEnergy through the Fortran wrappers
    assert abs(part_charge[0] - o.deposited_charge()) <= 1e-12 * abs(o.deposited_charge())
    qp = o.particle_charge()
    assert abs(part_charge[1] - abs(qp - part_charge[0])) <= 1e-12 * abs(qp)
    if depo == stage.DEPO_CVWM:   # node-volume weighting conserves charge to round-off
        assert part_charge[2] <= 1e-12
    assert abs(wel - o.field_energy()) <= 1e-12 * o.field_energy()


@pytest.mark.parametrize("mode", [0, 1])
def test_fortran_host_refmapping_annulus(oracle_mod, tmp_path, mode):
    """The Fortran host on curved (NGeo=2) elements with RefMapping; the library
    places the particles (no element handed over)."""
    w = W.annulus_plasma(shape=(3, 12, 3), n_total=2000, dt=2e-9)
    case, res = str(tmp_path / "case.bin"), str(tmp_path / "res.bin")
    write_case(case, w, 5, mode, w.deposition)
    print("case fortran host refmapping mode=%d" % mode, flush=True)
    out = subprocess.run([DRIVER, case, res], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    ps, el, ins, PS, part_charge, wel = read_result(res, len(w.PartState), w.mesh.nElems, (w.N + 1) ** 3)
    o = oracle_mod.Oracle(w.mesh, w.N, tracking=w.tracking, refmapping_guess=w.refmapping_guess,
                          deposition=w.deposition, sf_radius=w.sf_radius, sf_alpha=w.sf_alpha,
                          external_field=w.external_field, charge=w.charge, mass=w.mass, mpf=w.mpf)
    o.set_field(w.E)
    o.set_particles(w.PartState, elem=w.elem, species=w.species)
    assert o.locate() == 0
    o.time_step(w.dt, 5)
    o.deposition()
    ops, oel, oins = o.particles()
    assert np.array_equal(ins, oins) and np.array_equal(el[ins], oel[ins])
    assert np.abs(ps[ins, :3] - ops[ins, :3]).max() <= 1e-12
    assert np.abs(ps[ins, 3:] - ops[ins, 3:]).max() <= 1e-12 * np.abs(ops[ins, 3:]).max()
    oPS = o.part_source()
    assert np.abs(PS - oPS).max() <= 1e-12 * np.abs(oPS).max()
    assert abs(wel - o.field_energy()) <= 1e-12 * o.field_energy()
