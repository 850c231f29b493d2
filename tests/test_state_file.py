"""State-file layer (piclas_amd/state_file.py), on the CPU.

Pins:
  * a state file the reference wrote (tests/golden/2Dplasma_test_State_*.h5,
    File_Version 3.21) reads into PartData / PartInt / DG_Source of the shapes
    the reference's reader expects (particle_readin.f90:503-637), with PartInt
    tiling the particle rows;
  * the plasma-ball fixture (File_Version 1.5) holds PartInt in the layout
    before 2.8, (2, nElems) on disk, and its ranges tile its 3,333 rows only
    in that reading;
  * what write_state writes reads back unchanged, with the datasets in the
    reference's on-disk orientation.
"""
import os

import numpy as np
import pytest

from piclas_amd import state_file as SF

HERE = os.path.dirname(os.path.abspath(__file__))
REF_STATE = os.path.join(HERE, "golden", "2Dplasma_test_State_000.00000005000000000.h5")

pytestmark = pytest.mark.skipif(not os.path.exists(SF.HDF5_LIB), reason="HDF5 C library absent")


def tiles(pi, n):
    pi = np.asarray(pi)
    return pi[0, 0] == 0 and (pi[1:, 0] == pi[:-1, 1]).all() and (pi[:, 1] >= pi[:, 0]).all() and pi[-1, 1] == n


def test_reads_reference_state_file():
    st = SF.read_state(REF_STATE)
    assert st["File_Type"] == "State"
    assert abs(float(st["File_Version"]) - 3.21) < 1e-12
    assert int(st["N"]) == 1
    pd, pi = st["PartData"], st["PartInt"]
    assert pd.shape == (654, 7) and pi.shape == (101, 2) and pi.dtype == np.int64
    assert tiles(pi, pd.shape[0])
    assert st["VarNamesParticles"] == list(SF.PART_DATA_VAR_NAMES)
    sp = pd[:, 6]
    assert (sp == np.round(sp)).all() and (sp >= 1).all()
    assert st["DG_Source"].shape == (101, 8, 4)


def test_old_partint_layout_of_plasma_ball_fixture(tmp_path):
    d = np.load(os.path.join(HERE, "golden", "cvwm_plasma_ball.npz"))
    raw = d["PartInt"]                       # as stored on disk: dims (2, nElems)
    assert float(d["File_Version"]) < 2.8 and raw.shape == (2, 1000)
    n = d["PartData"].shape[0]
    assert tiles(raw.T, n) and not tiles(raw.reshape(1000, 2), n)
    # a file in that layout reads back as (nElems, 2)
    p = str(tmp_path / "old.h5")
    with SF._File(p, write=True) as f:
        f.write_attr("File_Version", float(d["File_Version"]))
        f.write_attr("N", 1)
        f.write_dataset("PartInt", raw.astype(np.int64))
        f.write_attr("VarNamesParticles", list(SF.PART_DATA_VAR_NAMES))
        f.write_dataset("PartData", d["PartData"])
        f.write_dataset("DG_Source", d["DG_Source"])
    st = SF.read_state(p)
    assert np.array_equal(st["PartInt"], raw.T)
    assert np.array_equal(st["PartData"], d["PartData"])
    assert np.array_equal(st["DG_Source"], d["DG_Source"].reshape(1000, 8, 4))


def test_write_read_round_trip(tmp_path):
    d = np.load(os.path.join(HERE, "golden", "cvwm_plasma_ball.npz"))
    pi = np.ascontiguousarray(d["PartInt"].T, np.int64)
    src = d["DG_Source"].reshape(1000, 8, 4)
    p = str(tmp_path / "s.h5")
    SF.write_state(p, d["PartData"], pi, 1, time=2.5e-9, mesh_file="Box_mesh.h5", DG_Source=src)
    st = SF.read_state(p)
    assert st["File_Type"] == "State" and st["MeshFile"] == "Box_mesh.h5"
    assert float(st["File_Version"]) == SF.FILE_VERSION and int(st["N"]) == 1 and float(st["Time"]) == 2.5e-9
    assert np.array_equal(st["PartInt"], pi) and np.array_equal(st["PartData"], d["PartData"])
    assert np.array_equal(st["DG_Source"], src)
    # on disk: PartInt(2,nElems) -> (nElems,2), PartData(7,n) -> (n,7), DG_Source(4,nDOF) -> (nDOF,4)
    with SF._File(p) as f:
        assert f.dataset("PartInt").shape == (1000, 2)
        assert f.dataset("PartData").shape == (3333, 7)
        assert f.dataset("DG_Source").shape == (8000, 4)
        assert f.attr("VarNamesSource") == list(SF.SOURCE_VAR_NAMES)


def test_write_rejects_bad_shapes(tmp_path):
    with pytest.raises(ValueError):
        SF.write_state(str(tmp_path / "x.h5"), np.zeros((3, 6)), np.zeros((2, 2), np.int64), 1)
    with pytest.raises(ValueError):
        SF.write_state(str(tmp_path / "x.h5"), np.zeros((3, 7)), np.zeros((2, 3), np.int64), 1)


class _Recorder:
    """Stands in for a stage: records what restart_stage hands over."""
    N = 1

    def restart_particles(self, PartData, PartInt):
        self.args = (PartData, PartInt)
        return 0, 0


def test_restart_maps_particle_variables_by_name(tmp_path):
    # particle_readin.f90:640-700: columns are matched by VarNamesParticles
    d = np.load(os.path.join(HERE, "golden", "cvwm_plasma_ball.npz"))
    perm = [6, 3, 4, 5, 0, 1, 2]
    names = [SF.PART_DATA_VAR_NAMES[k] for k in perm]
    p = str(tmp_path / "perm.h5")
    with SF._File(p, write=True) as f:
        f.write_attr("File_Version", SF.FILE_VERSION)
        f.write_attr("N", 1)
        f.write_dataset("PartInt", np.ascontiguousarray(d["PartInt"].T, np.int64))
        f.write_attr("VarNamesParticles", names)
        f.write_dataset("PartData", np.ascontiguousarray(d["PartData"][:, perm]))
    r = _Recorder()
    SF.restart_stage(r, p)
    assert np.array_equal(r.args[0], d["PartData"])
    assert np.array_equal(r.args[1], d["PartInt"].T)
