"""The C-ABI library: it loads, exports every entry point include/pic_stage.h
declares, and rejects bad arguments before touching a GPU."""
import ctypes
import os
import re

import numpy as np
import pytest

from piclas_amd import mesh as M
from piclas_amd import stage

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_entry_points():
    txt = open(os.path.join(ROOT, "include", "pic_stage.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char|void)\s*\**\s*(pic_\w+)\s*\(", txt, re.M)))


def test_library_exports_every_declared_symbol():
    L = stage.load_library()
    names = declared_entry_points()
    assert len(names) >= 18
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(stage.EXPORTS), set(names) ^ set(stage.EXPORTS)


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(RuntimeError):
        stage._lib = None
        try:
            stage.load_library(str(tmp_path / "nope.so"))
        finally:
            stage._lib = None
    stage.load_library()


def _create(**over):
    m = M.cartesian_box((2, 2, 2))
    kw = dict(N=3)
    kw.update(over)
    return stage.PICStage(m, **kw)


@pytest.mark.parametrize("bad,msg", [
    (dict(tracking=2), "TriaTracking"),
    (dict(charge=tuple([1.0] * 9), mass=tuple([1.0] * 9), mpf=tuple([1.0] * 9)), None),
    (dict(refmapping_guess=3), "RefMappingGuess"),
    (dict(refmapping_guess=5, tracking=stage.TRACK_REFMAPPING), "RefMappingGuess"),
    (dict(tracking=stage.TRACK_REFMAPPING, deposition=stage.DEPO_CVWM), "cell_volweight_mean needs TriaTracking"),
    (dict(deposition=stage.DEPO_SF, sf_radius=0.1, sf_alpha=1), "alpha"),
    (dict(max_particles=0), "max_particles"),
    (dict(precision=stage.PREC_MIXED), "mixed precision is implemented for shape_function"),
    (dict(precision=2), "precision must be"),
    (dict(N=9), "N out of range"),
])
def test_create_rejects_bad_parameters(bad, msg):
    if "charge" in bad:
        with pytest.raises((IndexError, stage.StageError)):
            _create(**bad)
        return
    with pytest.raises(stage.StageError) as ei:
        _create(**bad)
    assert msg in str(ei.value)


def test_create_rejects_curved_triatracking():
    m = M.annulus((2, 4, 2))
    with pytest.raises(stage.StageError, match="NGeo"):
        stage.PICStage(m, 3, tracking=stage.TRACK_TRIA)


def test_create_rejects_reflective_and_mortars():
    m = M.cartesian_box((2, 2, 2), periodic=(False, True, True), wall_bc=M.PART_BC_REFLECTIVE)
    with pytest.raises(stage.StageError, match="boundary condition"):
        stage.PICStage(m, 1)
    m = M.cartesian_box((2, 2, 2))
    m.SideInfo = m.SideInfo.copy()
    m.SideInfo[0, 0] = 104 + 100
    with pytest.raises(stage.StageError, match="mortar"):
        stage.PICStage(m, 1)

