"""The Fortran host module's BIND(C) types (mod_pic_stage_c.f90) have the
field offsets and sizes of the C structs of pic_stage.h (as ctypes lays
them out): a program built with flang against the module prints them."""
import os
import shutil
import subprocess

import pytest

from piclas_amd import stage as S

HERE = os.path.dirname(os.path.abspath(__file__))
FDIR = os.path.join(os.path.dirname(HERE), "piclas_amd", "fortran")
TYPES = [("tPicMeshDesc", S.PicMeshDesc), ("tPicParams", S.PicParams), ("tPicEmission", S.PicEmission)]


@pytest.mark.skipif(shutil.which("flang") is None, reason="flang not installed")
def test_bind_c_types_match_c_structs(tmp_path):
    obj = tmp_path / "mod_pic_stage_c.o"
    subprocess.run(["flang", "-c", os.path.join(FDIR, "mod_pic_stage_c.f90"), "-o", str(obj),
                    "-module-dir", str(tmp_path)], check=True, capture_output=True)
    lines = ["PROGRAM layout", "USE ISO_C_BINDING", "USE MOD_PIC_Stage_C", "IMPLICIT NONE",
             "INTEGER(C_INTPTR_T) :: b"]
    for i, (ft, _) in enumerate(TYPES):
        lines.append("TYPE(%s), TARGET :: v%d" % (ft, i))
    for i, (ft, ct) in enumerate(TYPES):
        lines.append("b = TRANSFER(C_LOC(v%d), b)" % i)
        lines.append("WRITE(*,'(A,I0)') '%s size ', C_SIZEOF(v%d)" % (ft, i))
        for name, _ in ct._fields_:
            lines.append("WRITE(*,'(A,I0)') '%s %s ', TRANSFER(C_LOC(v%d%%%s), b) - b" % (ft, name, i, name))
    lines.append("END PROGRAM layout")
    src = tmp_path / "layout.f90"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    # the program calls no stage routine: the module's references to the C ABI stay unresolved
    subprocess.run(["flang", str(src), str(obj), "-I", str(tmp_path), "-o", str(exe),
                    "-Wl,--unresolved-symbols=ignore-all"], check=True, capture_output=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l.strip()}
    import ctypes
    for ft, ct in TYPES:
        assert got[(ft, "size")] == ctypes.sizeof(ct), ft
        for name, _ in ct._fields_:
            assert got[(ft, name)] == getattr(ct, name).offset, (ft, name)
