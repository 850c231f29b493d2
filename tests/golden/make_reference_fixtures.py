"""Extract small golden fixtures from data files that the reference's own
regression tests hold (HDF5 meshes and restart states).

The reference stores these as HDF5; this script reads them with the HDF5 C
library of the image (ctypes, no h5py) and writes compressed ``.npz`` fixtures
next to itself.  It runs only where the reference tree is present; the
fixtures it writes are committed, so tests never read the reference itself.

Fixtures written:
  2Dplasma_test_State_000.00000005000000000.h5
      regressioncheck/NIG_PIC_poisson_Leapfrog/2D_innerBC_dielectric_surface_charge/
      a state file the reference wrote, its userblock (parameter file and
      build notes, not data) zeroed: File_Version 3.21,
      PartInt (nElems, 2), PartData, VarNamesParticles, 5-D DG_Source), the
      layout check of piclas_amd/state_file.py.
  cvwm_plasma_ball.npz
      regressioncheck/NIG_PIC_Deposition/Plasma_Ball_cell_volweight_mean/
        Box_mesh.h5  (HOPR mesh, 10x10x10 periodic box [-1,1]^3, N=1)
        plasma_wave_State_000.00000000000000000_restart.h5
          PartData (x,y,z,vx,vy,vz,species), PartInt, DG_Source (the
          cell_volweight_mean source the reference deposited for PartData)
      plus the species / BC constants of that case's parameter.ini and
      the total-charge pin of PartAnalyze_ref.csv.
"""
import ctypes
import os
import sys

import numpy as np

REF = os.environ.get("PICLAS_REFERENCE", "/root/reference")
HERE = os.path.dirname(os.path.abspath(__file__))

_hid = ctypes.c_int64


def _lib():
    lib = ctypes.CDLL("/opt/conda/lib/libhdf5.so")
    lib.H5open()
    lib.H5Fopen.restype = _hid
    lib.H5Fopen.argtypes = [ctypes.c_char_p, ctypes.c_uint, _hid]
    lib.H5Fclose.argtypes = [_hid]
    lib.H5Dopen2.restype = _hid
    lib.H5Dopen2.argtypes = [_hid, ctypes.c_char_p, _hid]
    lib.H5Dclose.argtypes = [_hid]
    lib.H5Dget_space.restype = _hid
    lib.H5Dget_space.argtypes = [_hid]
    lib.H5Dget_type.restype = _hid
    lib.H5Dget_type.argtypes = [_hid]
    lib.H5Sget_simple_extent_ndims.argtypes = [_hid]
    lib.H5Sget_simple_extent_dims.argtypes = [
        _hid, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)]
    lib.H5Tget_class.argtypes = [_hid]
    lib.H5Tget_size.restype = ctypes.c_size_t
    lib.H5Tget_size.argtypes = [_hid]
    lib.H5Dread.argtypes = [_hid, _hid, _hid, _hid, _hid, ctypes.c_void_p]
    return lib


def read_dataset(lib, path, name):
    """Read one dataset; returns a numpy array in HDF5 (C) dimension order."""
    f = lib.H5Fopen(path.encode(), 0, 0)
    if f < 0:
        raise IOError(path)
    d = lib.H5Dopen2(f, name.encode(), 0)
    if d < 0:
        raise KeyError(name)
    s = lib.H5Dget_space(d)
    nd = lib.H5Sget_simple_extent_ndims(s)
    dims = (ctypes.c_uint64 * nd)()
    lib.H5Sget_simple_extent_dims(s, dims, None)
    shape = tuple(int(x) for x in dims)
    t = lib.H5Dget_type(d)
    cls, size = lib.H5Tget_class(t), lib.H5Tget_size(t)
    if cls == 0:      # H5T_INTEGER
        mt = ctypes.c_int64.in_dll(lib, "H5T_NATIVE_INT_g" if size == 4 else "H5T_NATIVE_LLONG_g").value
        arr = np.empty(shape, dtype=np.int32 if size == 4 else np.int64)
    elif cls == 1:    # H5T_FLOAT
        mt = ctypes.c_int64.in_dll(lib, "H5T_NATIVE_DOUBLE_g").value
        arr = np.empty(shape, dtype=np.float64)
    elif cls == 3:    # H5T_STRING, fixed length
        mt = t
        arr = np.empty(shape, dtype="S%d" % size)
    else:
        raise TypeError((name, cls))
    rc = lib.H5Dread(d, mt, 0, 0, 0, arr.ctypes.data)
    if rc < 0:
        raise IOError(name)
    lib.H5Dclose(d)
    lib.H5Fclose(f)
    return arr


def cvwm_plasma_ball(lib):
    case = os.path.join(REF, "regressioncheck/NIG_PIC_Deposition/Plasma_Ball_cell_volweight_mean")
    mesh = os.path.join(case, "Box_mesh.h5")
    state = os.path.join(case, "plasma_wave_State_000.00000000000000000_restart.h5")
    out = {}
    for n in ("ElemInfo", "SideInfo", "NodeCoords", "GlobalNodeIDs", "BCType"):
        out[n] = read_dataset(lib, mesh, n)
    out["BCNames"] = np.array([s.rstrip(b"\0 ").decode() for s in read_dataset(lib, mesh, "BCNames")])
    for n in ("PartData", "PartInt", "DG_Source"):
        out[n] = read_dataset(lib, state, n)
    # parameter.ini of the case: species 1, MPF, charge; periodic vectors from hopr.ini (VV, la=2)
    out["ChargeIC"] = np.array([1.60217653e-5, -1.60217653e-19])
    out["MassIC"] = np.array([1.0, 9.1093826e-31])
    out["MacroParticleFactor"] = np.array([200.0, 200.0])
    out["PeriodicVectors"] = np.array([[2.0, 0, 0], [0, 2.0, 0], [0, 0, 2.0]])
    out["N"] = np.array(1)
    # PartAnalyze_ref.csv: deposited charge and its absolute error
    out["TotalChargeRef"] = np.array([1.06801087489800e1, 1.49213974509621e-13])
    # File_Version of the state file: 1.5, so PartInt is stored as PartInt(nElems,2)
    sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
    from piclas_amd import state_file
    with state_file._File(state) as f:
        out["File_Version"] = np.array(float(f.attr("File_Version")))
    np.savez_compressed(os.path.join(HERE, "cvwm_plasma_ball.npz"), **out)
    return out


HDF5_SIGNATURE = b"\x89HDF\r\n\x1a\n"


def blank_userblock(data):
    """Zero the HDF5 userblock (everything before the superblock).

    The reference writes its parameter file, build configuration and a source
    diff into the userblock of every state file.  That is not data: only the
    datasets and attributes after the superblock are kept.  The HDF5 library
    ignores the userblock's content, so the file still opens unchanged.  The
    superblock sits at offset 0, 512, 1024, 2048, ... (HDF5 file format)."""
    data = bytearray(data)
    off = 0
    while data[off:off + 8] != HDF5_SIGNATURE:
        off = 512 if off == 0 else 2 * off
        if off >= len(data):
            raise ValueError("no HDF5 superblock found")
    data[:off] = bytes(off)
    return bytes(data)


def copy_state_file():
    """The state file's datasets and attributes, with its userblock zeroed."""
    src = os.path.join(REF, "regressioncheck/NIG_PIC_poisson_Leapfrog/2D_innerBC_dielectric_surface_charge",
                       "2Dplasma_test_State_000.00000005000000000.h5")
    with open(src, "rb") as f, open(os.path.join(HERE, os.path.basename(src)), "wb") as g:
        g.write(blank_userblock(f.read()))


if __name__ == "__main__":
    if not os.path.isdir(REF):
        sys.exit("reference tree not found at %s" % REF)
    lib = _lib()
    copy_state_file()
    o = cvwm_plasma_ball(lib)
    for k, v in o.items():
        print(k, getattr(v, "shape", None), getattr(v, "dtype", None))
