"""Python mirror of the reference's particle-stage interface over the C ABI.

The reference's stepper (src/timedisc/timedisc_TimeStepPoissonByBorisLeapfrog.f90)
calls Deposition(), InterpolateFieldToParticle(), an inline Boris push and
PerformTracking() on module arrays.  ``PICStage`` exposes the same four
operations (under the reference names and snake_case aliases) on a handle of
libpicstage.so, with the arrays passed in the reference's layout:

  PartState (nPart, 6)   E (nElems, (N+1)^3, 3)   PartSource (nElems, (N+1)^3, 4)

The HIP library is required: loading fails loudly when it is missing.
"""
import ctypes
import os

import numpy as np

from .mesh import HoprMesh

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("PICLAS_AMD_LIB", os.path.join(HERE, "libpicstage.so"))

TRACK_REFMAPPING, TRACK_TRIA = 1, 3
DEPO_NONE, DEPO_CVWM, DEPO_SF = 0, 1, 2
PREC_FP64, PREC_MIXED = 0, 1
# dominant deposition kernel (pic_stats.deposit_kernel)
KERNEL_NONE, KERNEL_CVWM, KERNEL_SF_VALU, KERNEL_SF_MATRIX, KERNEL_SF_CARTESIAN, KERNEL_SF_MIXED = range(6)

ELEMENTARY_CHARGE = 1.602176634e-19   # globals_vars.f90:54
ELECTRON_MASS = 9.1093826e-31         # globals_vars.f90:57
EPS0 = 8.8541878176e-12               # globals_vars.f90:84

_i32p = ctypes.POINTER(ctypes.c_int32)
_f64p = ctypes.POINTER(ctypes.c_double)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_i64p = ctypes.POINTER(ctypes.c_int64)


class PicMeshDesc(ctypes.Structure):
    _fields_ = [("nElems", ctypes.c_int32), ("nSides", ctypes.c_int32), ("nNodes", ctypes.c_int32),
                ("NGeo", ctypes.c_int32), ("N", ctypes.c_int32),
                ("ElemInfo", _i32p), ("SideInfo", _i32p), ("NodeCoords", _f64p), ("GlobalNodeIDs", _i32p),
                ("nBCs", ctypes.c_int32), ("BCType", _i32p), ("PartBoundCond", _i32p),
                ("nPeriodicVectors", ctypes.c_int32), ("PeriodicVectors", _f64p)]


class PicParams(ctypes.Structure):
    _fields_ = [("tracking", ctypes.c_int32), ("deposition", ctypes.c_int32),
                ("sf_radius", ctypes.c_double), ("sf_alpha", ctypes.c_int32),
                ("refmapping_guess", ctypes.c_int32), ("refmapping_eps", ctypes.c_double),
                ("external_field", ctypes.c_double * 6), ("nSpecies", ctypes.c_int32),
                ("charge", ctypes.c_double * 8), ("mass", ctypes.c_double * 8), ("mpf", ctypes.c_double * 8),
                ("precision", ctypes.c_int32), ("max_particles", ctypes.c_int64), ("device", ctypes.c_int32)]


SPACE_CUBOID, SPACE_SIN_DEVIATION = 1, 2
VELO_CONSTANT, VELO_MAXWELL_LPN = 1, 2


class PicEmission(ctypes.Structure):
    _fields_ = [("species", ctypes.c_int32), ("space", ctypes.c_int32), ("particle_number", ctypes.c_int64),
                ("base_point", ctypes.c_double * 3), ("base_vector1", ctypes.c_double * 3),
                ("base_vector2", ctypes.c_double * 3), ("cuboid_height", ctypes.c_double),
                ("max_particle_number", ctypes.c_int32 * 3), ("amplitude", ctypes.c_double),
                ("wave_number", ctypes.c_double), ("velocity", ctypes.c_int32), ("velo_ic", ctypes.c_double),
                ("velo_vec", ctypes.c_double * 3), ("temperature", ctypes.c_double)]


class PicStats(ctypes.Structure):
    _fields_ = [("n_particles", ctypes.c_int64), ("n_lost", ctypes.c_int64), ("n_steps", ctypes.c_int64),
                ("ms_deposit", ctypes.c_double), ("ms_push_track", ctypes.c_double), ("ms_sort", ctypes.c_double),
                ("n_deposit_launches", ctypes.c_int32), ("ms_deposit_main", ctypes.c_double),
                ("deposit_kernel", ctypes.c_int32), ("ms_push_main", ctypes.c_double)]


class PicTuning(ctypes.Structure):
    _fields_ = [("use_graph", ctypes.c_int32), ("sf_kernel", ctypes.c_int32), ("sf_persistent", ctypes.c_int32),
                ("sf_shared_loads", ctypes.c_int32)]


# Entry points declared in include/pic_stage.h: name -> (restype, argtypes)
_H = ctypes.c_void_p
EXPORTS = {
    "pic_create": (ctypes.c_int, [ctypes.POINTER(PicMeshDesc), ctypes.POINTER(PicParams), ctypes.POINTER(ctypes.c_void_p)]),
    "pic_destroy": (ctypes.c_int, [_H]),
    "pic_last_error": (ctypes.c_char_p, []),
    "pic_set_field": (ctypes.c_int, [_H, _f64p]),
    "pic_set_field_device": (ctypes.c_int, [_H, ctypes.c_void_p]),
    "pic_set_particles": (ctypes.c_int, [_H, ctypes.c_int64, _f64p, _i32p, _i32p, _u8p]),
    "pic_deposition": (ctypes.c_int, [_H]),
    "pic_interpolate_field_to_particle": (ctypes.c_int, [_H]),
    "pic_push_boris_leapfrog": (ctypes.c_int, [_H, ctypes.c_double]),
    "pic_perform_tracking": (ctypes.c_int, [_H]),
    "pic_time_step": (ctypes.c_int, [_H, ctypes.c_double, ctypes.c_int32]),
    "pic_get_particles": (ctypes.c_int, [_H, _f64p, _i32p, _u8p]),
    "pic_get_last_part_pos": (ctypes.c_int, [_H, _f64p]),
    "pic_get_field_at_particle": (ctypes.c_int, [_H, _f64p]),
    "pic_get_part_pos_ref": (ctypes.c_int, [_H, _f64p]),
    "pic_get_part_source": (ctypes.c_int, [_H, _f64p]),
    "pic_part_source_device": (ctypes.c_int, [_H, ctypes.POINTER(ctypes.c_void_p)]),
    "pic_deposited_charge": (ctypes.c_int, [_H, _f64p]),
    "pic_field_energy": (ctypes.c_int, [_H, _f64p]),
    "pic_set_tuning": (ctypes.c_int, [_H, ctypes.POINTER(PicTuning)]),
    "pic_set_field_operator": (ctypes.c_int, [_H, _f64p]),
    "pic_sc_time_step": (ctypes.c_int, [_H, ctypes.c_double, ctypes.c_int32, _f64p]),
    "pic_adjoint": (ctypes.c_int, [_H, ctypes.c_double, ctypes.c_int32, ctypes.c_int32, _f64p, _f64p, _f64p]),
    "pic_get_stats": (ctypes.c_int, [_H, ctypes.POINTER(PicStats)]),
    "pic_emit_particles": (ctypes.c_int, [_H, ctypes.c_int32, ctypes.POINTER(PicEmission), ctypes.c_uint64,
                                          ctypes.POINTER(ctypes.c_int64)]),
    "pic_emit_particles_noalloc": (ctypes.c_int, [_H, ctypes.c_int32, ctypes.POINTER(PicEmission), ctypes.c_uint64,
                                                  ctypes.POINTER(ctypes.c_int64)]),
    "pic_get_particle_data": (ctypes.c_int, [_H, ctypes.c_int64, _f64p, _i64p, ctypes.POINTER(ctypes.c_int64)]),
    "pic_restart_particles": (ctypes.c_int, [_H, ctypes.c_int64, _f64p, _i64p, ctypes.POINTER(ctypes.c_int64),
                                             ctypes.POINTER(ctypes.c_int64)]),
    "pic_reset_stats": (ctypes.c_int, [_H]),
    "pic_synchronize": (ctypes.c_int, [_H]),
}

PART_DATA_SIZE = 7   # PIC_PART_DATA_SIZE: x y z vx vy vz species

_lib = None


def load_library(path=LIB_PATH):
    """Load libpicstage.so; raises if it was not built (no fallback path exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise RuntimeError("libpicstage.so not found at %s: build it with __graft_entry__.build()" % path)
        L = ctypes.CDLL(path)
        for name, (res, args) in EXPORTS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _hip_runtime_paths():
    """Paths of the HIP runtime libraries mapped into this process."""
    try:
        with open("/proc/self/maps") as f:
            return {line.split()[-1] for line in f if "libamdhip64" in line}
    except OSError:
        return set()


def _check_one_hip_runtime():
    if len(_hip_runtime_paths()) > 1:
        raise RuntimeError("two HIP runtimes are loaded (the stage library's and torch's): import torch before "
                           "piclas_amd loads the stage library, or hand device memory over with DeviceBuffer")


def _hip():
    """The HIP runtime the stage library is linked against (already loaded)."""
    load_library()
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return hip


class DeviceBuffer:
    """Device memory of the HIP runtime the stage uses, for handing U_N%E and
    PartSource over without host copies (pic_set_field_device,
    pic_part_source_device), e.g. to and from a field solver on the GPU."""
    H2D, D2H = 1, 2

    def __init__(self, nbytes):
        self.hip = _hip()
        p = ctypes.c_void_p()
        if self.hip.hipMalloc(ctypes.byref(p), int(nbytes)) != 0:
            raise StageError("hipMalloc of %d bytes failed" % nbytes)
        self.ptr, self.nbytes = p.value, int(nbytes)

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a, np.float64)
        b = cls(a.nbytes)
        if b.hip.hipMemcpy(ctypes.c_void_p(b.ptr), a.ctypes.data_as(ctypes.c_void_p), a.nbytes, cls.H2D) != 0:
            raise StageError("hipMemcpy to the device failed")
        return b

    def free(self):
        if self.ptr:
            self.hip.hipFree(ctypes.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        self.free()


def device_to_host(ptr, shape):
    """float64 array of the given shape copied from device address ptr."""
    out = np.empty(shape, np.float64)
    if _hip().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(ptr), out.nbytes, DeviceBuffer.D2H) != 0:
        raise StageError("hipMemcpy from the device failed")
    return out


class StageError(RuntimeError):
    pass


def _chk(rc, what):
    if rc != 0:
        raise StageError("%s failed (%d): %s" % (what, rc, load_library().pic_last_error().decode()))


def _p(a, t):
    return a.ctypes.data_as(t)


class PICStage:
    """One particle stage on one GPU (a pic_handle)."""

    def __init__(self, mesh: HoprMesh, N, deposition=DEPO_NONE, sf_radius=0.0, sf_alpha=2, tracking=TRACK_TRIA,
                 refmapping_guess=1, refmapping_eps=1e-4, external_field=(0, 0, 0, 0, 0, 0),
                 charge=(-ELEMENTARY_CHARGE,), mass=(ELECTRON_MASS,), mpf=(1.0,), max_particles=1 << 20,
                 precision=PREC_FP64, device=0):
        L = load_library()
        self._keep = [np.ascontiguousarray(mesh.ElemInfo, np.int32), np.ascontiguousarray(mesh.SideInfo, np.int32),
                      np.ascontiguousarray(mesh.NodeCoords, np.float64),
                      np.ascontiguousarray(mesh.GlobalNodeIDs, np.int32),
                      np.ascontiguousarray(mesh.BCType, np.int32), np.ascontiguousarray(mesh.PartBoundCond, np.int32),
                      np.ascontiguousarray(mesh.PeriodicVectors, np.float64).reshape(-1, 3)]
        ei, si, nc, gid, bct, pbc, pv = self._keep
        md = PicMeshDesc(ei.shape[0], si.shape[0], nc.shape[0], mesh.NGeo, N, _p(ei, _i32p), _p(si, _i32p),
                         _p(nc, _f64p), _p(gid, _i32p), bct.shape[0], _p(bct, _i32p), _p(pbc, _i32p),
                         pv.shape[0], _p(pv, _f64p))
        par = PicParams()
        par.tracking, par.deposition = tracking, deposition
        par.sf_radius, par.sf_alpha = sf_radius, sf_alpha
        par.refmapping_guess, par.refmapping_eps = refmapping_guess, refmapping_eps
        for i in range(6):
            par.external_field[i] = external_field[i]
        par.nSpecies = len(charge)
        for i in range(len(charge)):
            par.charge[i], par.mass[i], par.mpf[i] = charge[i], mass[i], mpf[i]
        par.precision, par.max_particles, par.device = precision, max_particles, device
        h = ctypes.c_void_p()
        _chk(L.pic_create(ctypes.byref(md), ctypes.byref(par), ctypes.byref(h)), "pic_create")
        self.h = h
        self.nElems, self.N, self.nDof = ei.shape[0], N, (N + 1) ** 3
        self.nPart = 0

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            load_library().pic_destroy(self.h)
            self.h = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---------------------------------------------------------------- inputs
    def set_field(self, E):
        E = np.ascontiguousarray(E, np.float64)
        if E.size != self.nElems * self.nDof * 3:
            raise ValueError("E must have nElems*(N+1)^3*3 entries")
        _chk(load_library().pic_set_field(self.h, _p(E, _f64p)), "pic_set_field")

    def set_field_device(self, ptr):
        """U_N%E from device memory: ptr is the device address of
        [nElems][(N+1)^3][3] float64 values (copied on the stage's stream)."""
        _chk(load_library().pic_set_field_device(self.h, ctypes.c_void_p(int(ptr))), "pic_set_field_device")

    def set_field_tensor(self, E):
        """U_N%E from a contiguous float64 tensor on the stage's GPU, without a
        host round trip; waits for the work torch has queued on its stream.
        torch and the stage must share one HIP runtime: import torch before
        this module loads the stage library."""
        import torch
        _check_one_hip_runtime()
        if not (E.is_cuda and E.dtype == torch.float64 and E.is_contiguous()):
            raise ValueError("E must be a contiguous float64 CUDA tensor")
        if E.numel() != self.nElems * self.nDof * 3:
            raise ValueError("E must have nElems*(N+1)^3*3 entries")
        torch.cuda.current_stream(E.device).synchronize()
        self.set_field_device(E.data_ptr())

    def part_source_device(self):
        """Device address of PartSource [nElems][(N+1)^3][4] (float64), valid
        until the next deposition."""
        out = ctypes.c_void_p()
        _chk(load_library().pic_part_source_device(self.h, ctypes.byref(out)), "pic_part_source_device")
        return out.value

    def part_source_tensor(self, device=None):
        """PartSource as a torch tensor viewing the stage's device memory (no
        copy), shape (nElems, (N+1)^3, 4); valid until the next deposition."""
        import torch
        _check_one_hip_runtime()

        class _View:
            pass
        v = _View()
        v.__cuda_array_interface__ = {"shape": (self.nElems, self.nDof, 4), "typestr": "<f8",
                                      "data": (self.part_source_device(), False), "version": 2, "strides": None}
        return torch.as_tensor(v, device=device or torch.device("cuda", 0))

    def emit_particles(self, inits, seed=1, noalloc=False):
        """Particle emission on the device (SetParticlePosition 'cuboid' /
        'sin_deviation', SetParticleVelocity 'constant' / 'maxwell_lpn'):
        inits is a list of dicts with the PicEmission fields (species is
        1-based); sets them as the particles, located on the device.
        noalloc selects pic_emit_particles_noalloc (fixed device staging, no
        allocation in the call, at most 524,288 particles).
        Returns the number of particles inside the mesh."""
        arr = (PicEmission * len(inits))()
        for a, d in zip(arr, inits):
            for k, v in d.items():
                if isinstance(v, (list, tuple, np.ndarray)):
                    getattr(a, k)[:] = list(v)
                else:
                    setattr(a, k, v)
        out = ctypes.c_int64(0)
        name = "pic_emit_particles_noalloc" if noalloc else "pic_emit_particles"
        _chk(getattr(load_library(), name)(self.h, len(inits), arr, ctypes.c_uint64(seed), ctypes.byref(out)), name)
        self.nPart = sum(d.get("particle_number", 0) if d.get("space", SPACE_CUBOID) == SPACE_CUBOID
                         else int(np.prod(d["max_particle_number"])) for d in inits)
        return out.value

    def set_particles(self, PartState, elem=None, species=None, is_new=None):
        ps = np.ascontiguousarray(PartState, np.float64).reshape(-1, 6)
        n = ps.shape[0]
        el = None if elem is None else np.ascontiguousarray(elem, np.int32)
        sp = None if species is None else np.ascontiguousarray(species, np.int32)
        nw = None if is_new is None else np.ascontiguousarray(is_new, np.uint8)
        _chk(load_library().pic_set_particles(self.h, n, _p(ps, _f64p), None if el is None else _p(el, _i32p),
                                              None if sp is None else _p(sp, _i32p),
                                              None if nw is None else _p(nw, _u8p)), "pic_set_particles")
        self.nPart = n

    def restart_particles(self, PartData, PartInt):
        """ParticleRestart (particles/restart/particle_restart.f90:134-440) from
        the state file's element-sorted arrays: PartData (n, 7) = x y z vx vy
        vz species, PartInt (nElems, 2) = first and last offset per element.
        The stored element is checked on the device and particles that fail
        the check are located again.  Returns (n_missing, n_lost); the
        particles are numbered in file order."""
        pd = np.ascontiguousarray(PartData, np.float64)
        if pd.ndim != 2 or pd.shape[1] < PART_DATA_SIZE:
            raise ValueError("PartData must have shape (n, >= %d)" % PART_DATA_SIZE)
        pd = np.ascontiguousarray(pd[:, :PART_DATA_SIZE])
        pi = np.ascontiguousarray(PartInt, np.int64)
        if pi.shape != (self.nElems, 2):
            raise ValueError("PartInt must have shape (nElems, 2)")
        miss, lost = ctypes.c_int64(0), ctypes.c_int64(0)
        _chk(load_library().pic_restart_particles(self.h, pd.shape[0], _p(pd, _f64p), _p(pi, _i64p),
                                                  ctypes.byref(miss), ctypes.byref(lost)), "pic_restart_particles")
        self.nPart = pd.shape[0]
        return miss.value, lost.value

    def particle_data(self):
        """FillParticleData (io_hdf5/hdf5_output_particle.f90:1514-1620): the
        particles inside as PartData (n, 7) sorted by element (caller order
        within an element) and PartInt (nElems, 2)."""
        pi = np.zeros((self.nElems, 2), np.int64)
        n = ctypes.c_int64(0)
        pd = np.zeros((max(self.nPart, 1), PART_DATA_SIZE))
        _chk(load_library().pic_get_particle_data(self.h, pd.shape[0], _p(pd, _f64p), _p(pi, _i64p), ctypes.byref(n)),
             "pic_get_particle_data")
        return pd[:n.value].copy(), pi

    # ---------------------------------------------------------------- the stage
    def deposition(self):
        _chk(load_library().pic_deposition(self.h), "pic_deposition")

    def interpolate_field_to_particle(self):
        _chk(load_library().pic_interpolate_field_to_particle(self.h), "pic_interpolate_field_to_particle")

    def push_boris_leapfrog(self, dt):
        _chk(load_library().pic_push_boris_leapfrog(self.h, dt), "pic_push_boris_leapfrog")

    def perform_tracking(self):
        _chk(load_library().pic_perform_tracking(self.h), "pic_perform_tracking")

    def time_step(self, dt, nsteps=1):
        _chk(load_library().pic_time_step(self.h, dt, nsteps), "pic_time_step")

    # reference names
    Deposition = deposition
    InterpolateFieldToParticle = interpolate_field_to_particle
    PerformTracking = perform_tracking
    TimeStepPoissonByBorisLeapfrog = time_step

    # ---------------------------------------------------------------- outputs
    def particles(self):
        ps = np.zeros((self.nPart, 6))
        el = np.zeros(self.nPart, np.int32)
        ins = np.zeros(self.nPart, np.uint8)
        _chk(load_library().pic_get_particles(self.h, _p(ps, _f64p), _p(el, _i32p), _p(ins, _u8p)), "pic_get_particles")
        return ps, el, ins.astype(bool)

    def last_part_pos(self):
        a = np.zeros((self.nPart, 3))
        _chk(load_library().pic_get_last_part_pos(self.h, _p(a, _f64p)), "pic_get_last_part_pos")
        return a

    def field_at_particle(self):
        a = np.zeros((self.nPart, 6))
        _chk(load_library().pic_get_field_at_particle(self.h, _p(a, _f64p)), "pic_get_field_at_particle")
        return a

    def part_pos_ref(self):
        """PartPosRef (RefMapping): reference coordinates in the owner element."""
        a = np.zeros((self.nPart, 3))
        _chk(load_library().pic_get_part_pos_ref(self.h, _p(a, _f64p)), "pic_get_part_pos_ref")
        return a

    def part_source(self):
        a = np.zeros((self.nElems, self.nDof, 4))
        _chk(load_library().pic_get_part_source(self.h, _p(a, _f64p)), "pic_get_part_source")
        return a

    def deposited_charge(self):
        """CalcDepositedCharge: (deposited charge, absolute error, relative error)."""
        a = np.zeros(3)
        _chk(load_library().pic_deposited_charge(self.h, _p(a, _f64p)), "pic_deposited_charge")
        return tuple(a)

    def field_energy(self):
        """CalcPotentialEnergy: eps0/2 * integral |E|^2 of the field last set."""
        a = np.zeros(1)
        _chk(load_library().pic_field_energy(self.h, _p(a, _f64p)), "pic_field_energy")
        return float(a[0])

    CalcDepositedCharge = deposited_charge
    CalcPotentialEnergy = field_energy

    # ---------------------------------------------------------------- self-consistent loop and adjoint
    def set_field_operator(self, G):
        self._G = np.ascontiguousarray(G, np.float64)
        _chk(load_library().pic_set_field_operator(self.h, _p(self._G, _f64p)), "pic_set_field_operator")

    def sc_time_step(self, dt, nsteps, with_loss=True):
        out = np.zeros(1)
        _chk(load_library().pic_sc_time_step(self.h, dt, nsteps, _p(out, _f64p) if with_loss else None),
             "pic_sc_time_step")
        return float(out[0]) if with_loss else None

    def adjoint(self, dt, K, every=10, gradients=True):
        """K self-consistent steps: (loss, dL/dx0, dL/dv0) with respect to the state at entry."""
        loss = np.zeros(1)
        gx = np.zeros((self.nPart, 3)) if gradients else None
        gv = np.zeros((self.nPart, 3)) if gradients else None
        _chk(load_library().pic_adjoint(self.h, dt, K, every, _p(loss, _f64p), None if gx is None else _p(gx, _f64p),
                                        None if gv is None else _p(gv, _f64p)), "pic_adjoint")
        return float(loss[0]), gx, gv

    def set_tuning(self, use_graph=False, sf_kernel=0, sf_persistent=False, sf_shared_loads=False):
        """Execution tuning: graph replay of pic_time_step, shape-function kernel (0 default, 1 VALU, 2 matrix),
        persistent matrix-core shape-function waves, record loads shared by a workgroup's four targets."""
        t = PicTuning(int(bool(use_graph)), int(sf_kernel), int(bool(sf_persistent)), int(bool(sf_shared_loads)))
        _chk(load_library().pic_set_tuning(self.h, ctypes.byref(t)), "pic_set_tuning")

    def stats(self):
        st = PicStats()
        _chk(load_library().pic_get_stats(self.h, ctypes.byref(st)), "pic_get_stats")
        return {k: getattr(st, k) for k, _ in PicStats._fields_}

    def reset_stats(self):
        _chk(load_library().pic_reset_stats(self.h), "pic_reset_stats")

    def synchronize(self):
        _chk(load_library().pic_synchronize(self.h), "pic_synchronize")
