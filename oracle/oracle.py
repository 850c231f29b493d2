"""ctypes binding of the CPU oracle (libpic_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() to
check its result, and by bench.py's cpu_baseline leg.  The product never imports
this module.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpic_oracle.so")

TRACK_REFMAPPING, TRACK_TRIA = 1, 3
DEPO_NONE, DEPO_CVWM, DEPO_SF = 0, 1, 2

_i32p = ctypes.POINTER(ctypes.c_int32)
_f64p = ctypes.POINTER(ctypes.c_double)
_u8p = ctypes.POINTER(ctypes.c_uint8)


class MeshDesc(ctypes.Structure):
    _fields_ = [("nElems", ctypes.c_int32), ("nSides", ctypes.c_int32), ("nNodes", ctypes.c_int32),
                ("NGeo", ctypes.c_int32), ("N", ctypes.c_int32),
                ("ElemInfo", _i32p), ("SideInfo", _i32p), ("NodeCoords", _f64p), ("GlobalNodeIDs", _i32p),
                ("nBCs", ctypes.c_int32), ("BCType", _i32p), ("PartBoundCond", _i32p),
                ("nPeriodicVectors", ctypes.c_int32), ("PeriodicVectors", _f64p)]


class Params(ctypes.Structure):
    _fields_ = [("tracking", ctypes.c_int32), ("deposition", ctypes.c_int32),
                ("sf_radius", ctypes.c_double), ("sf_alpha", ctypes.c_int32),
                ("refmapping_guess", ctypes.c_int32), ("refmapping_eps", ctypes.c_double),
                ("external_field", ctypes.c_double * 6), ("nSpecies", ctypes.c_int32),
                ("charge", ctypes.c_double * 8), ("mass", ctypes.c_double * 8), ("mpf", ctypes.c_double * 8)]


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.om_create.restype = ctypes.c_void_p
        L.om_create.argtypes = [ctypes.POINTER(MeshDesc), ctypes.POINTER(Params)]
        L.om_destroy.argtypes = [ctypes.c_void_p]
        L.om_last_error.restype = ctypes.c_char_p
        for name in ("om_deposition", "om_interpolate_field_to_particle", "om_perform_tracking",
                     "om_locate_particles"):
            getattr(L, name).argtypes = [ctypes.c_void_p]
        L.om_push_boris_leapfrog.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.om_time_step.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int32]
        L.om_set_threads.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.om_set_field.argtypes = [ctypes.c_void_p, _f64p]
        L.om_set_particles.argtypes = [ctypes.c_void_p, ctypes.c_int32, _f64p, _i32p, _i32p, _u8p]
        L.om_get_particles.argtypes = [ctypes.c_void_p, _f64p, _i32p, _u8p]
        for name in ("om_get_last_part_pos", "om_get_field_at_particle", "om_get_part_source", "om_get_elem_xgp",
                     "om_get_sj", "om_get_node_volume"):
            getattr(L, name).argtypes = [ctypes.c_void_p, _f64p]
        L.om_get_elem_side_node_id.argtypes = [ctypes.c_void_p, _i32p]
        L.om_get_concave.argtypes = [ctypes.c_void_p, _u8p]
        L.om_get_elem_bary.argtypes = [ctypes.c_void_p, _f64p, _f64p]
        L.om_get_xieta_basis.argtypes = [ctypes.c_void_p, _f64p, _f64p]
        L.om_get_part_pos_ref.argtypes = [ctypes.c_void_p, _f64p]
        L.om_sc_set_operator.argtypes = [ctypes.c_void_p, _f64p]
        L.om_sc_field_from_source.argtypes = [ctypes.c_void_p]
        L.om_sc_forward.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int32, _f64p]
        L.om_get_refmapping_tables.argtypes = [ctypes.c_void_p, _u8p, _f64p]
        L.om_position_in_ref_elem.argtypes = [ctypes.c_void_p, ctypes.c_int32, _f64p, _i32p, _f64p, _u8p]
        L.om_eval_field_at_ref_pos.argtypes = [ctypes.c_void_p, ctypes.c_int32, _f64p, _i32p, _f64p]
        L.om_em_field_dw.argtypes = [ctypes.c_void_p, ctypes.c_int32, _f64p, _i32p, _f64p]
        L.om_get_n_bc_sides.restype = ctypes.c_int32
        L.om_get_n_bc_sides.argtypes = [ctypes.c_void_p]
        L.om_get_bc_sides.argtypes = [ctypes.c_void_p, _i32p, _u8p, _f64p, _f64p, _f64p]
        L.om_get_bc_list_size.restype = ctypes.c_int32
        L.om_get_bc_list_size.argtypes = [ctypes.c_void_p]
        L.om_get_bc_elem_lists.argtypes = [ctypes.c_void_p, _i32p, _i32p, _f64p]
        L.om_bc_side_intersection.argtypes = [ctypes.c_void_p, ctypes.c_int32, _f64p, _f64p, _f64p]
        L.om_set_part_source.argtypes = [ctypes.c_void_p, _f64p]
        L.om_get_field.argtypes = [ctypes.c_void_p, _f64p]
        L.om_legendre_gauss_nodes.argtypes = [ctypes.c_int32, _f64p, _f64p]
        L.om_cheby_gauss_lobatto_nodes.argtypes = [ctypes.c_int32, _f64p]
        L.om_barycentric_weights.argtypes = [ctypes.c_int32, _f64p, _f64p]
        L.om_lagrange_polys.argtypes = [ctypes.c_double, ctypes.c_int32, _f64p, _f64p, _f64p]
        L.om_get_nparticles.restype = ctypes.c_int32
        L.om_get_nparticles.argtypes = [ctypes.c_void_p]
        L.om_get_private_reduce_seconds.restype = ctypes.c_double
        L.om_get_private_reduce_seconds.argtypes = [ctypes.c_void_p]
        L.om_get_lost_count.restype = ctypes.c_int64
        L.om_get_lost_count.argtypes = [ctypes.c_void_p]
        L.om_get_multi_crossings.restype = ctypes.c_int64
        L.om_get_multi_crossings.argtypes = [ctypes.c_void_p]
        L.om_plant_mistake.restype = None
        L.om_plant_mistake.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.om_set_tiebreak_log.argtypes = [ctypes.c_void_p, ctypes.c_int64]
        L.om_get_tiebreak_log.restype = ctypes.c_int64
        L.om_get_tiebreak_log.argtypes = [ctypes.c_void_p, _f64p]
        L.om_get_n_unique_nodes.restype = ctypes.c_int32
        L.om_get_n_unique_nodes.argtypes = [ctypes.c_void_p]
        for name in ("om_deposited_charge", "om_particle_charge", "om_field_energy"):
            getattr(L, name).restype = ctypes.c_double
            getattr(L, name).argtypes = [ctypes.c_void_p]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


class Oracle:
    """One oracle instance over a HOPR mesh (see piclas_amd.mesh.HoprMesh)."""

    def __init__(self, mesh, N, tracking=TRACK_TRIA, deposition=DEPO_NONE, sf_radius=0.0, sf_alpha=2,
                 refmapping_guess=1, refmapping_eps=1e-4, external_field=(0, 0, 0, 0, 0, 0),
                 charge=(-1.602176634e-19,), mass=(9.1093826e-31,), mpf=(1.0,)):
        L = lib()
        self._keep = [np.ascontiguousarray(mesh.ElemInfo, np.int32), np.ascontiguousarray(mesh.SideInfo, np.int32),
                      np.ascontiguousarray(mesh.NodeCoords, np.float64),
                      np.ascontiguousarray(mesh.GlobalNodeIDs, np.int32),
                      np.ascontiguousarray(mesh.BCType, np.int32), np.ascontiguousarray(mesh.PartBoundCond, np.int32),
                      np.ascontiguousarray(mesh.PeriodicVectors, np.float64).reshape(-1, 3)]
        ei, si, nc, gid, bct, pbc, pv = self._keep
        md = MeshDesc(ei.shape[0], si.shape[0], nc.shape[0], mesh.NGeo, N, _p(ei, _i32p), _p(si, _i32p),
                      _p(nc, _f64p), _p(gid, _i32p), bct.shape[0], _p(bct, _i32p), _p(pbc, _i32p),
                      pv.shape[0], _p(pv, _f64p))
        par = Params()
        par.tracking, par.deposition = tracking, deposition
        par.sf_radius, par.sf_alpha = sf_radius, sf_alpha
        par.refmapping_guess, par.refmapping_eps = refmapping_guess, refmapping_eps
        for i in range(6):
            par.external_field[i] = external_field[i]
        par.nSpecies = len(charge)
        for i in range(len(charge)):
            par.charge[i], par.mass[i], par.mpf[i] = charge[i], mass[i], mpf[i]
        self.h = L.om_create(ctypes.byref(md), ctypes.byref(par))
        if not self.h:
            raise RuntimeError("om_create: " + L.om_last_error().decode())
        self.nElems, self.N = ei.shape[0], N
        self.NGeo = mesh.NGeo
        self.nDof = (N + 1) ** 3

    def __del__(self):
        if getattr(self, "h", None):
            lib().om_destroy(self.h)
            self.h = None

    def _check(self, rc, what):
        if rc != 0:
            raise RuntimeError("%s: %s" % (what, lib().om_last_error().decode()))

    def set_threads(self, n):
        lib().om_set_threads(self.h, n)

    def set_field(self, E):
        E = np.ascontiguousarray(E, np.float64)
        assert E.size == self.nElems * self.nDof * 3
        self._E = E
        self._check(lib().om_set_field(self.h, _p(E, _f64p)), "set_field")

    def set_particles(self, PartState, elem=None, species=None, is_new=None):
        ps = np.ascontiguousarray(PartState, np.float64)
        n = ps.shape[0]
        el = np.ascontiguousarray(elem if elem is not None else np.zeros(n), np.int32)
        sp = np.ascontiguousarray(species if species is not None else np.ones(n), np.int32)
        nw = np.ascontiguousarray(is_new if is_new is not None else np.zeros(n), np.uint8)
        self._check(lib().om_set_particles(self.h, n, _p(ps, _f64p), _p(el, _i32p), _p(sp, _i32p), _p(nw, _u8p)),
                    "set_particles")

    def locate(self):
        return lib().om_locate_particles(self.h)

    def deposition(self):
        self._check(lib().om_deposition(self.h), "deposition")

    def interpolate(self):
        self._check(lib().om_interpolate_field_to_particle(self.h), "interpolate")

    def push(self, dt):
        self._check(lib().om_push_boris_leapfrog(self.h, dt), "push")

    def track(self):
        self._check(lib().om_perform_tracking(self.h), "tracking")

    def time_step(self, dt, nsteps=1):
        self._check(lib().om_time_step(self.h, dt, nsteps), "time_step")

    def particles(self):
        n = lib().om_get_nparticles(self.h)
        ps = np.zeros((n, 6))
        el = np.zeros(n, np.int32)
        ins = np.zeros(n, np.uint8)
        lib().om_get_particles(self.h, _p(ps, _f64p), _p(el, _i32p), _p(ins, _u8p))
        return ps, el, ins.astype(bool)

    def last_part_pos(self):
        n = lib().om_get_nparticles(self.h)
        a = np.zeros((n, 3))
        lib().om_get_last_part_pos(self.h, _p(a, _f64p))
        return a

    def field_at_particle(self):
        n = lib().om_get_nparticles(self.h)
        a = np.zeros((n, 6))
        lib().om_get_field_at_particle(self.h, _p(a, _f64p))
        return a

    def part_source(self):
        a = np.zeros((self.nElems, self.nDof, 4))
        lib().om_get_part_source(self.h, _p(a, _f64p))
        return a

    def private_reduce_seconds(self):
        """Threaded shape-function deposit: seconds spent zeroing and summing
        the per-thread source copies (summed over calls)."""
        return lib().om_get_private_reduce_seconds(self.h)

    def lost(self):
        return lib().om_get_lost_count(self.h)

    def multi_crossings(self):
        return lib().om_get_multi_crossings(self.h)

    def tiebreak_log(self, cap=None):
        """With cap: start recording up to cap multi-crossing events of the
        TriaTracking walk.  Without: the records so far, shape (n, 29)."""
        if cap is not None:
            lib().om_set_tiebreak_log(self.h, int(cap))
            return None
        n = lib().om_get_tiebreak_log(self.h, None)
        a = np.zeros((max(n, 1), 29))
        lib().om_get_tiebreak_log(self.h, _p(a, _f64p))
        return a[:n]

    def plant_mistake(self, which):
        """Test only: 1 makes the TriaTracking tie-break take the side crossed last."""
        lib().om_plant_mistake(self.h, which)

    def elem_xgp(self):
        a = np.zeros((self.nElems, self.nDof, 3))
        lib().om_get_elem_xgp(self.h, _p(a, _f64p))
        return a

    def sj(self):
        a = np.zeros((self.nElems, self.nDof))
        lib().om_get_sj(self.h, _p(a, _f64p))
        return a

    def elem_side_node_id(self):
        a = np.zeros((self.nElems, 6, 4), np.int32)
        lib().om_get_elem_side_node_id(self.h, _p(a, _i32p))
        return a

    def concave(self):
        a = np.zeros((self.nElems, 6), np.uint8)
        lib().om_get_concave(self.h, _p(a, _u8p))
        return a

    def node_volume(self):
        a = np.zeros(lib().om_get_n_unique_nodes(self.h))
        lib().om_get_node_volume(self.h, _p(a, _f64p))
        return a

    def elem_bary(self):
        b = np.zeros((self.nElems, 3))
        r = np.zeros(self.nElems)
        lib().om_get_elem_bary(self.h, _p(b, _f64p), _p(r, _f64p))
        return b, r

    def xieta_basis(self):
        b = np.zeros((self.nElems, 6, 3))
        s = np.zeros((self.nElems, 6))
        lib().om_get_xieta_basis(self.h, _p(b, _f64p), _p(s, _f64p))
        return b, s

    def sc_set_operator(self, G):
        self._G = np.ascontiguousarray(G, np.float64)
        self._check(lib().om_sc_set_operator(self.h, _p(self._G, _f64p)), "sc_set_operator")

    def set_part_source(self, PS):
        """Test only: overwrite PartSource [nElems][(N+1)^3][4]."""
        PS = np.ascontiguousarray(PS, np.float64)
        assert PS.size == self.nElems * self.nDof * 4
        self._check(lib().om_set_part_source(self.h, _p(PS, _f64p)), "set_part_source")

    def field(self):
        """E at the Gauss points [nElems][(N+1)^3][3] as the oracle holds it."""
        E = np.zeros((self.nElems, self.nDof, 3))
        lib().om_get_field(self.h, _p(E, _f64p))
        return E

    def sc_field_from_source(self):
        self._check(lib().om_sc_field_from_source(self.h), "sc_field_from_source")

    def sc_forward(self, dt, K):
        """K self-consistent steps through the field operator; returns the final field energy."""
        out = np.zeros(1)
        self._check(lib().om_sc_forward(self.h, dt, K, _p(out, _f64p)), "sc_forward")
        return float(out[0])

    def part_pos_ref(self):
        n = lib().om_get_nparticles(self.h)
        a = np.zeros((n, 3))
        lib().om_get_part_pos_ref(self.h, _p(a, _f64p))
        return a

    def refmapping_tables(self):
        """(ElemCurved, ElemEpsOneCell) per element."""
        c = np.zeros(self.nElems, np.uint8)
        e = np.zeros(self.nElems)
        lib().om_get_refmapping_tables(self.h, _p(c, _u8p), _p(e, _f64p))
        return c.astype(bool), e

    def position_in_ref_elem(self, x, elem):
        x = np.ascontiguousarray(x, np.float64)
        elem = np.ascontiguousarray(elem, np.int32)
        n = x.shape[0]
        xi = np.zeros((n, 3))
        ok = np.zeros(n, np.uint8)
        lib().om_position_in_ref_elem(self.h, n, _p(x, _f64p), _p(elem, _i32p), _p(xi, _f64p), _p(ok, _u8p))
        return xi, ok.astype(bool)

    def eval_field_at_ref_pos(self, xi, elem):
        xi = np.ascontiguousarray(xi, np.float64)
        elem = np.ascontiguousarray(elem, np.int32)
        n = xi.shape[0]
        U = np.zeros((n, 3))
        lib().om_eval_field_at_ref_pos(self.h, n, _p(xi, _f64p), _p(elem, _i32p), _p(U, _f64p))
        return U

    def em_field_dw(self, x, elem):
        """GetEMFieldDW (E part) at physical positions x in elements elem (1-based)."""
        x = np.ascontiguousarray(x, np.float64)
        elem = np.ascontiguousarray(elem, np.int32)
        n = x.shape[0]
        U = np.zeros((n, 3))
        lib().om_em_field_dw(self.h, n, _p(x, _f64p), _p(elem, _i32p), _p(U, _f64p))
        return U

    def deposited_charge(self):
        return lib().om_deposited_charge(self.h)

    def particle_charge(self):
        return lib().om_particle_charge(self.h)

    def field_energy(self):
        return lib().om_field_energy(self.h)

    # --- RefMapping boundary tracking tables (meshes with non-periodic boundaries)
    def bc_sides(self):
        """(side rows, SideType, Bezier control points [nBC][g1][g1][3], origin, radius)"""
        n = lib().om_get_n_bc_sides(self.h)
        g1 = self.NGeo + 1
        side = np.zeros(n, np.int32)
        typ = np.zeros(n, np.uint8)
        bez = np.zeros((n, g1, g1, 3))
        org = np.zeros((n, 3))
        rad = np.zeros(n)
        lib().om_get_bc_sides(self.h, _p(side, _i32p), _p(typ, _u8p), _p(bez, _f64p), _p(org, _f64p), _p(rad, _f64p))
        return side, typ, bez, org, rad

    def bc_elem_lists(self):
        m = lib().om_get_bc_list_size(self.h)
        off = np.zeros(self.nElems + 1, np.int32)
        bs = np.zeros(m, np.int32)
        dist = np.zeros(m)
        lib().om_get_bc_elem_lists(self.h, _p(off, _i32p), _p(bs, _i32p), _p(dist, _f64p))
        return off, bs, dist

    def bc_side_intersection(self, b, last, x):
        """(alpha, xi, eta) of the segment last -> x with boundary side b; alpha = -1: no hit"""
        last = np.ascontiguousarray(last, np.float64)
        x = np.ascontiguousarray(x, np.float64)
        out = np.zeros(3)
        rc = lib().om_bc_side_intersection(self.h, int(b), _p(last, _f64p), _p(x, _f64p), _p(out, _f64p))
        if rc != 0:
            raise RuntimeError(lib().om_last_error().decode())
        return out


def legendre_gauss_nodes(N):
    x = np.zeros(N + 1)
    w = np.zeros(N + 1)
    lib().om_legendre_gauss_nodes(N, _p(x, _f64p), _p(w, _f64p))
    return x, w


def cheby_gauss_lobatto_nodes(N):
    x = np.zeros(N + 1)
    lib().om_cheby_gauss_lobatto_nodes(N, _p(x, _f64p))
    return x


def barycentric_weights(x):
    x = np.ascontiguousarray(x, np.float64)
    w = np.zeros_like(x)
    lib().om_barycentric_weights(len(x) - 1, _p(x, _f64p), _p(w, _f64p))
    return w


def lagrange_polys(xv, nodes, wb):
    nodes = np.ascontiguousarray(nodes, np.float64)
    wb = np.ascontiguousarray(wb, np.float64)
    L = np.zeros_like(nodes)
    lib().om_lagrange_polys(float(xv), len(nodes) - 1, _p(nodes, _f64p), _p(wb, _f64p), _p(L, _f64p))
    return L
