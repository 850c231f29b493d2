"""HOPR-format hexahedral meshes for the particle stage.

The reference reads its mesh from HOPR files (ElemInfo, SideInfo,
NodeCoords, GlobalNodeIDs, BCType; src/mesh/mesh_readin.f90 and the index
constants in src/piclas.h:136-176).  The stage consumes exactly those arrays,
so this module produces them: a periodic/open Cartesian box generator (the
synthetic-plasma workloads) and a loader for arrays that come from a HOPR
file.

Array conventions (identical bytes to the reference's Fortran arrays):
  ElemInfo[nElems][6]  = (type, zone, firstSide, lastSide, firstNode, lastNode)
                         with 0-based offsets: sides firstSide+1..lastSide.
  SideInfo[nSides][5]  = (type, sideID, nbElemID, 10*nbLocSide+flip, BCID)
                         sideID > 0 master, < 0 slave; nbElemID 1-based.
  NodeCoords[nNodes][3], element-local nodes in tensor order (i fastest).
  GlobalNodeIDs[nNodes] 1-based unique node ids.
  BCType[nBCs][4]      = (type, curve, state, alpha); alpha = +-periodic vector id.
"""
from dataclasses import dataclass, field

import numpy as np

# GetCornerNodeMapCGNS (src/mesh/mesh_tools.f90:740-769) for NGeo=1: the CGNS
# corners as tensor indices (i,j,k) and the 4 corners of the 6 local sides.
CGNS_CORNERS = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
SIDE_CORNERS = [(1, 4, 3, 2), (1, 2, 6, 5), (2, 3, 7, 6), (3, 4, 8, 7), (1, 5, 8, 4), (5, 6, 7, 8)]
# local side -> (axis, +1/-1) outward direction
SIDE_DIR = [(2, -1), (1, -1), (0, +1), (1, +1), (0, -1), (2, +1)]

PART_BC_OPEN, PART_BC_REFLECTIVE, PART_BC_PERIODIC = 1, 2, 3


@dataclass
class HoprMesh:
    ElemInfo: np.ndarray
    SideInfo: np.ndarray
    NodeCoords: np.ndarray
    GlobalNodeIDs: np.ndarray
    BCType: np.ndarray
    PartBoundCond: np.ndarray
    PeriodicVectors: np.ndarray
    NGeo: int = 1
    BCNames: list = field(default_factory=list)
    shape: tuple = None          # (nx, ny, nz) for generated Cartesian boxes
    lo: tuple = None
    hi: tuple = None

    @property
    def nElems(self):
        return self.ElemInfo.shape[0]

    def corner_coords(self):
        """[nElems][8][3] element corners in CGNS order (NGeo=1)."""
        nc = self.NodeCoords.reshape(self.nElems, -1, 3)
        idx = [i + 2 * j + 4 * k for (i, j, k) in CGNS_CORNERS]
        return nc[:, idx, :]

    def element_of(self, x):
        """Closed-form owner element (1-based) of points in a generated Cartesian box."""
        nx, ny, nz = self.shape
        lo, hi = np.asarray(self.lo), np.asarray(self.hi)
        h = (hi - lo) / np.array([nx, ny, nz])
        ijk = np.floor((x - lo) / h).astype(np.int64)
        ijk = np.clip(ijk, 0, np.array([nx, ny, nz]) - 1)
        return (self._elem_index(ijk[:, 0], ijk[:, 1], ijk[:, 2]) + 1).astype(np.int32)

    def _elem_index(self, i, j, k):
        nx, ny, nz = self.shape
        return i + nx * (j + ny * k)


def cartesian_box(shape, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), periodic=(True, True, True),
                  wall_bc=PART_BC_OPEN):
    """Periodic (or open) Cartesian box of nx*ny*nz linear hexahedra in HOPR layout.

    Element order is lexicographic (i fastest).  The master side of every
    interior face belongs to the element with the smaller id; a periodic face
    is mastered by its element on the upper boundary.  Boundary ids: per axis
    one BC for the lower face (alpha=+axis: a particle leaving through it is
    shifted by +VV) and one for the upper face (alpha=-axis), as in the
    reference's periodic-box regression meshes.
    """
    nx, ny, nz = shape
    lo = np.asarray(lo, np.float64)
    hi = np.asarray(hi, np.float64)
    n = np.array(shape)
    nE = nx * ny * nz
    # node grid (shared coordinates computed once so equal nodes are bit-equal)
    gx = [lo[a] + (hi[a] - lo[a]) * np.arange(n[a] + 1) / n[a] for a in range(3)]
    for a in range(3):
        gx[a][-1] = hi[a]
    ii, jj, kk = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ii, jj, kk = ii.ravel(order="F"), jj.ravel(order="F"), kk.ravel(order="F")  # i fastest
    e = ii + nx * (jj + ny * kk)
    order = np.argsort(e)
    ii, jj, kk = ii[order], jj[order], kk[order]

    ElemInfo = np.zeros((nE, 6), np.int32)
    ElemInfo[:, 0] = 108
    ElemInfo[:, 1] = 1
    ElemInfo[:, 2] = 6 * np.arange(nE)
    ElemInfo[:, 3] = 6 * np.arange(nE) + 6
    ElemInfo[:, 4] = 8 * np.arange(nE)
    ElemInfo[:, 5] = 8 * np.arange(nE) + 8

    NodeCoords = np.zeros((nE, 8, 3))
    GlobalNodeIDs = np.zeros((nE, 8), np.int32)
    for t in range(8):
        di, dj, dk = t & 1, (t >> 1) & 1, (t >> 2) & 1
        NodeCoords[:, t, 0] = gx[0][ii + di]
        NodeCoords[:, t, 1] = gx[1][jj + dj]
        NodeCoords[:, t, 2] = gx[2][kk + dk]
        GlobalNodeIDs[:, t] = (ii + di) + (nx + 1) * ((jj + dj) + (ny + 1) * (kk + dk)) + 1

    # boundary conditions: per axis (lower face, upper face)
    BCType, PartBC, names, pv = [], [], [], []
    bc_of = {}
    for a, ax in enumerate("xyz"):
        if periodic[a]:
            v = np.zeros(3)
            v[a] = hi[a] - lo[a]
            pv.append(v)
            pid = len(pv)
            bc_of[(a, -1)] = len(BCType) + 1
            BCType.append((1, 0, 0, +pid)); PartBC.append(PART_BC_PERIODIC); names.append("BC_%s+" % ax)
            bc_of[(a, +1)] = len(BCType) + 1
            BCType.append((1, 0, 0, -pid)); PartBC.append(PART_BC_PERIODIC); names.append("BC_%s-" % ax)
        else:
            for sgn, nm in ((-1, "min"), (+1, "max")):
                bc_of[(a, sgn)] = len(BCType) + 1
                BCType.append((4 if wall_bc == PART_BC_REFLECTIVE else 2, 0, 0, 0))
                PartBC.append(wall_bc)
                names.append("BC_%s%s" % (ax, nm))

    # sides
    SideInfo = np.zeros((nE, 6, 5), np.int32)
    SideInfo[:, :, 0] = 4
    ijk = np.stack([ii, jj, kk], 1)
    face_id = {}
    next_id = [0]

    def nb_of(a, sgn):
        t = ijk.copy()
        t[:, a] += sgn
        wrap = (t[:, a] < 0) | (t[:, a] >= n[a])
        t[:, a] %= n[a]
        return t[:, 0] + nx * (t[:, 1] + ny * t[:, 2]), wrap

    # local side pairs across a face: (ls, opposite ls)
    opposite = {0: 5, 5: 0, 1: 3, 3: 1, 2: 4, 4: 2}
    sid = np.zeros((nE, 6), np.int64)
    for ls in range(6):
        a, sgn = SIDE_DIR[ls]
        nb, wrap = nb_of(a, sgn)
        on_bnd = ((ijk[:, a] == 0) & (sgn < 0)) | ((ijk[:, a] == n[a] - 1) & (sgn > 0))
        bcid = bc_of.get((a, sgn), 0) if True else 0
        SideInfo[:, ls, 4] = np.where(on_bnd, bcid, 0)
        if periodic[a]:
            SideInfo[:, ls, 2] = nb + 1
        else:
            SideInfo[:, ls, 2] = np.where(on_bnd, 0, nb + 1)
        # master: interior -> smaller element id; periodic -> element on the upper boundary
        me = np.arange(nE)
        master = np.where(on_bnd, sgn > 0, me < nb)
        if not periodic[a]:
            master = np.where(on_bnd, True, master)
        # face key: (min elem, max elem, axis, wrap) identifies the face
        key_lo = np.where(master, me, nb)
        sid[:, ls] = key_lo * 6 + ls  # temporary per-side key from the master's side
        SideInfo[:, ls, 1] = np.where(master, 1, -1)
        # flip: 10*nbLocSide + (1 + position, in this side's CGNS node order,
        # of the node that coincides with the first node of the side across)
        own = _side_corner_grid(ijk, ls, n, periodic)
        other = _side_corner_grid(np.stack(_unravel(nb, nx, ny), 1), opposite[ls], n, periodic)
        first = other[:, 0, :]
        pos = np.full(nE, -1)
        for q in range(4):
            pos = np.where((pos < 0) & np.all(own[:, q, :] == first, axis=1), q, pos)
        connected = SideInfo[:, ls, 2] > 0
        assert np.all(pos[connected] >= 0)
        SideInfo[:, ls, 3] = np.where(connected, 10 * (opposite[ls] + 1) + pos + 1, 0)
    # global side ids: number master sides consecutively
    is_master = SideInfo[:, :, 1] > 0
    ids = np.zeros((nE, 6), np.int64)
    ids[is_master] = np.arange(1, is_master.sum() + 1)
    # slave gets -id of the master side across the face
    for ls in range(6):
        a, sgn = SIDE_DIR[ls]
        nb, _ = nb_of(a, sgn)
        slave = ~is_master[:, ls]
        ids[slave, ls] = -ids[nb[slave], opposite[ls]]
    SideInfo[:, :, 1] = ids
    # bounded (non-periodic) boundary sides keep flip 0
    mesh = HoprMesh(ElemInfo=ElemInfo, SideInfo=SideInfo.reshape(nE * 6, 5),
                    NodeCoords=NodeCoords.reshape(nE * 8, 3), GlobalNodeIDs=GlobalNodeIDs.reshape(-1),
                    BCType=np.array(BCType, np.int32).reshape(-1, 4), PartBoundCond=np.array(PartBC, np.int32),
                    PeriodicVectors=np.array(pv, np.float64).reshape(-1, 3), NGeo=1, BCNames=names,
                    shape=tuple(shape), lo=tuple(lo), hi=tuple(hi))
    return mesh


def _unravel(e, nx, ny):
    return e % nx, (e // nx) % ny, e // (nx * ny)


def _side_corner_grid(ijk, ls, n, periodic):
    """Grid indices [nE][4][3] of the 4 corners of local side ls (CGNS order), wrapped on periodic axes."""
    out = np.zeros((ijk.shape[0], 4, 3), np.int64)
    for q, c in enumerate(SIDE_CORNERS[ls]):
        off = np.array(CGNS_CORNERS[c - 1])
        g = ijk + off
        for a in range(3):
            if periodic[a]:
                g[:, a] %= n[a]
        out[:, q, :] = g
    return out


def annulus(shape, r0=0.5, r1=1.0, length=1.0, NGeo=2, wall_bc=PART_BC_OPEN):
    """Annular cylinder r0 <= r <= r1, 0 <= z <= length of nr*nt*nz curved
    hexahedra (NGeo >= 1) in HOPR layout.

    Element (i, j, k) spans [r0 + i dr, r0 + (i+1) dr] x [j dt, (j+1) dt] x
    [k dz, (k+1) dz] in (r, theta, z) and is numbered i + nr*(j + nt*k);
    xi, eta, zeta run along r, theta, z, so every Jacobian is positive.  Its
    (NGeo+1)^3 nodes are the images of the equidistant reference points,
    stored in tensor order (i fastest) as HOPR stores curved elements; nodes
    shared by two elements are computed once, so their coordinates are equal
    bit for bit.  The mesh closes on itself in theta (conforming interior
    faces), z is periodic (vector (0,0,length); lower face alpha=+1, upper
    face alpha=-1) and the inner / outer cylinder walls carry wall_bc.
    """
    nr, nt, nz = shape
    n = np.array(shape)
    nE = nr * nt * nz
    g = NGeo
    n1 = g + 1
    # fine node grid: radius, angle, height of every geometry node
    R = r0 + (r1 - r0) * np.arange(g * nr + 1) / (g * nr)
    R[-1] = r1
    T = 2.0 * np.pi * np.arange(g * nt) / (g * nt)
    Z = length * np.arange(g * nz + 1) / (g * nz)
    Z[-1] = length
    cosT, sinT = np.cos(T), np.sin(T)
    e = np.arange(nE)
    ii, jj, kk = e % nr, (e // nr) % nt, e // (nr * nt)

    ElemInfo = np.zeros((nE, 6), np.int32)
    ElemInfo[:, 0] = 208 if g > 1 else 108
    ElemInfo[:, 1] = 1
    ElemInfo[:, 2] = 6 * e
    ElemInfo[:, 3] = 6 * e + 6
    ElemInfo[:, 4] = n1 ** 3 * e
    ElemInfo[:, 5] = n1 ** 3 * (e + 1)

    NodeCoords = np.zeros((nE, n1 ** 3, 3))
    GlobalNodeIDs = np.zeros((nE, n1 ** 3), np.int32)
    for c in range(n1):
        for b in range(n1):
            for a in range(n1):
                q = (c * n1 + b) * n1 + a
                I = g * ii + a
                J = (g * jj + b) % (g * nt)
                K = g * kk + c
                NodeCoords[:, q, 0] = R[I] * cosT[J]
                NodeCoords[:, q, 1] = R[I] * sinT[J]
                NodeCoords[:, q, 2] = Z[K]
                GlobalNodeIDs[:, q] = I + (g * nr + 1) * (J + g * nt * K) + 1

    BCType = [(2 if wall_bc != PART_BC_REFLECTIVE else 4, 0, 0, 0)] * 2 + [(1, 0, 0, +1), (1, 0, 0, -1)]
    PartBC = [wall_bc, wall_bc, PART_BC_PERIODIC, PART_BC_PERIODIC]
    names = ["BC_inner", "BC_outer", "BC_z+", "BC_z-"]
    bc_of = {(0, -1): 1, (0, +1): 2, (2, -1): 3, (2, +1): 4}
    wrap = (False, True, True)
    SideInfo = _structured_sides(np.stack([ii, jj, kk], 1), n, wrap, bc_of, 24 if g > 1 else 4)
    pv = np.array([[0.0, 0.0, length]])
    return HoprMesh(ElemInfo=ElemInfo, SideInfo=SideInfo, NodeCoords=NodeCoords.reshape(-1, 3),
                    GlobalNodeIDs=GlobalNodeIDs.reshape(-1), BCType=np.array(BCType, np.int32),
                    PartBoundCond=np.array(PartBC, np.int32), PeriodicVectors=pv, NGeo=g, BCNames=names)


def _structured_sides(ijk, n, wrap, bc_of, side_type):
    """SideInfo [nE*6][5] of a structured hexahedral grid numbered i + n0*(j + n1*k).

    wrap[a]: the grid closes on itself along axis a (through a periodic BC
    when bc_of has entries for that axis, else through interior faces).
    Master side of an interior face: the element with the smaller id; of a
    periodic face: the element on the upper boundary."""
    nE = ijk.shape[0]
    nx, ny = n[0], n[1]
    me = np.arange(nE)
    SideInfo = np.zeros((nE, 6, 5), np.int32)
    SideInfo[:, :, 0] = side_type
    opposite = {0: 5, 5: 0, 1: 3, 3: 1, 2: 4, 4: 2}

    def nb_of(a, sgn):
        t = ijk.copy()
        t[:, a] += sgn
        t[:, a] %= n[a]
        return t[:, 0] + nx * (t[:, 1] + ny * t[:, 2])

    for ls in range(6):
        a, sgn = SIDE_DIR[ls]
        nb = nb_of(a, sgn)
        on_edge = ((ijk[:, a] == 0) & (sgn < 0)) | ((ijk[:, a] == n[a] - 1) & (sgn > 0))
        bcid = bc_of.get((a, sgn), 0)
        on_bc = on_edge & (bcid > 0)
        SideInfo[:, ls, 4] = np.where(on_bc, bcid, 0)
        connected = ~on_edge | bool(wrap[a])
        SideInfo[:, ls, 2] = np.where(connected, nb + 1, 0)
        master = np.where(on_bc, sgn > 0, me < nb)
        master = np.where(connected, master, True)
        SideInfo[:, ls, 1] = np.where(master, 1, -1)
        own = _side_corner_grid(ijk, ls, n, wrap)
        other = _side_corner_grid(np.stack(_unravel(nb, nx, ny), 1), opposite[ls], n, wrap)
        first = other[:, 0, :]
        pos = np.full(nE, -1)
        for q in range(4):
            pos = np.where((pos < 0) & np.all(own[:, q, :] == first, axis=1), q, pos)
        assert np.all(pos[connected] >= 0)
        SideInfo[:, ls, 3] = np.where(connected, 10 * (opposite[ls] + 1) + pos + 1, 0)
    is_master = SideInfo[:, :, 1] > 0
    ids = np.zeros((nE, 6), np.int64)
    ids[is_master] = np.arange(1, is_master.sum() + 1)
    for ls in range(6):
        a, sgn = SIDE_DIR[ls]
        nb = nb_of(a, sgn)
        slave = ~is_master[:, ls]
        ids[slave, ls] = -ids[nb[slave], opposite[ls]]
    SideInfo[:, :, 1] = ids
    return SideInfo.reshape(nE * 6, 5)


def from_hopr_arrays(ElemInfo, SideInfo, NodeCoords, GlobalNodeIDs, BCType, part_bc_by_name, BCNames,
                     PeriodicVectors, NGeo=1):
    """Mesh from arrays read out of a HOPR file; particle BCs are given per BC name."""
    pbc = np.array([part_bc_by_name[nm] for nm in BCNames], np.int32)
    return HoprMesh(ElemInfo=np.asarray(ElemInfo, np.int32), SideInfo=np.asarray(SideInfo, np.int32),
                    NodeCoords=np.asarray(NodeCoords, np.float64), GlobalNodeIDs=np.asarray(GlobalNodeIDs, np.int32),
                    BCType=np.asarray(BCType, np.int32), PartBoundCond=pbc,
                    PeriodicVectors=np.asarray(PeriodicVectors, np.float64), NGeo=NGeo, BCNames=list(BCNames))
