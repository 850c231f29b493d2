"""Seeded synthetic plasmas of the benchmark configurations.

Configuration 1 (reference-runnable): periodic box [0,4pi]x[0,1]x[0,1],
16x4x4 linear hexahedra, N=3, 64 electrons/cell with density
n0(1+0.01cos(0.5x)), Maxwellian 1 eV, cell_volweight_mean, TriaTracking,
dt = 0.05/omega_pe.
Configuration 2 (throughput): periodic unit cube of 48^3 linear hexahedra,
N=3, 200 electrons/cell uniform, Maxwellian 2 eV, frozen smooth E (8 random
Fourier modes, |E| <= 1e3 V/m) at the Gauss points, uniform B = 0.01 T as
external field, shape_function R = 1.5h, alpha = 2, dt with omega_c dt = 0.2.

Uneven occupancy on curved elements: annular cylinder r in [0.5, 1] m,
length 1 m, of curved (NGeo = 2) hexahedra, N = 3, RefMapping tracking;
electrons (5 eV) and Ar+ ions (39.95 amu, 0.05 eV) with radial density
exp(-(r-0.5)/0.1), so elements hold from a few to several hundred particles;
frozen smooth E as above, B = 0.01 T along z, shape_function deposition, open
walls, periodic in z.

Every array is a pure function of (configuration, seed); the generator is
numpy's PCG64.
"""
from dataclasses import dataclass

import numpy as np

from . import mesh as M
from .stage import ELECTRON_MASS, ELEMENTARY_CHARGE, EPS0, DEPO_CVWM, DEPO_SF, TRACK_TRIA, TRACK_REFMAPPING

KB_EV = ELEMENTARY_CHARGE   # 1 eV in J


@dataclass
class Workload:
    name: str
    mesh: M.HoprMesh
    N: int
    PartState: np.ndarray
    elem: np.ndarray
    species: np.ndarray
    E: np.ndarray
    external_field: tuple
    deposition: int
    sf_radius: float
    sf_alpha: int
    dt: float
    steps: int
    charge: tuple
    mass: tuple
    mpf: tuple
    tracking: int = TRACK_TRIA
    refmapping_guess: int = 1


def gauss_nodes(N):
    """Gauss-Legendre nodes on [-1,1] (for sampling fields at the DOFs)."""
    x, _ = np.polynomial.legendre.leggauss(N + 1)
    return x


def dof_positions(mesh, N):
    """Physical Gauss-point positions [nElems][(N+1)^3][3] of a Cartesian box (i fastest)."""
    c = mesh.corner_coords()               # CGNS corner order
    lo, hi = c[:, 0, :], c[:, 6, :]          # (0,0,0) and (1,1,1) corners
    xg = (gauss_nodes(N) + 1.0) / 2.0
    k, j, i = np.meshgrid(xg, xg, xg, indexing="ij")
    ref = np.stack([i.ravel(), j.ravel(), k.ravel()], 1)   # r = (k*n1+j)*n1+i
    return lo[:, None, :] + ref[None, :, :] * (hi - lo)[:, None, :]


def fourier_field(x, L, nmodes, emax, rng):
    """Sum of nmodes smooth periodic modes sampled at x; |E| <= emax."""
    E = np.zeros_like(x)
    for _ in range(nmodes):
        kvec = rng.integers(-2, 3, size=3)
        if not kvec.any():
            kvec[0] = 1
        amp = rng.normal(size=3)
        amp *= (emax / nmodes) / np.linalg.norm(amp)
        phase = rng.uniform(0, 2 * np.pi)
        arg = 2 * np.pi * (x @ (kvec / np.asarray(L))) + phase
        E += np.sin(arg)[..., None] * amp
    return E


def maxwellian(rng, n, T_eV, m):
    vth = np.sqrt(T_eV * KB_EV / m)
    return rng.normal(scale=vth, size=(n, 3))


def plasma_wave_box(seed=20261203, ppc=64):
    mesh = M.cartesian_box((16, 4, 4), lo=(0.0, 0.0, 0.0), hi=(4 * np.pi, 1.0, 1.0))
    N = 3
    rng = np.random.Generator(np.random.PCG64(seed))
    n = 16 * 4 * 4 * ppc
    Lx = 4 * np.pi
    # inverse-CDF sampling of n0(1+0.01cos(0.5x)) on [0,Lx]: F(x) = x + 0.02 sin(0.5x)
    u = rng.uniform(0, Lx, size=n)
    x = u.copy()
    for _ in range(30):
        x -= (x + 0.02 * np.sin(0.5 * x) - u) / (1 + 0.01 * np.cos(0.5 * x))
    pos = np.stack([np.clip(x, 0, np.nextafter(Lx, 0)), rng.uniform(0, 1, n), rng.uniform(0, 1, n)], 1)
    vel = maxwellian(rng, n, 1.0, ELECTRON_MASS)
    n0 = 1e12
    mpf = n0 * Lx / n
    wpe = np.sqrt(n0 * ELEMENTARY_CHARGE ** 2 / (EPS0 * ELECTRON_MASS))
    dt = 0.05 / wpe
    # frozen field of the perturbation: Gauss's law for the 1% cos mode
    xg = dof_positions(mesh, N)
    E = np.zeros_like(xg)
    E[..., 0] = -ELEMENTARY_CHARGE * n0 * 0.01 / (EPS0 * 0.5) * np.sin(0.5 * xg[..., 0])
    return Workload("plasma_wave_box16x4x4_N3_cvwm", mesh, N, np.hstack([pos, vel]), mesh.element_of(pos),
                    np.ones(n, np.int32), E, (0, 0, 0, 0, 0, 0), DEPO_CVWM, 0.0, 2, dt, 200,
                    (-ELEMENTARY_CHARGE,), (ELECTRON_MASS,), (mpf,))


def thermal_cube(nx=48, ppc=200, seed=20261204, deposition=DEPO_SF, steps=100):
    mesh = M.cartesian_box((nx, nx, nx), lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0))
    N = 3
    rng = np.random.Generator(np.random.PCG64(seed))
    nE = nx ** 3
    n = nE * ppc
    pos = rng.uniform(0.0, 1.0, size=(n, 3))
    vel = maxwellian(rng, n, 2.0, ELECTRON_MASS)
    B = 0.01
    wc = ELEMENTARY_CHARGE * B / ELECTRON_MASS
    dt = 0.2 / wc
    n0 = 1e12
    mpf = n0 * 1.0 / n
    xg = dof_positions(mesh, N)
    E = fourier_field(xg.reshape(-1, 3), (1.0, 1.0, 1.0), 8, 1e3, rng).reshape(xg.shape)
    h = 1.0 / nx
    return Workload("thermal_cube%d_N3_sf" % nx, mesh, N, np.hstack([pos, vel]), mesh.element_of(pos),
                    np.ones(n, np.int32), E, (0, 0, 0, 0, 0, B), deposition, 1.5 * h, 2, dt, steps,
                    (-ELEMENTARY_CHARGE,), (ELECTRON_MASS,), (mpf,))


def curved_dof_positions(mesh, N):
    """Physical Gauss-point positions [nElems][(N+1)^3][3] of a mesh whose
    element nodes are equidistant in reference space (tensor order)."""
    g = mesh.NGeo
    nodes = mesh.NodeCoords.reshape(mesh.nElems, g + 1, g + 1, g + 1, 3)   # [e][k][j][i][c]
    xe = np.linspace(-1.0, 1.0, g + 1)
    xg = gauss_nodes(N)
    L = np.ones((N + 1, g + 1))
    for a in range(g + 1):
        for b in range(g + 1):
            if a != b:
                L[:, a] *= (xg - xe[b]) / (xe[a] - xe[b])
    x = np.einsum("ri,sj,tk,ekjic->etsrc", L, L, L, nodes)      # [e][t=k][s=j][r=i][c]
    return x.reshape(mesh.nElems, (N + 1) ** 3, 3)


AMU = 1.66053906660e-27


def annulus_plasma(shape=(24, 96, 24), n_total=30_000_000, seed=20261205, sf_radius=0.03, steps=100, dt=2e-10,
                   NGeo=2, deposition=DEPO_SF):
    mesh = M.annulus(shape, 0.5, 1.0, 1.0, NGeo=NGeo)
    N = 3
    rng = np.random.Generator(np.random.PCG64(seed))
    # r with density exp(-(r-0.5)/0.1) in the annulus (pdf ~ r exp(..)), by a tabulated inverse CDF
    rr = np.linspace(0.5, 1.0, 200001)
    pdf = rr * np.exp(-(rr - 0.5) / 0.1)
    cdf = np.concatenate([[0.0], np.cumsum(0.5 * (pdf[1:] + pdf[:-1]) * np.diff(rr))])
    cdf /= cdf[-1]
    r = np.interp(rng.uniform(0.0, 1.0, n_total), cdf, rr)
    th = rng.uniform(0.0, 2.0 * np.pi, n_total)
    z = rng.uniform(0.0, 1.0, n_total)
    pos = np.stack([r * np.cos(th), r * np.sin(th), z], 1)
    ne = n_total // 2
    species = np.ones(n_total, np.int32)
    species[ne:] = 2
    m_ar = 39.95 * AMU
    vel = np.concatenate([maxwellian(rng, ne, 5.0, ELECTRON_MASS), maxwellian(rng, n_total - ne, 0.05, m_ar)])
    vol = np.pi * (1.0 - 0.25) * 1.0
    mpf = 1e12 * vol / n_total
    xg = curved_dof_positions(mesh, N)
    E = fourier_field(xg.reshape(-1, 3), (2.0, 2.0, 1.0), 8, 1e3, rng).reshape(xg.shape)
    return Workload("annulus%dx%dx%d_NGeo%d_N3_refmap" % (shape + (NGeo,)), mesh, N, np.hstack([pos, vel]),
                    np.zeros(n_total, np.int32), species, E, (0, 0, 0, 0, 0, 0.01), deposition, sf_radius, 2, dt,
                    steps, (-ELEMENTARY_CHARGE, ELEMENTARY_CHARGE), (ELECTRON_MASS, m_ar), (mpf, mpf),
                    tracking=TRACK_REFMAPPING, refmapping_guess=2 if NGeo > 1 else 1)


def stage_kwargs(w: Workload):
    return dict(deposition=w.deposition, sf_radius=w.sf_radius, sf_alpha=w.sf_alpha,
                external_field=w.external_field, charge=w.charge, mass=w.mass, mpf=w.mpf,
                tracking=w.tracking, refmapping_guess=w.refmapping_guess)
