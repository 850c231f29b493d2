"""The differentiable workload: a perturbed periodic plasma whose field is
closed by a fixed linear (spectral, periodic) Poisson operator.

  G = periodic_poisson_operator(mesh)   E at the unique corner nodes from the
                                        element means of rho
  w, dxda = landau_box(...)             16^3 box, electrons with density
                                        n0 (1 + alpha cos(k x)), and d x0 / d alpha

The stage's pic_adjoint gives dL/dx0 and dL/dv0 of the final field energy;
dL/dalpha = sum_p dL/dx0_p . d x0_p / d alpha.
"""
import numpy as np

from . import mesh as M
from . import workloads as W
from .stage import ELECTRON_MASS, ELEMENTARY_CHARGE, EPS0, DEPO_SF


def periodic_poisson_operator(mesh):
    """Dense G[3*nUnique][nElems] of a generated periodic Cartesian box:
    -lap phi = rho/eps0 solved spectrally on the element grid (element means
    at the cell centres), E = -grad phi evaluated spectrally at the cell
    corners.  Row 3*(GlobalNodeID-1)+c holds E_c at that node."""
    nx, ny, nz = mesh.shape
    lo, hi = np.asarray(mesh.lo), np.asarray(mesh.hi)
    Lx, Ly, Lz = hi - lo
    hx, hy, hz = Lx / nx, Ly / ny, Lz / nz
    kx = 2 * np.pi * np.fft.fftfreq(nx, d=hx)
    ky = 2 * np.pi * np.fft.fftfreq(ny, d=hy)
    kz = 2 * np.pi * np.fft.fftfreq(nz, d=hz)
    KZ, KY, KX = np.meshgrid(kz, ky, kx, indexing="ij")
    K2 = KX ** 2 + KY ** 2 + KZ ** 2
    K2[0, 0, 0] = 1.0
    unit = np.zeros((nz, ny, nx))
    unit[0, 0, 0] = 1.0
    phih = np.fft.fftn(unit) / (EPS0 * K2)
    phih[0, 0, 0] = 0.0
    # corner (i,j,k) lies half a cell below the centre of element (i,j,k)
    shift = np.exp(-1j * (KX * hx / 2 + KY * hy / 2 + KZ * hz / 2))
    E0 = [np.fft.ifftn(-1j * Kc * phih * shift).real for Kc in (KX, KY, KZ)]   # [k][j][i], unit charge at element 0
    # unique nodes of the generated box: id = I + (nx+1)(J + (ny+1)K), I in 0..nx
    nu = (nx + 1) * (ny + 1) * (nz + 1)
    I, J, Kk = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    uid = (I + (nx + 1) * (J + (ny + 1) * Kk)).ravel()
    ci, cj, ck = (I.ravel() % nx), (J.ravel() % ny), (Kk.ravel() % nz)
    e = np.arange(nx * ny * nz)
    ea, eb, ec = e % nx, (e // nx) % ny, e // (nx * ny)
    G = np.zeros((3 * nu, nx * ny * nz))
    for c in range(3):
        G[3 * uid + c, :] = E0[c][(ck[:, None] - ec[None, :]) % nz, (cj[:, None] - eb[None, :]) % ny,
                                  (ci[:, None] - ea[None, :]) % nx]
    return G


def landau_box(n_el=16, n_part=2_000_000, alpha=0.01, T_eV=8.0, seed=20261206, sf_radius_h=1.5, sf_alpha=3,
               wpe_dt=0.1, steps=50):
    """Periodic unit cube, n_el^3 linear hexahedra, N = 3, electrons with
    density n0 (1 + alpha cos(2 pi x)) (inverse-CDF sampling), Maxwellian T_eV;
    shape_function deposition.  Returns (workload, d x0 / d alpha)."""
    mesh = M.cartesian_box((n_el, n_el, n_el))
    N = 3
    rng = np.random.Generator(np.random.PCG64(seed))
    k = 2 * np.pi
    u = rng.uniform(0.0, 1.0, n_part)
    x = u.copy()
    for _ in range(40):   # x + alpha/k sin(k x) = u
        x -= (x + alpha / k * np.sin(k * x) - u) / (1 + alpha * np.cos(k * x))
    dxda = -(np.sin(k * x) / k) / (1 + alpha * np.cos(k * x))
    pos = np.stack([x, rng.uniform(0, 1, n_part), rng.uniform(0, 1, n_part)], 1)
    vel = W.maxwellian(rng, n_part, T_eV, ELECTRON_MASS)
    n0 = 1e12
    mpf = n0 / n_part
    wpe = np.sqrt(n0 * ELEMENTARY_CHARGE ** 2 / (EPS0 * ELECTRON_MASS))
    w = W.Workload("landau_box%d_N3_sf_adjoint" % n_el, mesh, N, np.hstack([pos, vel]), mesh.element_of(pos),
                   np.ones(n_part, np.int32), np.zeros((mesh.nElems, (N + 1) ** 3, 3)), (0, 0, 0, 0, 0, 0), DEPO_SF,
                   sf_radius_h / n_el, sf_alpha, wpe_dt / wpe, steps, (-ELEMENTARY_CHARGE,), (ELECTRON_MASS,), (mpf,))
    dx0 = np.zeros((n_part, 3))
    dx0[:, 0] = dxda
    return w, dx0


def stage_loss(s, dt, K, every=10, species=None, locate=None):
    """The loss of K self-consistent steps of stage ``s`` (whose field operator
    is set) as a function of the particle state, for torch.autograd:

        f = stage_loss(s, dt, K, species=w.species)
        L = f(x0, v0)            # x0, v0: (nPart, 3) float64 tensors
        L.backward()             # x0.grad, v0.grad from pic_adjoint

    Every call sets the particles from (x0, v0) and runs pic_adjoint forward
    and backward; ``locate(x) -> 1-based element ids`` gives the elements,
    else the stage locates the particles on the device."""
    import torch

    class _StageLoss(torch.autograd.Function):
        @staticmethod
        def forward(ctx, x0, v0):
            ps = torch.cat([x0, v0], 1).detach().cpu().numpy()
            elem = None if locate is None else locate(ps[:, :3])
            s.set_particles(ps, elem=elem, species=species)
            loss, gx, gv = s.adjoint(dt, K, every)
            ctx.save_for_backward(torch.from_numpy(gx).to(x0.device), torch.from_numpy(gv).to(v0.device))
            return torch.tensor(loss, dtype=torch.float64, device=x0.device)

        @staticmethod
        def backward(ctx, g):
            gx, gv = ctx.saved_tensors
            return g * gx, g * gv

    return _StageLoss.apply
