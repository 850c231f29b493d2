"""The fixed linear field operator of the differentiable workload, and the
oracle's self-consistent loop through it (CPU)."""
import numpy as np

from piclas_amd import adjoint as A
from piclas_amd import mesh as M
from piclas_amd.stage import EPS0


def test_poisson_operator_single_mode():
    """A cosine charge mode gives E_x = rho0/(eps0 k) sin(k x) at the corners."""
    n = 8
    m = M.cartesian_box((n, n, n))
    G = A.periodic_poisson_operator(m)
    nu = (n + 1) ** 3
    assert G.shape == (3 * nu, n ** 3)
    e = np.arange(n ** 3)
    xc = (e % n + 0.5) / n
    rho0, k = 1e-9, 2 * np.pi
    E = (G @ (rho0 * np.cos(k * xc))).reshape(nu, 3)
    I = np.arange(nu) % (n + 1)
    xn = I / n
    ref = rho0 / (EPS0 * k) * np.sin(k * xn)
    assert np.allclose(E[:, 0], ref, rtol=0, atol=1e-12 * np.abs(ref).max())
    assert np.abs(E[:, 1:]).max() <= 1e-12 * np.abs(ref).max()
    # a uniform charge has no field (the k = 0 mode is dropped)
    assert np.abs(G @ np.ones(n ** 3)).max() <= 1e-12 * np.abs(G).max() * n ** 3


def test_oracle_self_consistent_loss_is_smooth_in_alpha(oracle_mod):
    """Central differences of the oracle's final field energy in alpha agree
    for two step sizes (the loss is smooth there), and the loss grows with
    the perturbation amplitude at t = 0."""
    def loss(alpha, K):
        w, _ = A.landau_box(n_el=4, n_part=600, alpha=alpha, steps=K)
        o = oracle_mod.Oracle(w.mesh, w.N, deposition=w.deposition, sf_radius=w.sf_radius, sf_alpha=w.sf_alpha,
                              charge=w.charge, mass=w.mass, mpf=w.mpf)
        o.sc_set_operator(A.periodic_poisson_operator(w.mesh))
        o.set_field(w.E)
        o.set_particles(w.PartState, elem=w.elem, species=w.species)
        return o.sc_forward(w.dt, K)
    assert loss(0.05, 0) > loss(0.01, 0)
    d1 = (loss(0.01 + 1e-5, 5) - loss(0.01 - 1e-5, 5)) / 2e-5
    d2 = (loss(0.01 + 1e-6, 5) - loss(0.01 - 1e-6, 5)) / 2e-6
    assert abs(d1 - d2) <= 1e-4 * abs(d1)
