"""The hand-written adjoint of interpolate + push + deposit (pic_adjoint),
wrapped in a torch.autograd.Function, against central finite differences of
the device forward and of the FP64 CPU oracle's forward.

Workload: perturbed periodic plasma (density n0 (1 + alpha cos 2 pi x)),
field closed by the fixed spectral Poisson operator of piclas_amd.adjoint,
loss = field energy after K self-consistent steps.  Finite-difference steps:
1e-6 in alpha (1e-4 relative); 1e-4 of the thermal speed for one particle's
velocity (with 1e-6 the loss change, ~1e-17 J, is within a few hundred
round-off units of the forward's sums and the difference quotient carries
~1e-5 noise); 1e-6 of the thermal speed along a random direction in all
velocities (the autograd check).
"""
import numpy as np
import pytest

from piclas_amd import adjoint as A
from piclas_amd import stage

pytestmark = pytest.mark.gpu

K, EVERY = 20, 10


def make_stage(w, G):
    s = stage.PICStage(w.mesh, w.N, max_particles=len(w.PartState), deposition=w.deposition, sf_radius=w.sf_radius,
                       sf_alpha=w.sf_alpha, charge=w.charge, mass=w.mass, mpf=w.mpf)
    s.set_field_operator(G)
    return s


def loss_of(s, w, ps):
    s.set_particles(ps, elem=w.mesh.element_of(ps[:, :3]), species=w.species)
    return s.adjoint(w.dt, K, EVERY, gradients=False)[0]


def test_sc_forward_matches_oracle(oracle_mod):
    w, _ = A.landau_box(n_el=6, n_part=4000, steps=K)
    G = A.periodic_poisson_operator(w.mesh)
    print("case sc forward landau6 n=%d K=%d" % (len(w.PartState), K), flush=True)
    with make_stage(w, G) as s:
        s.set_particles(w.PartState, elem=w.elem, species=w.species)
        loss = s.sc_time_step(w.dt, K)
        ps, el, ins = s.particles()
    o = oracle_mod.Oracle(w.mesh, w.N, deposition=w.deposition, sf_radius=w.sf_radius, sf_alpha=w.sf_alpha,
                          charge=w.charge, mass=w.mass, mpf=w.mpf)
    o.sc_set_operator(G)
    o.set_field(w.E)
    o.set_particles(w.PartState, elem=w.elem, species=w.species)
    oloss = o.sc_forward(w.dt, K)
    ops, oel, oins = o.particles()
    print("loss %.12e oracle %.12e" % (loss, oloss))
    assert abs(loss - oloss) <= 1e-10 * abs(oloss)
    assert np.abs(ps[:, :3] - ops[:, :3]).max() <= 1e-12
    assert np.array_equal(el, oel)


def test_adjoint_matches_finite_differences(oracle_mod):
    w, dxda = A.landau_box(n_el=6, n_part=4000, steps=K)
    G = A.periodic_poisson_operator(w.mesh)
    print("case adjoint landau6 n=%d K=%d every=%d" % (len(w.PartState), K, EVERY), flush=True)
    with make_stage(w, G) as s:
        s.set_particles(w.PartState, elem=w.elem, species=w.species)
        loss, gx, gv = s.adjoint(w.dt, K, EVERY)
        # dL/dalpha from the adjoint and from central differences of the device forward
        dLda = float((gx * dxda).sum())
        h = 1e-6
        wp, _ = A.landau_box(n_el=6, n_part=4000, alpha=0.01 + h, steps=K)
        wm, _ = A.landau_box(n_el=6, n_part=4000, alpha=0.01 - h, steps=K)
        fd = (loss_of(s, wp, wp.PartState) - loss_of(s, wm, wm.PartState)) / (2 * h)
        print("dL/dalpha adjoint %.10e  FD(device) %.10e" % (dLda, fd))
        assert abs(dLda - fd) <= 1e-6 * abs(fd)
        # velocities of a few particles
        vth = np.abs(w.PartState[:, 3:]).std()
        rng = np.random.default_rng(5)
        for p, c in zip(rng.integers(0, len(w.PartState), 4), rng.integers(0, 3, 4)):
            hv = 1e-4 * vth
            up, dn = w.PartState.copy(), w.PartState.copy()
            up[p, 3 + c] += hv
            dn[p, 3 + c] -= hv
            fdv = (loss_of(s, w, up) - loss_of(s, w, dn)) / (2 * hv)
            print("dL/dv[%d,%d] adjoint %.10e  FD %.10e" % (p, c, gv[p, c], fdv))
            assert abs(gv[p, c] - fdv) <= 1e-6 * max(abs(fdv), 1e-3 * np.abs(gv).max())
    # the FP64 CPU oracle's forward gives the same derivative in alpha
    def oloss(a):
        wa, _ = A.landau_box(n_el=6, n_part=4000, alpha=a, steps=K)
        o = oracle_mod.Oracle(wa.mesh, wa.N, deposition=wa.deposition, sf_radius=wa.sf_radius,
                              sf_alpha=wa.sf_alpha, charge=wa.charge, mass=wa.mass, mpf=wa.mpf)
        o.sc_set_operator(G)
        o.set_field(wa.E)
        o.set_particles(wa.PartState, elem=wa.elem, species=wa.species)
        return o.sc_forward(wa.dt, K)
    ofd = (oloss(0.01 + 1e-6) - oloss(0.01 - 1e-6)) / 2e-6
    print("dL/dalpha FD(oracle) %.10e" % ofd)
    assert abs(dLda - ofd) <= 1e-6 * abs(ofd)


def test_autograd_function_wrapper():
    """piclas_amd.adjoint.stage_loss (a torch.autograd.Function around
    pic_adjoint): backward returns the adjoint gradients, and a random
    directional derivative matches central differences of the forward."""
    import torch

    w, _ = A.landau_box(n_el=6, n_part=3000, steps=K)
    G = A.periodic_poisson_operator(w.mesh)
    s = make_stage(w, G)

    StageLoss = A.stage_loss(s, w.dt, K, EVERY, species=w.species, locate=w.mesh.element_of)

    x0 = torch.tensor(w.PartState[:, :3], dtype=torch.float64, requires_grad=True)
    v0 = torch.tensor(w.PartState[:, 3:], dtype=torch.float64, requires_grad=True)
    L = StageLoss(x0, v0)
    L.backward()
    rng = np.random.default_rng(11)
    dv = torch.tensor(rng.normal(size=v0.shape) * 1e-6 * float(v0.detach().std()), dtype=torch.float64)
    with torch.no_grad():
        lp = StageLoss(x0, v0 + dv)
        lm = StageLoss(x0, v0 - dv)
    fd = float(lp - lm) / 2.0
    ad = float((v0.grad * dv).sum())
    print("directional derivative adjoint %.10e  FD %.10e" % (ad, fd))
    assert abs(ad - fd) <= 1e-6 * abs(fd)
    s.close()


def test_adjoint_stated_size():
    """The differentiable workload at its stated size: 16^3 linear hexes,
    N = 3, 2,000,000 electrons at 8 eV, alpha = 0.01, K = 50 steps with
    checkpoints every 10; dL/dalpha from the adjoint against central
    differences (h = 1e-6) of the device forward."""
    Kf = 50
    w, dxda = A.landau_box(n_el=16, n_part=2_000_000, steps=Kf)
    G = A.periodic_poisson_operator(w.mesh)
    print("case adjoint landau16 n=%d K=%d every=%d" % (len(w.PartState), Kf, EVERY), flush=True)
    with make_stage(w, G) as s:
        s.set_particles(w.PartState, elem=w.elem, species=w.species)
        loss, gx, gv = s.adjoint(w.dt, Kf, EVERY)
        dLda = float((gx * dxda).sum())
        h = 1e-6
        fdl = []
        for a in (0.01 + h, 0.01 - h):
            wa, _ = A.landau_box(n_el=16, n_part=2_000_000, alpha=a, steps=Kf)
            s.set_particles(wa.PartState, elem=wa.mesh.element_of(wa.PartState[:, :3]), species=wa.species)
            fdl.append(s.adjoint(w.dt, Kf, EVERY, gradients=False)[0])
        fd = (fdl[0] - fdl[1]) / (2 * h)
        print("dL/dalpha adjoint %.10e  FD(device) %.10e  rel %.2e" % (dLda, fd, abs(dLda - fd) / abs(fd)))
        assert abs(dLda - fd) <= 1e-6 * abs(fd)
