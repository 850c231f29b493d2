"""Particle emission on the device (pic_emit_particles): SetParticlePosition
'cuboid' / 'sin_deviation' and SetParticleVelocity 'constant' /
'maxwell_lpn' (particles/emission/particle_position_and_velocity.f90,
particle_emission_tools.f90), localization on the device, and a time step
of the emitted particles against the oracle.

Checks: cuboid positions bit-exact against the formula evaluated with the
CPU Philox (tests/philox_ref.py, pinned by known-answer vectors);
sin_deviation positions against the lattice formula; Maxwell velocities
against the polar method on the CPU Philox to 1e-13, and mean / variance
against kT/m; reproducibility in the seed; elements against the closed-form
Cartesian owner."""
import numpy as np
import pytest

from philox_ref import uniform2
from piclas_amd import mesh as M
from piclas_amd import stage, workloads as W
from test_gpu_parity import compare

pytestmark = pytest.mark.gpu

KB = 1.380648813e-23


def make_stage(mesh, n, **kw):
    return stage.PICStage(mesh, 3, max_particles=n, deposition=stage.DEPO_SF, sf_radius=0.3, sf_alpha=2,
                          charge=(-stage.ELEMENTARY_CHARGE,), mass=(stage.ELECTRON_MASS,), mpf=(1e3,), **kw)


def maxwell_cpu(seed, pid, T, m, drift, draw0):
    """CalcVelocity_maxwell_lpn on the CPU Philox, draw by draw as the device
    (the position took draws 0 .. draw0-1)."""
    out = np.zeros((len(pid), 3))
    for j, p in enumerate(pid):
        draw, w = draw0, []
        for _ in range(2):
            s = 2.0
            while s >= 1.0 or s == 0.0:
                r0, r1 = uniform2(seed, np.array([p]), draw)
                draw += 1
                a0, a1 = 2. * r0[0] - 1., 2. * r1[0] - 1.
                s = a0 * a0 + a1 * a1
            f = np.sqrt(-2 * KB * T / m * np.log(s) / s)
            w += [a0 * f, a1 * f]
        out[j] = np.array(w[:3]) + np.asarray(drift)
    return out


def test_cuboid_maxwell():
    mesh = M.cartesian_box((6, 6, 6))
    n, seed, T = 20000, 20261210, 11604.5
    init = dict(species=1, space=stage.SPACE_CUBOID, particle_number=n, base_point=(0.1, 0.15, 0.2),
                base_vector1=(0.7, 0.0, 0.0), base_vector2=(0.0, 0.65, 0.0), cuboid_height=0.6,
                velocity=stage.VELO_MAXWELL_LPN, velo_ic=2e4, velo_vec=(1.0, 0.0, 0.0), temperature=T)
    print("case emission cuboid maxwell n=%d" % n, flush=True)
    with make_stage(mesh, n) as s:
        assert s.emit_particles([init], seed=seed) == n
        ps, el, ins = s.particles()
        assert ins.all()
        # positions: BasePointIC + BaseVector1IC r1 + BaseVector2IC r2 + normal * height * r3
        pid = np.arange(n, dtype=np.uint64)
        r0, r1 = uniform2(seed, pid, 0)
        r2, _ = uniform2(seed, pid, 1)
        x = np.stack([0.1 + 0.7 * r0 + 0.0 * r1 + 0.0 * r2, 0.15 + 0.0 * r0 + 0.65 * r1 + 0.0 * r2,
                      0.2 + 0.0 * r0 + 0.0 * r1 + 0.6 * r2], 1)
        assert np.array_equal(ps[:, :3], x)
        assert np.array_equal(el, mesh.element_of(ps[:, :3]))
        # velocities: the polar method on the same numbers, a sample of particles
        m = stage.ELECTRON_MASS
        sample = np.arange(0, n, 997)
        vcpu = maxwell_cpu(seed, sample, T, m, (2e4, 0.0, 0.0), draw0=2)
        assert np.abs(ps[sample, 3:] - vcpu).max() <= 1e-13 * np.abs(vcpu).max()
        # moments: drift and kT/m
        v = ps[:, 3:]
        var = KB * T / m
        assert abs(v[:, 0].mean() - 2e4) <= 5 * np.sqrt(var / n)
        assert np.abs(v[:, 1:].mean(0)).max() <= 5 * np.sqrt(var / n)
        assert np.abs(v.var(0) / var - 1).max() <= 5 * np.sqrt(2.0 / n)
        # the same seed gives the same particles, another seed others
        s.emit_particles([init], seed=seed)
        assert np.array_equal(s.particles()[0], ps)
        s.emit_particles([init], seed=seed + 1)
        assert not np.array_equal(s.particles()[0], ps)


def test_sin_deviation_constant_velocity_and_step(oracle_mod):
    w = W.plasma_wave_box(ppc=1)
    mesh = w.mesh
    nx, ny, nz = 64, 8, 8
    amp, wn = 0.05, 1.0
    inits = [dict(species=1, space=stage.SPACE_SIN_DEVIATION, max_particle_number=(nx, ny, nz), amplitude=amp,
                  wave_number=wn, velocity=stage.VELO_CONSTANT, velo_ic=1e5, velo_vec=(0.0, 0.6, 0.8))]
    n = nx * ny * nz
    print("case emission sin_deviation n=%d" % n, flush=True)
    kw = W.stage_kwargs(w)
    with stage.PICStage(mesh, w.N, max_particles=n, **kw) as s:
        s.set_field(w.E)
        assert s.emit_particles(inits, seed=7) == n
        ps, el, ins = s.particles()
        # the lattice of SetParticlePositionSinDeviation, k (z) innermost
        lo, hi = np.asarray(mesh.lo), np.asarray(mesh.hi)
        L = np.abs(hi - lo)
        i, j, k = np.meshgrid(np.arange(1, nx + 1), np.arange(1, ny + 1), np.arange(1, nz + 1), indexing="ij")
        xs, ys, zs = L[0] / nx, L[1] / ny, L[2] / nz
        xp = (i * xs - xs * 0.5)
        x = np.stack([(lo[0] + xp + amp * np.sin(wn * (2.0 * np.pi / L[0]) * xp)).ravel(),
                      (lo[1] + j * ys - ys * 0.5).ravel(), (lo[2] + k * zs - zs * 0.5).ravel()], 1)
        assert np.abs(ps[:, :3] - x).max() <= 1e-14 * L.max()
        assert np.array_equal(ps[:, 3:], np.tile([0.0, 0.6 * 1e5, 0.8 * 1e5], (n, 1)))
        assert ins.all() and np.array_equal(el, mesh.element_of(ps[:, :3]))
        # the emitted particles (IsNewPart) step like the oracle's
        s.time_step(w.dt, 5)
        s.deposition()
        o = oracle_mod.Oracle(mesh, w.N, deposition=kw["deposition"], sf_radius=kw["sf_radius"],
                              sf_alpha=kw["sf_alpha"], external_field=kw["external_field"], charge=kw["charge"],
                              mass=kw["mass"], mpf=kw["mpf"])
        o.set_field(w.E)
        o.set_particles(ps, elem=el, species=np.ones(n, np.int32), is_new=np.ones(n, np.uint8))
        o.time_step(w.dt, 5)
        o.deposition()
        compare(s, o, L.max(), mesh=mesh)
