"""pic_emit_particles_noalloc: the emission with its device staging in the
library's fixed per-GPU array (no allocation in the call) gives the same
particles as pic_emit_particles, bit for bit, on two calls in a row (the
second smaller, so stale staging past its end must not matter), and rejects
emissions beyond its staging limit with PIC_ECAPACITY before any launch."""
import numpy as np
import pytest

from piclas_amd import mesh as M
from piclas_amd import stage

pytestmark = pytest.mark.gpu

NOALLOC_MAX = 1 << 19


def make_stage(mesh, n):
    return stage.PICStage(mesh, 3, max_particles=n, deposition=stage.DEPO_SF, sf_radius=0.3, sf_alpha=2,
                          charge=(-stage.ELEMENTARY_CHARGE,), mass=(stage.ELECTRON_MASS,), mpf=(1e3,))


def inits_for(n_cuboid, nxyz):
    return [
        dict(species=1, space=stage.SPACE_CUBOID, particle_number=n_cuboid, base_point=(0.05, 0.1, 0.15),
             base_vector1=(0.8, 0.0, 0.0), base_vector2=(0.0, 0.75, 0.0), cuboid_height=0.7,
             velocity=stage.VELO_MAXWELL_LPN, velo_ic=1e4, velo_vec=(0.0, 1.0, 0.0), temperature=23209.0),
        dict(species=1, space=stage.SPACE_SIN_DEVIATION, max_particle_number=nxyz, amplitude=0.01,
             wave_number=1.0, velocity=stage.VELO_CONSTANT, velo_ic=2e5, velo_vec=(1.0, 0.0, 0.0)),
    ]


def emitted(s, inits, seed, noalloc):
    inside = s.emit_particles(inits, seed=seed, noalloc=noalloc)
    ps, el, ins = s.particles()
    return inside, ps, el, ins


def test_noalloc_emission_matches_allocating_emission():
    mesh = M.cartesian_box((6, 6, 6))
    cases = [(60000, (24, 12, 10)), (7000, (6, 5, 4))]   # second call smaller than the first
    cap = max(c + int(np.prod(x)) for c, x in cases)
    with make_stage(mesh, cap) as a, make_stage(mesh, cap) as b:
        for seed, (nc, nxyz) in zip((20261203, 977), cases):
            inits = inits_for(nc, nxyz)
            n = nc + int(np.prod(nxyz))
            print("case noalloc emission seed=%d n=%d" % (seed, n), flush=True)
            ra = emitted(a, inits, seed, noalloc=False)
            rb = emitted(b, inits, seed, noalloc=True)
            assert ra[0] == rb[0] == n
            assert ra[1].shape == (n, 6)
            assert np.array_equal(ra[1], rb[1]), "PartState differs"
            assert np.array_equal(ra[2], rb[2]), "elements differ"
            assert np.array_equal(ra[3], rb[3]) and rb[3].all()
            # the emitted particles step alike
            a.time_step(1e-11, 2)
            b.time_step(1e-11, 2)
            assert np.array_equal(a.particles()[0], b.particles()[0])
            assert np.array_equal(a.part_source(), b.part_source())


def test_noalloc_emission_rejects_beyond_staging_limit():
    mesh = M.cartesian_box((4, 4, 4))
    n = NOALLOC_MAX + 1
    init = dict(species=1, space=stage.SPACE_CUBOID, particle_number=n, base_point=(0.0, 0.0, 0.0),
                base_vector1=(1.0, 0.0, 0.0), base_vector2=(0.0, 1.0, 0.0), cuboid_height=1.0,
                velocity=stage.VELO_CONSTANT, velo_ic=0.0, velo_vec=(1.0, 0.0, 0.0))
    print("case noalloc emission beyond limit n=%d" % n, flush=True)
    with make_stage(mesh, n) as s:
        with pytest.raises(RuntimeError, match="staging limit"):
            s.emit_particles([init], seed=1, noalloc=True)
        # the stage is still usable: the allocating emission takes the same request
        assert s.emit_particles([init], seed=1) == n
