"""Choice of execution variants for a workload by measurement.

The stage has a few execution knobs that do not change its results beyond
round-off in the order of the shape-function sums (pic_set_tuning): graph
replay of the time step, the shape-function kernel, persistent waves for the
matrix-core shape-function sums, and record loads shared by the four targets
of a workgroup (both bit-identical to one wave per target).
``autotune`` times each candidate on a scratch stage built for the same mesh
and particles (the caller's stage is not touched) and returns the fastest as
keyword arguments of ``PICStage.set_tuning``.

    best, table = autotune(w, steps=5)
    s.set_tuning(**best)
"""
import time

from . import stage
from . import workloads as W


def candidates(N, deposition=stage.DEPO_SF):
    out = []
    for g in (False, True):
        for k in ((1, 2) if N == 3 and deposition == stage.DEPO_SF else (0,)):
            for p, c in (((False, False), (True, False), (False, True)) if k == 2 else ((False, False),)):
                out.append({"use_graph": g, "sf_kernel": k, "sf_persistent": p, "sf_shared_loads": c})
    return out


def autotune(w, steps=4, warmup=2, precision=stage.PREC_FP64, cands=None, verbose=False):
    """Returns (best tuning dict, [(tuning, ms per step), ...])."""
    cands = cands or candidates(w.N, w.deposition)
    n = len(w.PartState)
    s = stage.PICStage(w.mesh, w.N, max_particles=n, precision=precision, **W.stage_kwargs(w))
    s.set_field(w.E)
    table = []
    try:
        for c in cands:
            s.set_particles(w.PartState, elem=w.elem, species=w.species)
            s.set_tuning(**c)
            s.time_step(w.dt, warmup)
            s.synchronize()
            t0 = time.perf_counter()
            s.time_step(w.dt, steps)
            s.synchronize()
            ms = (time.perf_counter() - t0) / steps * 1e3
            table.append((c, ms))
            if verbose:
                print("autotune %s: %.3f ms/step" % (c, ms), flush=True)
    finally:
        s.close()
    best = min(table, key=lambda r: r[1])[0]
    return best, table
