"""Throughput of the PIC particle stage on one MI355X.

Workload (the throughput configuration): periodic unit cube of 48^3 linear
hexahedra, N=3, 200 electrons per element (22,118,400), Maxwellian 2 eV,
frozen smooth E (8 Fourier modes, |E| <= 1e3 V/m), uniform B = 0.01 T,
shape_function deposition R = 1.5h, alpha = 2, TriaTracking, FP64.
A step = Deposition + interpolate + Boris push + TriaTracking (+ element sort),
inputs resident in HBM.  Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-steps/s (interp+push+track+deposit) FP64 vs mixed; % HBM peak"
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0
# compulsory HBM bytes of one particle-step at 200 particles per element:
# x,v read + written and the element id (104 B) plus the per-element tables
# (E 1,536 B, Elem_xGP 1,536 B, source 2,048 B, tracking geometry 216 B) / 200
STEP_BYTES_PER_PARTICLE = 104 + (1536 + 1536 + 2048 + 216) / 200.0


def sf_flops_per_particle(R_over_h, rad_over_h, dofs_per_elem):
    """Algorithmic FP64 flops of one shape-function deposit of one particle
    (the reference's algorithm, depoChargeOnDOFsSF): every DOF of every element
    whose barycentre lies within R + ElemRadius is distance-tested (8 flops),
    and every DOF within R receives S^2 * (q v, q) (4 flops)."""
    cand = dofs_per_elem * 4.0 / 3.0 * math.pi * (R_over_h + rad_over_h) ** 3
    inrad = dofs_per_elem * 4.0 / 3.0 * math.pi * R_over_h ** 3
    return 8.0 * cand + 4.0 * inrad, cand, inrad


def measured_traffic(kernel, workload="cube48_N3_sf"):
    """HBM bytes per launch of `kernel` on a workload, as recorded by the
    FETCH_SIZE / WRITE_SIZE counter passes in profiles/ (None if absent)."""
    pdir = os.path.join(ROOT, "profiles")
    try:
        names = sorted(f for f in os.listdir(pdir) if f.endswith("_hbm_traffic.md") and workload in f)
    except OSError:
        return None
    for f in names:
        for line in open(os.path.join(pdir, f)):
            if line.startswith("traffic_bytes_per_launch %s:" % kernel):
                return int(float(line.split(":")[1]))
    return None


def run_cube(w, args, precision, keep=False):
    """Warm-up, then the timed steps; with keep, the initial deposition and the
    final particle state for the mixed-precision error report (outside the
    timed region)."""
    from piclas_amd import stage, workloads as W
    n = len(w.PartState)
    s = stage.PICStage(w.mesh, w.N, max_particles=n, precision=precision, **W.stage_kwargs(w))
    s.set_field(w.E)
    s.set_particles(w.PartState, elem=w.elem, species=w.species)
    s.synchronize()
    res = {}
    if keep:
        s.deposition()
        res["PS0"] = s.part_source()
    if args.warmup:
        s.time_step(w.dt, args.warmup)
    s.synchronize()
    s.reset_stats()
    t0 = time.perf_counter()
    s.time_step(w.dt, args.steps)
    s.synchronize()
    res["t"] = time.perf_counter() - t0
    res["stats"] = s.stats()
    if keep:
        res["state"] = s.particles()[0]
    s.close()
    return res


def mixed_report(w, args, r64):
    """The same workload in mixed precision: throughput, speedup and the
    error against the FP64 run (deposited source of the initial state, and
    positions / kinetic energy after warm-up + timed steps)."""
    import numpy as np
    from piclas_amd import stage
    rm = run_cube(w, args, stage.PREC_MIXED, keep=True)
    n = len(w.PartState)
    a, b = rm["PS0"], r64["PS0"]
    ps_m, ps_d = rm["state"], r64["state"]
    m_e = w.mass[0]
    ke_m = 0.5 * m_e * float((ps_m[:, 3:] ** 2).sum())
    ke_d = 0.5 * m_e * float((ps_d[:, 3:] ** 2).sum())
    stm = rm["stats"]
    return {"value": n * args.steps / rm["t"], "ms_per_step": rm["t"] / args.steps * 1e3,
            "speedup_vs_fp64": r64["t"] / rm["t"],
            "deposit_ms_per_step": stm["ms_deposit"] / args.steps,
            "rho_rel_err_L2": float(np.linalg.norm(a - b) / np.linalg.norm(b)),
            "rho_rel_err_Linf": float(np.abs(a - b).max() / np.abs(b).max()),
            "pos_err_max_over_L": float(np.abs(ps_m[:, :3] - ps_d[:, :3]).max()),
            "kinetic_energy_rel_diff": abs(ke_m - ke_d) / ke_d,
            "after_steps": args.warmup + args.steps,
            "note": "FP32 Lagrange basis, field interpolation and shape weights, FP32 partial sums per source chunk "
                    "into FP64; FP64 positions, Newton and face tests"}


def exchange_timing(w, reps=10):
    """The rho/E hand-over between the stage and a field solver: U_N%E from
    device memory (a CUDA tensor, device-to-device) and from host memory, and
    PartSource to host memory (the device path hands over a pointer)."""
    import numpy as np
    from piclas_amd import stage, workloads as W
    s = stage.PICStage(w.mesh, w.N, max_particles=1, **W.stage_kwargs(w))
    E_host = np.ascontiguousarray(w.E)
    E_dev = stage.DeviceBuffer.from_array(E_host)
    s.set_particles(w.PartState[:1], elem=w.elem[:1], species=w.species[:1])
    s.deposition()

    def timed(f):   # every call returns after its copies have finished
        f()
        t0 = time.perf_counter()
        for _ in range(reps):
            f()
        return (time.perf_counter() - t0) / reps * 1e3

    out = {"field_bytes": int(E_host.nbytes), "source_bytes": int(w.mesh.nElems * (w.N + 1) ** 3 * 4 * 8),
           "set_field_device_ms": timed(lambda: s.set_field_device(E_dev.ptr)),
           "set_field_host_ms": timed(lambda: s.set_field(E_host)),
           "part_source_host_ms": timed(s.part_source),
           "part_source_device": "pointer (pic_part_source_device), no copy"}
    s.close()
    E_dev.free()
    return out


def host_threads():
    """Host threads the CPU baseline uses: every core this process may run on,
    capped by OMP_NUM_THREADS when the host sets it (a GPU box grants one
    GPU's share of its cores that way)."""
    cores = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS", "")
    return min(cores, int(omp)) if omp.isdigit() and int(omp) > 0 else cores


def host_info():
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "affinity_cores": len(os.sched_getaffinity(0)),
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS"), "cpu_model": model}


def cpu_baseline(w, fraction, steps, threads):
    """The FP64 CPU oracle (the reference's algorithm restated in C, OpenMP)
    on a bounded sample of the same workload: the first `fraction` of the
    particles on the whole mesh (the per-particle work of every stage does
    not depend on the particle density).  The threaded shape-function
    deposit sums per-thread copies of PartSource (226 MB each at 48^3); that
    zeroing and summing has no counterpart in the reference's
    domain-decomposed MPI path, so it is timed apart and left out of
    `value` (`value_with_private_copies` keeps it)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle as O
    n = int(len(w.PartState) * fraction)
    idx = np.arange(n)
    o = O.Oracle(w.mesh, w.N, deposition=w.deposition, sf_radius=w.sf_radius, sf_alpha=w.sf_alpha,
                 external_field=w.external_field, charge=w.charge, mass=w.mass, mpf=w.mpf)
    o.set_threads(threads)
    o.set_field(w.E)
    o.set_particles(w.PartState[idx], elem=w.elem[idx], species=w.species[idx])
    t0 = time.perf_counter()
    o.time_step(w.dt, steps)
    t = time.perf_counter() - t0
    t_priv = o.private_reduce_seconds()
    return {"value": n * steps / (t - t_priv), "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "value_with_private_copies": n * steps / t, "private_copies_s": round(t_priv, 3),
            "host": host_info(),
            "sample": "%d of %d particles (every element's mesh, same field), %d step(s), %.1f s wall, of which "
                      "%.1f s zeroing and summing the per-thread source copies" % (n, len(w.PartState), steps, t,
                                                                                     t_priv)}


# Algorithmic bytes of the push kernels (interpolation, Boris push, the test
# in the last element) per particle: reads x, v (48), elem (4), species (4),
# IsNewPart (1); writes x, v (48), IsNewPart (1), the new owner (4); with
# RefMapping also PartPosRef read and written (24 + 24).  Element tables are
# counted once per element: field at the Gauss points (1,536 B), affine map
# (96 B) and side corners (576 B) on the affine mesh; field, XCL_NGeo (648 B),
# dXCL_NGeo (1,944 B) and the Gauss points of the start guess (1,536 B) with
# RefMapping at NGeo = 2.
PUSH_BYTES_AFFINE = 110
PUSH_BYTES_REFMAP = 158


def push_roofline(st, n, steps, kernel, bytes_per_particle, table_bytes):
    """HBM roofline of the push kernel (kernel A) from its HIP-event time over
    the eager steps."""
    ms = st["ms_push_main"] / max(steps, 1)
    if ms <= 0:
        return None
    gbs = (bytes_per_particle * n + table_bytes) / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(gbs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": round(gbs / PEAK_HBM_GBS, 4), "kernel": kernel, "avg_launch_ms": round(ms, 4),
            "algorithmic_bytes_per_particle": bytes_per_particle, "element_table_bytes": table_bytes}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100, help="timed steps (the cube and annulus workloads state 100)")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nx", type=int, default=48)
    ap.add_argument("--ppc", type=int, default=200)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision run")
    ap.add_argument("--no-exchange", action="store_true", help="skip the rho/E hand-over timing (profiling runs)")
    ap.add_argument("--autotune", action="store_true", help="pick graph replay / shape-function kernel by measurement")
    ap.add_argument("--cpu-fraction", type=float, default=1.0 / 4)
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--workload", choices=["cube", "annulus", "adjoint", "wave"], default="cube",
                    help="cube: 48^3 linear box (default); annulus: 24x96x24 curved NGeo=2, RefMapping, 30M "
                         "particles; adjoint: 16^3 perturbed plasma, 2M particles, K self-consistent steps and "
                         "their adjoint; wave: 16x4x4 plasma-wave box, 16,384 particles, CVWM (launch-bound: "
                         "eager steps against graph replay)")
    ap.add_argument("--adjoint-steps", type=int, default=50, help="adjoint workload: K")
    ap.add_argument("--adjoint-every", type=int, default=10, help="adjoint workload: checkpoint interval")
    ap.add_argument("--particles", type=int, default=30_000_000, help="annulus workload only")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; multi-GPU runs are not part of it (got --gpus %d)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    from piclas_amd import stage, workloads as W
    if args.workload == "annulus":
        return bench_annulus(args)
    if args.workload == "adjoint":
        return bench_adjoint(args)
    if args.workload == "wave":
        return bench_wave(args)
    w = W.thermal_cube(nx=args.nx, ppc=args.ppc)
    n = len(w.PartState)
    print("bench: %s, %d particles, %d warmup + %d timed steps" % (w.name, n, args.warmup, args.steps),
          file=sys.stderr, flush=True)
    r64 = run_cube(w, args, stage.PREC_FP64, keep=not args.no_mixed)
    t, st = r64["t"], r64["stats"]
    n_now = st["n_particles"]
    value = n * args.steps / t

    # roofline of the dominant kernel (shape-function gather), timed with HIP
    # events on the stage's own stream
    h = 1.0 / args.nx
    fpp, cand, inrad = sf_flops_per_particle(w.sf_radius / h, math.sqrt(3) / 2, (w.N + 1) ** 3)
    avg_ms = st["ms_deposit_main"] / max(st["n_deposit_launches"], 1)
    achieved = fpp * n_now / (avg_ms * 1e-3) / 1e12
    a = w.sf_alpha if w.sf_alpha == 2 else 0
    kname, note = {
        stage.KERNEL_SF_CARTESIAN: ("k_sf_gather_cart<%d, false>" % a,
                                    "FP64 issue bound: per pair a subtraction, the clamp and the square on the VALU "
                                    "(axis-aligned tensor grid of Gauss points), products summed on FP64 MFMA (4x4x4)"),
        stage.KERNEL_SF_MATRIX: ("k_sf_gather_mfma<%d, true, false>" % a,
                                 "FP64 issue bound: S on the VALU, its products summed on FP64 MFMA (4x4x4)"),
    }.get(st["deposit_kernel"], ("k_sf_gather<%d, %d>" % (w.N, a), "FP64 issue bound: all on the VALU"))
    note += "; peak = MI355X FP64 vector peak (the FP64 matrix peak is the same 78.6 TFLOP/s)"
    traffic = measured_traffic(kname) if (args.nx, args.ppc) == (48, 200) else None
    roof = {"bound": "mfma", "achieved": round(achieved, 3), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_FP64_TFLOPS, 4), "traffic": traffic,
            "kernel": kname, "avg_launch_ms": round(avg_ms, 4),
            "algorithmic_flops_per_particle": round(fpp, 1), "note": note}
    ms_step = t / args.steps * 1e3
    push = push_roofline(st, n_now, args.steps, "k_push_inside<3, false, true>", PUSH_BYTES_AFFINE,
                         (args.nx ** 3) * 2208)
    if push is not None:
        push["traffic"] = measured_traffic(push["kernel"]) if (args.nx, args.ppc) == (48, 200) else None
    out = {"metric": METRIC, "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": w.name, "elements": args.nx ** 3, "N": w.N, "particles": n,
                      "deposition": "shape_function R=1.5h alpha=2", "tracking": "triatracking",
                      "parallelism": "1 GPU"},
           "step_hbm_frac": round(STEP_BYTES_PER_PARTICLE * value / (PEAK_HBM_GBS * 1e9), 5),
           "stage_ms_per_step": {"deposit": st["ms_deposit"] / args.steps,
                                 "interp_push_track": st["ms_push_track"] / args.steps,
                                 "sort": st["ms_sort"] / args.steps},
           "roofline": roof, "push_roofline": push}
    if not args.no_mixed:
        out["mixed"] = mixed_report(w, args, r64)
    if not args.no_exchange:
        out["exchange"] = exchange_timing(w)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(w, args.cpu_fraction, args.cpu_steps, host_threads())
    print(json.dumps(out), flush=True)


def bench_annulus(args):
    """Uneven occupancy on curved elements: 24x96x24 NGeo=2 annulus, RefMapping,
    electrons and Ar+ ions (2 to ~800 per element), shape-function deposition."""
    from piclas_amd import stage, workloads as W
    w = W.annulus_plasma(n_total=args.particles)
    n = len(w.PartState)
    print("bench: %s, %d particles, %d warmup + %d timed steps" % (w.name, n, args.warmup, args.steps),
          file=sys.stderr, flush=True)
    tuning = {}
    if args.autotune:
        from piclas_amd.autotune import autotune
        tuning, table = autotune(w, steps=3, warmup=1, verbose=True)
    s = stage.PICStage(w.mesh, w.N, max_particles=n, **W.stage_kwargs(w))
    if tuning:
        s.set_tuning(**tuning)
    s.set_field(w.E)
    t0 = time.perf_counter()
    s.set_particles(w.PartState, elem=w.elem, species=w.species)
    s.synchronize()
    t_locate = time.perf_counter() - t0
    n_in = s.stats()["n_particles"]
    if args.warmup:
        s.time_step(w.dt, args.warmup)
    s.synchronize()
    s.reset_stats()
    t0 = time.perf_counter()
    s.time_step(w.dt, args.steps)
    s.synchronize()
    t = time.perf_counter() - t0
    st = s.stats()
    nb = args.steps
    if tuning.get("use_graph"):
        # a replayed graph is timed as one piece: the per-stage times and the
        # push kernel's events come from a few eager steps after the timed ones
        nb = 3
        s.set_tuning(**dict(tuning, use_graph=False))
        s.reset_stats()
        s.time_step(w.dt, nb)
        s.synchronize()
        sb = s.stats()
        st = dict(sb, n_particles=st["n_particles"], n_lost=st["n_lost"])
    s.close()
    ms_step = t / args.steps * 1e3
    out = {"metric": METRIC, "value": n_in * args.steps / t, "unit": "particle-steps/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": w.name, "elements": w.mesh.nElems, "NGeo": w.mesh.NGeo, "N": w.N,
                      "particles": n, "particles_placed": n_in, "deposition": "shape_function R=0.03 m alpha=2",
                      "tracking": "refmapping", "parallelism": "1 GPU"},
           "setup_s_locate_and_sort": round(t_locate, 3), "tuning": tuning,
           "push_roofline": push_roofline(st, st["n_particles"], nb, "k_rm_push_inside<3, 3, false>",
                                          PUSH_BYTES_REFMAP, w.mesh.nElems * 5664),
           "stage_ms_per_step": {"deposit": st["ms_deposit"] / nb,
                                 "interp_push_track": st["ms_push_track"] / nb,
                                 "sort": st["ms_sort"] / nb,
                                 "from": "the timed steps" if nb == args.steps else
                                         "%d eager steps after the timed (replayed) ones" % nb},
           "n_removed_or_lost": {"lost": st["n_lost"], "inside_after": st["n_particles"]}}
    print(json.dumps(out), flush=True)


def bench_wave(args):
    """A small, launch-bound step: 16x4x4 periodic box of [0, 4 pi] x [0,1]^2,
    N = 3, 16,384 electrons with density n0 (1 + 0.01 cos(x/2)), CVWM,
    TriaTracking, frozen field of the perturbation.  Times eager steps and
    the replay of the captured two-step graph (the tuner's table)."""
    from piclas_amd import stage, workloads as W
    from piclas_amd.autotune import autotune
    w = W.plasma_wave_box()
    n = len(w.PartState)
    print("bench: %s, %d particles, %d timed steps per candidate" % (w.name, n, args.steps), file=sys.stderr,
          flush=True)
    best, table = autotune(w, steps=args.steps, warmup=max(args.warmup, 2), verbose=True)
    ms = {("graph" if c["use_graph"] else "eager"): t for c, t in table}
    out = {"metric": "particle-steps/s (interp+push+track+deposit), launch-bound small step",
           "value": n / (ms[("graph" if best["use_graph"] else "eager")] * 1e-3), "unit": "particle-steps/s",
           "n_gpus": 1, "steps": args.steps, "warmup": max(args.warmup, 2),
           "ms_per_step": ms[("graph" if best["use_graph"] else "eager")], "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": w.name, "elements": w.mesh.nElems, "N": w.N, "particles": n,
                      "deposition": "cell_volweight_mean", "tracking": "triatracking", "parallelism": "1 GPU"},
           "tuning": best, "ms_per_step_eager": ms["eager"], "ms_per_step_graph": ms["graph"],
           "graph_speedup": ms["eager"] / ms["graph"]}
    print(json.dumps(out), flush=True)


def bench_adjoint(args):
    """The differentiable workload at full size: periodic 16^3 box, N = 3,
    2 M electrons with density n0 (1 + alpha cos 2 pi x), field closed by the
    fixed spectral Poisson operator, loss = field energy after K
    self-consistent steps, shape function alpha = 3.  Times the forward
    alone and the forward with checkpoints plus the reverse sweep
    (pic_adjoint), and checks dL/dalpha against central differences of the
    forward (h = 1e-6)."""
    from piclas_amd import adjoint as A, stage
    K, every = args.adjoint_steps, args.adjoint_every
    w, dxda = A.landau_box(n_el=16, n_part=2_000_000, steps=K)
    n = len(w.PartState)
    print("bench: %s, %d particles, K = %d, checkpoint every %d" % (w.name, n, K, every), file=sys.stderr, flush=True)
    G = A.periodic_poisson_operator(w.mesh)
    s = stage.PICStage(w.mesh, w.N, max_particles=n, deposition=w.deposition, sf_radius=w.sf_radius,
                       sf_alpha=w.sf_alpha, charge=w.charge, mass=w.mass, mpf=w.mpf)
    s.set_field_operator(G)

    def forward(ws):
        s.set_particles(ws.PartState, elem=ws.elem, species=ws.species)
        s.synchronize()
        t0 = time.perf_counter()
        loss = s.sc_time_step(ws.dt, K)
        return loss, time.perf_counter() - t0

    forward(w)   # warm-up
    loss, t_fwd = forward(w)
    s.set_particles(w.PartState, elem=w.elem, species=w.species)
    s.synchronize()
    t0 = time.perf_counter()
    loss_a, gx, gv = s.adjoint(w.dt, K, every)
    t_adj = time.perf_counter() - t0
    dLda = float((gx * dxda).sum())
    h = 1e-6
    wp, _ = A.landau_box(n_el=16, n_part=2_000_000, alpha=0.01 + h, steps=K)
    wm, _ = A.landau_box(n_el=16, n_part=2_000_000, alpha=0.01 - h, steps=K)
    fd = (forward(wp)[0] - forward(wm)[0]) / (2 * h)
    s.close()
    out = {"metric": "particle-steps/s of forward + adjoint (K self-consistent steps, checkpointed reverse sweep)",
           "value": n * K / t_adj, "unit": "particle-steps/s", "n_gpus": 1, "steps": K, "warmup": 1,
           "ms_per_step": t_adj / K * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64", "data": "synthetic",
           "config": {"workload": w.name, "elements": w.mesh.nElems, "N": w.N, "particles": n,
                      "deposition": "shape_function R=1.5h alpha=3", "tracking": "triatracking",
                      "checkpoint_every": every, "parallelism": "1 GPU"},
           "forward_ms_per_step": t_fwd / K * 1e3, "adjoint_over_forward": t_adj / t_fwd,
           "loss": loss_a, "loss_forward": loss,
           "dL_dalpha": {"adjoint": dLda, "central_difference": fd, "rel_err": abs(dLda - fd) / abs(fd)}}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
