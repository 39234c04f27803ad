"""FP64 RK4 point-updates/s of the fused generalized-harmonic step on one MI355X.

Workload (the plain method at real shape): a static spinning Kerr-Schild hole,
M = 1, a = 0.7, in the Dirichlet box [-8M, 8M]^3 at h = M/16 -- a uniform
forest of 16^3 blocks of 16^3 cells (4096 blocks, 16.8 M points by the
reference's N^3-per-block convention), q = 4, G = 3, the :damped interior
(r_0 = 0.71, r_1 = 1.21, rho_max = 4/M, ramps 1/2), Gaussian constraint
damping (near 1/M, far 1/(10M), width 3M), gamma2 = 0, eps_KO = 1/2, the
closed-form Kerr-Schild gauge source, analytic initial data, cfl = 1/4.

A step is one classical RK4 step (4 stages, each a ghost fill and the fused
right-hand side with the stage update in its epilogue) captured in a hipGraph.
The state is resident in HBM before the timed region.

Prints one JSON line (rank 0) with the metric, a roofline object for the
dominant kernel, the other forms of the same step (the plain verbs, and the
plain per-block form: the fused block kernel launched once per block and
stage), a batch of four holes of uneven block counts stepped by one launch per
kernel against the same holes stepped alone and per block, the three-level
octree around the hole (hole_forest, finest h = 5M/256, planned fill with an
order-6 prolongation) alone and batched with the four holes on the block
layout, and a CPU baseline (the oracle, OpenMP, on a bounded sample).
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

# Algorithmic work per point-update (one point, N^3 per block, advanced by one
# full RK4 step), frozen from the oracle's counting scalar: F_rhs FP64 ops per
# point per right-hand side (the :none branch, q = 4, dissipation on, no gauge
# source), and the stage combination's ops per point per step.
F_RHS = None  # filled from the oracle count at import time of the roofline
B_POINT_UPDATE = 2560.0  # bytes: 4 stages x 640 B (a fused stepper's reads and writes)
FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0


# tgh_rk4 flags per form (include/tgh.h)
FORMS = {"brick-split": 8, "brick-fused": 2, "blocks-fused": 4 | 2, "blocks-split": 4 | 8}
KERNEL = {"brick-fused": "k_brick_rhs<4, KS, hole>", "blocks-fused": "k_rhs<4, KS, hole, RK4 epilogue>",
          "brick-split:head": "k_brick_head<4, KS, hole>", "brick-split:pi": "k_brick_pi<4>",
          "blocks-split:head": "k_head<4, KS, hole, RK4 epilogue>", "blocks-split:pi": "k_pi<4, RK4 epilogue>"}


def note(msg):
    """A progress line on stderr (the JSON line stays alone on stdout)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-sampled", action="store_true", help="skip the sampled-gauge-source line")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint's cost line")
    ap.add_argument("--nblocks", type=int, default=16, help="blocks per axis (16 = the full workload)")
    ap.add_argument("--form", choices=tuple(FORMS), default="brick-split",
                    help="the stepper's layout and kernels: the brick layout (default), or the "
                         "per-block work arrays with one fused kernel or head + Pi launches")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain (verbs) form")
    ap.add_argument("--no-per-block", action="store_true", help="skip the plain per-block form")
    ap.add_argument("--no-batch", action="store_true", help="skip the four-hole batch")
    ap.add_argument("--no-forest", action="store_true", help="skip the three-level forest")
    ap.add_argument("--no-tune", action="store_true", help="keep the default Pi tile shape (no autotuner)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="end with a 1 GiB device-to-device copy, the known-size reference for "
                         "rocprofv3's FETCH_SIZE / WRITE_SIZE counters")
    return ap.parse_args()


def kerr_schild_desc(tgh, nblocks):
    d = tgh.TghDesc()
    d.float_bits, d.q, d.N = 64, 4, 16
    d.nblocks[:] = [nblocks] * 3
    d.periodic[:] = [0, 0, 0]
    L = 8.0 * nblocks / 16
    d.lo[:] = [-L] * 3
    d.hi[:] = [L] * 3
    d.background, d.gauge, d.damping, d.interior = 2, 2, 1, 1
    d.M, d.a = 1.0, 0.7
    d.gamma0_near, d.gamma0_far, d.gamma0_width, d.gamma2 = 1.0, 0.1, 3.0, 0.0
    d.eps_ko = 0.5
    d.r0, d.r1, d.rho_max, d.w_ramp, d.rho_ramp = 0.71, 1.21, 4.0, 0.5, 0.5
    return d


def measured_traffic(kernel, nblocks):
    """HBM bytes per launch of `kernel` on this workload, from the committed
    rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE passes (profiles/traffic.json), each
    scaled by what the same counter reads for a 1 GiB device copy in the same
    run; None where that kernel and workload were not measured."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        rec = json.load(fh).get(kernel)
    if not rec or rec.get("blocks") != nblocks:
        return None
    return rec["fetch_bytes"] + rec["write_bytes"]


def flop_count():
    """F_rhs and F_combine from the oracle's counting instantiation."""
    import oracle_lib
    import tgh_desc as TD
    d = TD.make_desc(q=4, N=8, nblocks=(1, 1, 1), background=TD.TGH_BG_GAUGE_WAVE, gamma0=1.0,
                     gamma2=0.0, eps_ko=0.5, wave_amp=0.05, wave_len=1.0)
    p = oracle_lib.OracleProblem(d)
    u = p.fill_exact(0.0)
    p.prepare(u, 0.0)
    f_rhs = p.count_flops(0, 3, 4, 5)
    f_head = p.count_flops(0, 3, 4, 5, head_only=True)
    # The stage combination: per variable, stage 1 acc = u + c k, Y = u + c k
    # (4 ops); stages 2, 3 acc += c k, Y = u + c k (4); stage 4 u = acc + c k (2).
    f_combine = 20 * (4 + 4 + 4 + 2)
    return f_rhs, f_head, f_combine


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _oracle_native():
    """The oracle built for this host's CPU (g++ -O3 -march=native, OpenMP) into
    a temporary directory, as the baseline is specified; the prebuilt
    oracle/liboracle.so (-O3, generic x86-64) when no compiler is at hand.
    Returns (library path, flags)."""
    import shutil
    import subprocess
    import tempfile
    here = os.path.dirname(os.path.abspath(__file__))
    gxx = shutil.which("g++")
    if gxx:
        out = os.path.join(tempfile.mkdtemp(prefix="tgh_cpu_"), "liboracle_native.so")
        flags = ["-O3", "-march=native", "-std=c++17", "-fPIC", "-fopenmp", "-w", "-shared"]
        r = subprocess.run([gxx] + flags + ["-o", out, os.path.join(here, "oracle", "oracle_capi.cpp")],
                           capture_output=True, text=True, timeout=180)
        if r.returncode == 0:
            return out, " ".join(flags[:2])
    return os.path.join(here, "oracle", "liboracle.so"), "-O3 generic x86-64 (prebuilt)"


def cpu_baseline():
    """The CPU baseline: the oracle's RK4 step (the reference's algorithm,
    OpenMP over blocks, one block per thread by owner) built for this host
    (-O3 -march=native) and run at the host's thread budget (OMP_NUM_THREADS,
    else its core count), one warm-up and the minimum of 10 timed steps, on
    two samples: the CPU-runnable workload (one 16^3 periodic Minkowski block
    with noise, q = 4) and the Kerr-Schild hole's own 64 blocks ([-2,2]^3 at
    h = 1/16: core, layer and exterior of the 4096-block workload's hole)."""
    import oracle_lib
    import tgh_desc as TD
    path, flags = _oracle_native()
    oracle_lib.LIB_PATH = path
    oracle_lib._lib = None
    threads = int(os.environ.get("OMP_NUM_THREADS") or (os.cpu_count() or 1))

    def min_step(desc, points, noise=False):
        p = oracle_lib.OracleProblem(desc)
        u = p.fill_exact(0.0)
        if noise:
            p.add_noise(u, 0x5EED20260917, 1e-6)
        dt = 0.25 * (desc.hi[0] - desc.lo[0]) / (desc.nblocks[0] * desc.N) / p.max_speed(u)
        p.rk4_step(u, 0.0, dt, threads)  # warm-up
        best = float("inf")
        for n in range(10):
            t0 = time.perf_counter()
            p.rk4_step(u, (n + 1) * dt, dt, threads)
            best = min(best, time.perf_counter() - t0)
        return points / best, best

    v1, t1 = min_step(TD.minkowski_noise_case(N=16, q=4), 16 ** 3, noise=True)
    out = {"cpu_runnable_block": {"value": v1, "step_ms": 1e3 * t1}}
    hole = TD.kerr_schild_hole_case(N=16, nblocks=4, halfwidth=2.0, a=0.7)
    v2, t2 = min_step(hole, 64 * 16 ** 3)
    return {"value": v2, "unit": "point-updates/s", "cores": threads, "kind": "port",
            "cpu": _cpu_model(), "flags": flags, "timing": "min of 10 steps after one warm-up",
            "configs": out,
            "sample": f"oracle RK4 step of the same case on the hole's 64 blocks ([-2,2]^3, h=1/16, "
                      f"{64 * 16 ** 3} points): min {1e3 * t2:.0f} ms over 10 steps, {threads} OpenMP "
                      f"threads, {flags}"}


def adjoint_bench(tgh, torch):
    """The differentiation workload: the a = 0.7 hole on 64 blocks of 16^3
    ([-2M, 2M]^3, h = M/16, q = 4, closed-form gauge source, Gaussian damping
    near 1/M far 1/(10M) width 3M, gamma2 = 0.1, the :damped layer), 200 RK4
    steps of dt = 1/100 to T = 2M, a checkpoint every 20 steps: the gradient
    of the masked L2 error with respect to (gamma0_near, gamma0_far, gamma2,
    rho_max, rho_ramp) by the checkpointed reverse sweep. The cost ratio is
    the forward sweep plus the reverse sweep (segment recomputation
    included) over the plain forward run (tgh_rk4's hipGraph) of the same
    200 steps."""
    case = tgh.kerr_schild_case(M=1.0, a=0.7, halfwidth=2.0, r_0=0.71, r_1=1.21, gauge_source="algebraic",
                                gamma2=0.1, rho_max=4.0, rho_ramp=0.5)
    case.gamma0 = tgh.GaussianDamping(near=1.0, far=0.1, width=3.0)
    p = tgh.GHProblem(tgh.gh_forest(None, case, N=16, roots=4), case, q=4)
    u = p.statevector()
    tgh.initial_state(u, p, 0.0)
    u0 = u.clone()
    dt, nsteps, every = 0.01, 200, 20
    tgh.adjoint(p, u, 0.0, dt, nsteps, every=every)  # the first call allocates the checkpoints
    u.copy_(u0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    J, grad, (fwd, rev) = tgh.adjoint(p, u, 0.0, dt, nsteps, every=every)
    total = (time.perf_counter() - t0) * 1e3
    u.copy_(u0)
    tgh.gh_solve(p, u, (0.0, nsteps * dt), dt=dt, graph=True)  # graph capture
    u.copy_(u0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    tgh.gh_solve(p, u, (0.0, nsteps * dt), dt=dt, graph=True)
    torch.cuda.synchronize()
    plain = (time.perf_counter() - t0) * 1e3
    return {"workload": "static Kerr-Schild hole M=1 a=0.7 on 64 blocks of 16^3 ([-2M,2M]^3, h=M/16), q=4, "
                        "closed-form gauge source, 200 RK4 steps of dt=1/100, checkpoints every 20; "
                        "J = masked L2 error at T=2M",
            "J": J, "grad": dict(zip(tgh.ADJOINT_PARAMS, grad)),
            "forward_sweep_ms": fwd, "reverse_sweep_ms": rev, "gradient_wall_ms": total,
            "plain_forward_ms": plain, "adjoint_over_forward": total / plain}


def moving_bench(tgh, torch, steps):
    """A time-dependent case where launch costs show: the boosted harmonic hole
    (translate(boost(Harmonic(M=1, a=0.3), -v), x0), v = 0.3 along x, x0 =
    0.25, the moving :damped core) on 64 blocks of 16^3 ([-2M, 2M]^3, h =
    M/16): the step's graph, captured once and replayed with the stage clock
    set before each replay, against the same launches without a graph."""
    L = tgh._load()
    case = tgh.harmonic_kerr_case(M=1.0, a=0.3, v=0.3, x0=0.25, axis=1, halfwidth=2.0, r_0=0.35, r_1=0.9)
    p = tgh.GHProblem(tgh.gh_forest(None, case, N=16, roots=4), case, q=4)
    u = p.statevector()
    tgh.initial_state(u, p, 0.0)
    dt = tgh.gh_dt(p, u, cfl=0.25)
    points = 64 * 16 ** 3
    out = {"workload": "boosted harmonic hole M=1 a=0.3 v=0.3 (moving core) on 64 blocks of 16^3, h=M/16, q=4",
           "steps": steps}
    for name, flags in (("graph", 1), ("launches", 0)):
        tgh._check(L.tgh_rk4(p.ptr, C.c_void_p(u.data_ptr()), 0.0, dt, 2, flags))  # warm-up, capture
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        tgh._check(L.tgh_rk4(p.ptr, C.c_void_p(u.data_ptr()), 2 * dt, dt, steps, flags))
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        out[name] = points * steps / el
        out[name + "_ms_per_step"] = el / steps * 1e3
    assert bool(torch.isfinite(u).all()), "the moving hole's state is not finite"
    out["graph_over_launches"] = out["graph"] / out["launches"]
    return out


# The batch workload: four Kerr-Schild holes a = 0, 0.3, 0.5, 0.7 (closed-form
# gauge source, damped interior r_0 = 0.71, r_1 = 1.21) on uneven forests of
# 7^3, 9^3, 11^3 and 13^3 blocks of 16^3 at h = M/16 (4600 blocks, 18.8 M points).
BATCH_HOLES = ((0.0, 7), (0.3, 9), (0.5, 11), (0.7, 13))


def batch_bench(tgh, torch, steps):
    """One batched hipGraph step over the four holes, against the same holes
    stepped one after another (each its own graph) and by the plain per-block
    form (one launch per block and stage)."""
    L = tgh._load()
    probs, us = [], []
    for a, roots in BATCH_HOLES:
        case = tgh.kerr_schild_case(M=1.0, a=a, halfwidth=roots / 2, r_0=0.71, r_1=1.21,
                                    gauge_source="algebraic")
        p = tgh.GHProblem(tgh.gh_forest(None, case, N=16, roots=roots), case, q=4)
        u = p.statevector()
        tgh.initial_state(u, p, 0.0)
        probs.append(p)
        us.append(u)
    dt = min(tgh.gh_dt(p, u, cfl=0.25) for p, u in zip(probs, us))
    points = sum(r ** 3 for _, r in BATCH_HOLES) * 16 ** 3
    batch = tgh.GHBatch(probs)
    uarr = (C.c_void_p * len(us))(*[u.data_ptr() for u in us])
    tgh._check(L.tgh_batch_rk4(batch.ptr, uarr, 0.0, dt, 2, 1))  # warm-up and capture
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    tgh._check(L.tgh_batch_rk4(batch.ptr, uarr, 2 * dt, dt, steps, 1))
    torch.cuda.synchronize()
    batched = points * steps / (time.perf_counter() - t0)
    for p, u in zip(probs, us):  # warm-up and capture of each problem's own graph
        tgh._check(L.tgh_rk4(p.ptr, C.c_void_p(u.data_ptr()), 0.0, dt, 1, 1))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for p, u in zip(probs, us):
        tgh._check(L.tgh_rk4(p.ptr, C.c_void_p(u.data_ptr()), 0.0, dt, steps, 1))
    torch.cuda.synchronize()
    sequential = points * steps / (time.perf_counter() - t0)
    # the plain form: each hole alone, its launches without a graph (one per
    # kernel per stage over all its blocks)
    t0 = time.perf_counter()
    for p, u in zip(probs, us):
        tgh._check(L.tgh_rk4(p.ptr, C.c_void_p(u.data_ptr()), 0.0, dt, steps, 0))
    torch.cuda.synchronize()
    plain = points * steps / (time.perf_counter() - t0)
    t0 = time.perf_counter()
    for p, u in zip(probs, us):
        tgh._check(L.tgh_rk4(p.ptr, C.c_void_p(u.data_ptr()), 0.0, dt, 1, tgh.TGH_RK4_PER_BLOCK))
    torch.cuda.synchronize()
    per_block = points / (time.perf_counter() - t0)
    assert all(bool(torch.isfinite(u).all()) for u in us), "a batch state is not finite"
    del batch
    return {"workload": "four static Kerr-Schild holes a = 0, 0.3, 0.5, 0.7 (closed-form gauge "
                        "source, damped interior) on 7^3, 9^3, 11^3, 13^3 blocks of 16^3 at h = M/16",
            "blocks": sum(r ** 3 for _, r in BATCH_HOLES), "points": points, "steps": steps,
            "batched_graph": batched, "sequential_graphs": sequential, "plain_launches": plain,
            "batched_over_plain": batched / plain, "per_block": per_block,
            "batched_over_per_block": batched / per_block}


def forest_bench(tgh, torch, steps):
    """The refined hierarchy of the batching workload: hole_forest around the
    a = 0.7 hole (three levels, finest h = 5M/256, 2416 leaves of 16^3), RK4 on
    the block layout with the planned fill, as one hipGraph per step (one RHS
    launch over all levels and one fill launch per pass) against the plain
    per-block form (one launch per leaf and stage); the fill's and the right-
    hand side's share from hipEvents."""
    L = tgh._load()
    case = tgh.kerr_schild_case(M=1.0, a=0.7, halfwidth=5.0, r_0=0.71, r_1=1.21, gauge_source="algebraic")
    forest = tgh.hole_forest(None, case, N=16, roots=8, radii=(3.0, 1.5))
    p = tgh.GHProblem(forest, case, q=4, ops=tgh.Operators(prolongation=6))
    u = p.statevector()
    tgh.initial_state(u, p, 0.0)
    dt = tgh.gh_dt(p, u, cfl=0.25)
    points = forest.nleaves() * 16 ** 3
    ptr = C.c_void_p(u.data_ptr())
    tgh._check(L.tgh_rk4(p.ptr, ptr, 0.0, dt, 1, 1))  # warm-up and capture
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    tgh._check(L.tgh_rk4(p.ptr, ptr, dt, dt, steps, 1))
    torch.cuda.synchronize()
    graph = points * steps / (time.perf_counter() - t0)
    # the plain form: the same launches without the graph (one launch per
    # kernel per stage over every block, as map_blocks! does)
    t0 = time.perf_counter()
    tgh._check(L.tgh_rk4(p.ptr, ptr, 0.0, dt, steps, 0))
    torch.cuda.synchronize()
    plain = points * steps / (time.perf_counter() - t0)
    tgh._check(L.tgh_profile(p.ptr, 1))
    tgh._check(L.tgh_rk4(p.ptr, ptr, 0.0, dt, 1, 0))
    prof = (C.c_double * 6)()
    tgh._check(L.tgh_profile_read(p.ptr, prof))
    tgh._check(L.tgh_profile(p.ptr, 0))
    rhs_ms = prof[0] / max(prof[1], 1)
    pi_ms = prof[4] / max(prof[5], 1)
    fill_ms = prof[2] / max(prof[3], 1)
    t0 = time.perf_counter()
    tgh._check(L.tgh_rk4(p.ptr, ptr, 0.0, dt, 1, tgh.TGH_RK4_PER_BLOCK))
    torch.cuda.synchronize()
    per_block = points / (time.perf_counter() - t0)
    assert bool(torch.isfinite(u).all()), "the forest's state is not finite"
    # one regrid: the leaves in the indicator's upper sixteenth refine once,
    # the level floor holds the horizon shell, nothing quiet coarsens
    # (marking, the new problem with its fill plan, the transfer), and the
    # transfer alone (its plan and kernel) onto the new forest again
    tau = tgh.refinement_indicator(p, u, dt, scale=tgh.field_scale(p, u, dt), eps=0.01)
    thr = sorted(tau)[len(tau) * 15 // 16]
    ref = tgh.Refinement(refine_tol=thr, coarsen_tol=min(t for t in tau if t > 0) / 2, maxlevel_cap=3,
                         floor_margin=1.0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    q, uq, _ = tgh.regrid(p, u, dt, refinement=ref)
    torch.cuda.synchronize()
    regrid_ms = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    tgh.transfer(p, u, dt, q, uq)
    torch.cuda.synchronize()
    transfer_ms = (time.perf_counter() - t0) * 1e3
    regrid_leaves = q.nblocks
    del q, uq
    levels = [sum(1 for lf in forest.leaves if lf[0] == k) for k in range(forest.maxlevel() + 1)]
    return {"workload": "static Kerr-Schild hole M=1 a=0.7 (closed-form gauge source, damped interior) on "
                        "hole_forest(N=16, roots=8 over [-5M,5M]^3, radii=(3M, 1.5M)): three levels, finest "
                        "h = 5M/256, prolongation p = 6, block layout",
            "leaves": forest.nleaves(), "leaves_per_level": levels, "points": points, "steps": steps,
            "graph": graph, "plain_launches": plain, "graph_over_plain": graph / plain,
            "per_block": per_block, "graph_over_per_block": graph / per_block,
            "rhs_ms": rhs_ms + pi_ms, "fill_ms": fill_ms,
            "fill_fraction": fill_ms / (fill_ms + rhs_ms + pi_ms),
            "regrid_ms": regrid_ms, "regrid_transfer_ms": transfer_ms, "regrid_leaves": regrid_leaves}


def moving_forest_bench(tgh, torch, steps):
    """The forward path of the regrid-cadence workload: the boosted Kerr-Schild
    hole (KerrSchildSource) M=1 a=0.7 at v=0.3 along x on hole_forest(N=16,
    roots=8 over [-16M,16M]^3, radii=(8M, 4M, 2M)): four levels, finest h=M/32,
    prolongation p=6, the block layout; RK4 steps as a graph and as plain
    launches, then the driver's cadence (chunks of 8 steps, each followed by
    a regrid: marking, the new problem and its fill plan, the transfer)."""
    L = tgh._load()
    case = tgh.kerr_schild_case(M=1.0, a=0.7, v=0.3, halfwidth=16.0, r_0=0.71, r_1=1.21,
                                gauge_source="algebraic")
    forest = tgh.hole_forest(None, case, N=16, roots=8, radii=(8.0, 4.0, 2.0))
    p = tgh.GHProblem(forest, case, q=4, ops=tgh.Operators(prolongation=6))
    u = p.statevector()
    tgh.initial_state(u, p, 0.0)
    dt = tgh.gh_dt(p, u, cfl=0.25)
    points = forest.nleaves() * 16 ** 3
    ptr = C.c_void_p(u.data_ptr())
    out = {"workload": "boosted Kerr-Schild hole M=1 a=0.7 v=0.3 (KerrSchildSource, moving damped core) on "
                       "hole_forest(N=16, roots=8 over [-16M,16M]^3, radii=(8M,4M,2M)): four levels, finest "
                       "h=M/32, p=6, block layout",
           "leaves": forest.nleaves(), "points": points, "steps": steps}
    for name, flags in (("graph", 1), ("plain_launches", 0)):
        tgh._check(L.tgh_rk4(p.ptr, ptr, 0.0, dt, 1, flags))  # warm-up, capture
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        tgh._check(L.tgh_rk4(p.ptr, ptr, dt, dt, steps, flags))
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        out[name] = points * steps / el
        out[name + "_ms_per_step"] = el / steps * 1e3
    assert bool(torch.isfinite(u).all()), "the moving forest's state is not finite"
    tau = tgh.refinement_indicator(p, u, dt, scale=tgh.field_scale(p, u, dt), eps=0.01)
    thr = sorted(tau)[len(tau) * 15 // 16]
    ref = tgh.Refinement(refine_tol=thr, coarsen_tol=min(t for t in tau if t > 0) / 2, maxlevel_cap=3,
                         floor_margin=1.0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    q, uq, log = tgh.gh_solve_regrid(p, u, (0.0, 16 * dt), cfl=0.25, every=8, refinement=ref, graph=True)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    assert bool(torch.isfinite(uq).all()), "the regridded state is not finite"
    out["cadence"] = {"steps": 16, "regrids": len(log), "wall_ms": el * 1e3,
                      "leaves_after": q.nblocks, "ms_per_step_with_regrids": el / 16 * 1e3}
    return out


def octree_batch_bench(tgh, torch, steps):
    """The batching workload whole: the three-level octree around the a = 0.7
    hole (forest_bench's hierarchy) and the four holes of batch_bench in one
    batch on the block layout (one fused launch over all 7016 blocks and every
    level per stage, the fills per problem, the step a hipGraph), against the
    same problems stepped one after another (each its own graph) and by the
    plain per-block form."""
    L = tgh._load()
    case = tgh.kerr_schild_case(M=1.0, a=0.7, halfwidth=5.0, r_0=0.71, r_1=1.21, gauge_source="algebraic")
    forest = tgh.hole_forest(None, case, N=16, roots=8, radii=(3.0, 1.5))
    probs = [tgh.GHProblem(forest, case, q=4, ops=tgh.Operators(prolongation=6))]
    for a, roots in BATCH_HOLES:
        c = tgh.kerr_schild_case(M=1.0, a=a, halfwidth=roots / 2, r_0=0.71, r_1=1.21, gauge_source="algebraic")
        probs.append(tgh.GHProblem(tgh.gh_forest(None, c, N=16, roots=roots), c, q=4))
    us = []
    for p in probs:
        u = p.statevector()
        tgh.initial_state(u, p, 0.0)
        us.append(u)
    dt = min(tgh.gh_dt(p, u, cfl=0.25) for p, u in zip(probs, us))
    blocks = sum(p.nblocks for p in probs)
    points = blocks * 16 ** 3
    batch = tgh.GHBatch(probs)
    uarr = (C.c_void_p * len(us))(*[u.data_ptr() for u in us])
    tgh._check(L.tgh_batch_rk4(batch.ptr, uarr, 0.0, dt, 1, 1))  # warm-up and capture
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    tgh._check(L.tgh_batch_rk4(batch.ptr, uarr, dt, dt, steps, 1))
    torch.cuda.synchronize()
    batched = points * steps / (time.perf_counter() - t0)
    flags = 4 | 2 | 1  # each alone as the batch runs it: block layout, fused, graph
    for p, u in zip(probs, us):
        tgh._check(L.tgh_rk4(p.ptr, C.c_void_p(u.data_ptr()), 0.0, dt, 1, flags))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for p, u in zip(probs, us):
        tgh._check(L.tgh_rk4(p.ptr, C.c_void_p(u.data_ptr()), 0.0, dt, steps, flags))
    torch.cuda.synchronize()
    sequential = points * steps / (time.perf_counter() - t0)
    t0 = time.perf_counter()
    for p, u in zip(probs, us):
        tgh._check(L.tgh_rk4(p.ptr, C.c_void_p(u.data_ptr()), 0.0, dt, steps, flags & ~1))
    torch.cuda.synchronize()
    plain = points * steps / (time.perf_counter() - t0)
    t0 = time.perf_counter()
    for p, u in zip(probs, us):
        tgh._check(L.tgh_rk4(p.ptr, C.c_void_p(u.data_ptr()), 0.0, dt, 1, tgh.TGH_RK4_PER_BLOCK))
    torch.cuda.synchronize()
    per_block = points / (time.perf_counter() - t0)
    assert all(bool(torch.isfinite(u).all()) for u in us), "a batch state is not finite"
    del batch
    return {"workload": "the three-level octree around the a = 0.7 hole (2416 leaves, finest h = 5M/256) and "
                        "four static Kerr-Schild holes a = 0, 0.3, 0.5, 0.7 on 7^3 .. 13^3 blocks, one batch on "
                        "the block layout",
            "blocks": blocks, "points": points, "steps": steps, "batched_graph": batched,
            "sequential_graphs": sequential, "plain_launches": plain, "batched_over_plain": batched / plain,
            "per_block": per_block, "batched_over_per_block": batched / per_block}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: every box here has one MI355X and this build has no "
              f"multi-GPU path (asked for --gpus {args.gpus})", file=sys.stderr)
        sys.exit(2)
    import torch
    import __graft_entry__
    tgh = __graft_entry__.load_package()
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    L = tgh._load()

    desc = kerr_schild_desc(tgh, args.nblocks)
    ptr = C.c_void_p()
    tgh._check(L.tgh_problem_create(C.byref(desc), C.byref(ptr)))
    # the library's work and torch's share one stream: the plain form below
    # interleaves them
    tgh._check(L.tgh_set_stream(ptr, C.c_void_p(torch.cuda.current_stream().cuda_stream)))
    n_state = int(L.tgh_state_length(ptr))
    u = torch.empty(n_state, dtype=torch.float64, device="cuda")
    uptr = C.c_void_p(u.data_ptr())
    tgh._check(L.tgh_initial_state(ptr, uptr, 0.0))
    lam = C.c_double()
    tgh._check(L.tgh_max_speed(ptr, uptr, 0.0, C.byref(lam)))
    h = 1.0 / 16
    dt = 0.25 * h / lam.value
    nb = args.nblocks ** 3
    points = nb * 16 ** 3

    form = FORMS[args.form]
    # the autotuner: each Pi tile shape timed on a scratch copy of the state,
    # the fastest kept for the steps below (all shapes give the same bits)
    tiles_ms = None
    if args.form == "brick-split" and not args.no_tune:
        scratch = u.clone()
        ms = (C.c_double * tgh.TGH_PI_TILES)()
        tgh._check(L.tgh_autotune(ptr, C.c_void_p(scratch.data_ptr()), 0.0, dt, ms))
        tiles_ms = list(ms)
        del scratch
    # warm-up (also captures the graph)
    note(f"{nb} blocks, {args.form}: warm-up")
    tgh._check(L.tgh_rk4(ptr, uptr, 0.0, dt, args.warmup, 1 | form))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    tgh._check(L.tgh_rk4(ptr, uptr, args.warmup * dt, dt, args.steps, 1 | form))
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    value = points * args.steps / el
    note(f"timed {args.steps} steps: {el / args.steps * 1e3:.2f} ms per step")
    assert bool(torch.isfinite(u).all()), "the state is not finite after the timed steps"

    # The same launches without the graph (one launch per kernel per stage
    # over the whole brick, as map_blocks! launches), wall-clock timed.
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    tgh._check(L.tgh_rk4(ptr, uptr, (args.warmup + args.steps) * dt, dt, args.steps, form))
    torch.cuda.synchronize()
    nograph_value = points * args.steps / (time.perf_counter() - t1)

    # Per-kernel durations from hipEvents on the problem's stream, over a
    # second pass of the same steps launched without the graph.
    tgh._check(L.tgh_profile(ptr, 1))
    tgh._check(L.tgh_rk4(ptr, uptr, (args.warmup + args.steps) * dt, dt, args.steps, form))
    prof = (C.c_double * 6)()
    tgh._check(L.tgh_profile_read(ptr, prof))
    tgh._check(L.tgh_profile(ptr, 0))
    head_ms = prof[0] / max(prof[1], 1)
    fill_ms = prof[2] / max(prof[3], 1)
    pi_ms = prof[4] / max(prof[5], 1)

    # The plain form: RK4 written with the verbs (gh_rhs! = prepare! + rhs!,
    # then the stage arithmetic as separate passes), as the reference's
    # integrator drives the kernel -- a few steps of it on the same state.
    plain_steps = 0 if args.no_plain else max(1, min(3, args.steps))
    k = [torch.empty_like(u) for _ in range(4)]
    y = torch.empty_like(u)

    def rhs_(du, uu, t):
        tgh._check(L.tgh_gh_rhs(ptr, C.c_void_p(du.data_ptr()), C.c_void_p(uu.data_ptr()), t, None))

    t = (args.warmup + 2 * args.steps) * dt
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for n in range(plain_steps):
        rhs_(k[0], u, t)
        torch.add(u, k[0], alpha=dt / 2, out=y)
        rhs_(k[1], y, t + dt / 2)
        torch.add(u, k[1], alpha=dt / 2, out=y)
        rhs_(k[2], y, t + dt / 2)
        torch.add(u, k[2], alpha=dt, out=y)
        rhs_(k[3], y, t + dt)
        u.add_(k[0] + 2 * k[1] + 2 * k[2] + k[3], alpha=dt / 6)
        t += dt
    torch.cuda.synchronize()
    plain_value = points * plain_steps / (time.perf_counter() - t1) if plain_steps else None

    # The plain per-block form: the fused block kernel launched once per block
    # and stage (4096 launches a stage), one step from the current state.
    per_block_value = None
    if not args.no_per_block:
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        tgh._check(L.tgh_rk4(ptr, uptr, t, dt, 1, tgh.TGH_RK4_PER_BLOCK))
        torch.cuda.synchronize()
        per_block_value = points / (time.perf_counter() - t1)
        t += dt

    f_rhs, f_head, f_combine = flop_count()
    f_pi = f_rhs - f_head
    # Per launch: F per point of that kernel's half over the N^3 points of each
    # block, plus its half of a quarter of the stage combination.
    if not args.form.endswith("split"):
        head_tf = (f_rhs + f_combine / 4) * points / (head_ms * 1e-3) / 1e12
        pi_tf = None
        dom = args.form
        achieved = head_tf
    else:
        head_tf = (f_head + f_combine / 8) * points / (head_ms * 1e-3) / 1e12
        pi_tf = (f_pi + f_combine / 8) * points / (pi_ms * 1e-3) / 1e12
        dom = args.form + (":head" if head_ms >= pi_ms else ":pi")
        achieved = head_tf if dom.endswith("head") else pi_tf
    step_ms = el / args.steps * 1e3
    t_roof = max((4 * f_rhs + f_combine) / (FP64_PEAK_TFLOPS * 1e12), B_POINT_UPDATE / (HBM_PEAK_GBS * 1e9))
    result = {
        "metric": "FP64 RK4 point-updates/s, 1 MI355X (batched vs plain); adjoint/forward cost",
        "value": value,
        "unit": "point-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": step_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (analytic Kerr-Schild initial data)",
        "config": {"workload": f"static Kerr-Schild hole M=1 a=0.7, [-8M,8M]^3 at h=M/16, "
                               f"{nb} blocks of 16^3, q=4, damped interior, KS closed-form gauge "
                               f"source, RK4 hipGraph step",
                   "blocks": nb, "points": points, "q": 4, "dt": dt},
        "roofline": {"bound": "fp64", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                     "traffic": measured_traffic(KERNEL[dom], nb),
                     "kernel": KERNEL[dom],
                     "kernel_ms": pi_ms if dom.endswith("pi") else head_ms,
                     "head_ms": head_ms, "head_tflops": head_tf, "pi_ms": pi_ms, "pi_tflops": pi_tf,
                     "fill_ms": fill_ms, "flops_per_point_rhs": f_rhs, "flops_per_point_head": f_head,
                     "flops_per_point_update": 4 * f_rhs + f_combine,
                     "step_roofline_frac": t_roof / (step_ms * 1e-3 / points),
                     "pi_tiles_ms": tiles_ms,
                     "pi_tile": (min(range(len(tiles_ms)), key=lambda i: tiles_ms[i]) if tiles_ms else 0)},
        "forms": {f"{args.form}_graph": value, "launches_without_graph": nograph_value,
                  "graph_over_launches": value / nograph_value, "plain_verbs": plain_value,
                  "per_block": per_block_value,
                  "graph_over_per_block": value / per_block_value if per_block_value else None},
    }
    # the two constraint monitors on the same state (each call: scatter, fill,
    # the monitor kernel and the fold), after one warm call
    monitors = {}
    l2m, lim = (C.c_double * 4)(), (C.c_double * 4)()
    for name, fn in (("gauge_constraint_ms", L.tgh_gauge_constraint), ("adm_constraint_ms", L.tgh_adm_constraint)):
        tgh._check(fn(ptr, uptr, t, l2m, lim))
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        tgh._check(fn(ptr, uptr, t, l2m, lim))
        monitors[name] = (time.perf_counter() - t2) * 1e3
    result["monitors"] = monitors
    L.tgh_problem_destroy(ptr)
    del k, y

    # The same workload with the gauge source sampled once at creation (the
    # reference's default H_a = -Gamma_a of the background) instead of the
    # closed form: the same stepper form, timed the same way.
    if not args.no_sampled:
        note("the sampled-source line")
        torch.cuda.empty_cache()
        ds = kerr_schild_desc(tgh, args.nblocks)
        ds.gauge = 1
        ps = C.c_void_p()
        tgh._check(L.tgh_problem_create(C.byref(ds), C.byref(ps)))
        tgh._check(L.tgh_set_stream(ps, C.c_void_p(torch.cuda.current_stream().cuda_stream)))
        tgh._check(L.tgh_initial_state(ps, uptr, 0.0))
        if tiles_ms:
            tgh._check(L.tgh_set_tile(ps, result["roofline"]["pi_tile"]))
        tgh._check(L.tgh_rk4(ps, uptr, 0.0, dt, args.warmup, 1 | form))
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        tgh._check(L.tgh_rk4(ps, uptr, args.warmup * dt, dt, args.steps, 1 | form))
        torch.cuda.synchronize()
        el_s = time.perf_counter() - t1
        assert bool(torch.isfinite(u).all()), "the sampled-source state is not finite after the timed steps"
        result["sampled_source"] = {"value": points * args.steps / el_s, "ms_per_step": el_s / args.steps * 1e3,
                                    "workload": "the same hole with H_a sampled at creation (20 more fields "
                                                "read per point and stage)"}
        L.tgh_problem_destroy(ps)
    if not args.no_adjoint:
        note("the adjoint's cost")
        torch.cuda.empty_cache()
        result["adjoint"] = adjoint_bench(tgh, torch)
    note("the moving hole")
    torch.cuda.empty_cache()
    result["moving"] = moving_bench(tgh, torch, max(1, args.steps))
    if not args.no_forest:
        note("the octree")
        torch.cuda.empty_cache()
        result["forest"] = forest_bench(tgh, torch, max(1, min(5, args.steps)))
    if not args.no_forest:
        note("the moving hole on its forest")
        torch.cuda.empty_cache()
        result["moving_forest"] = moving_forest_bench(tgh, torch, max(1, min(5, args.steps)))
    if not args.no_batch:
        note("the four-hole batch")
        torch.cuda.empty_cache()
        result["batch"] = batch_bench(tgh, torch, max(1, min(5, args.steps)))
        if not args.no_forest:
            note("the octree with the four holes")
            torch.cuda.empty_cache()
            result["octree_batch"] = octree_batch_bench(tgh, torch, max(1, min(3, args.steps)))
    if not args.no_cpu:
        note("the CPU baseline")
        result["cpu_baseline"] = cpu_baseline()
    if args.calibrate_copy:
        x = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        y.copy_(x)
        torch.cuda.synchronize()
        del x, y
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
