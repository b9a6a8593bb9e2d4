"""Benchmark of the PyQuake3D hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

Workloads
  hstep     (default) H-matrix time stepping at N = 100,352 triangles on the
            undulating 100 x 25 km fault with BP5-like heterogeneous friction
            (leaf 64, eta 2, ACA eps 1e-4). A step is one accepted RK45 time
            step (one six-operator H-matvec, the RK45 attempts, commit). The
            metric is simulated years per wall-hour over K timed steps after W
            warm-up steps, all from the same seeded initial state:
              plain   : FP64 operator storage, one simulation
              fast    : FP32 operator storage, one simulation
              batched : FP32 storage, --nb friction draws (default 16) on the
                        same mesh stepped together, one multi-RHS H-matvec
                        per step; rate = sum over the draws
            ``value`` is the batched variant; all three are reported.
  example   The reference's BP5-QD example model (N = 9,250, half space)
            from its initial state: dense FP64 vs FP64 / FP32 H-matrices,
            event times and simulated years per wall-hour.
  accuracy  Event times of the hstep problem with FP32 and with FP64
            operator storage (one draw, device controller): the relative
            event-time error behind "matched accuracy".
  ragged    --count simulations (default 64) on meshes of uneven size as
            one block-diagonal FP32 H-matrix with the device controller;
            per-instance accuracy of every 8th against its FP64 solo run.
  learned   Learned step size against the reference controller at equal
            rtol (--nb draws, default 8): rejected-step fraction and wall
            time of each over --steps steps from the same state.
  matvec    H-matvec only, --nb right-hand sides, on the hstep operator
            (--dtype f32|f64); the command the hardware-counter profiles use.
  assembly  Kernel assembly only: all N x N pairs at N = 20,000 (FP64, six
            traction operators, 19.2 GB of output in HBM). One step = one full
            assembly.
  disp      Surface displacement: a 256 x 256 grid on the free surface
            (x in [-60, 60] km, y in [-20, 20] km) from every element of the
            N = 20,000 undulating fault with seeded slip, half space, FP64
            (pq3d_tde_disp_sum). One step = one full grid.

Prints ONE JSON line. Inputs are resident in HBM before the timed region;
the timed region is bracketed by device synchronisation; kernel times come
from HIP events recorded on the launch stream.
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO))

MI355X_HBM_GBS = 8000.0
MI355X_FP64_TFLOPS = 78.6
YEAR = 3.15576e7

# HBM bytes per H-matvec (all three phases) at N = 100,352 from two
# rocprofv3 --pmc passes (FETCH_SIZE, WRITE_SIZE) of the matvec workload,
# FETCH scaled by 2.0 and WRITE by 1.0: the factors the same counters read
# for a 2 GiB device-to-device copy (profiles/hmatvec_traffic_N100352.md).
MEASURED_TRAFFIC = {("f32", 1): 6_276_266_539, ("f32", 16): 7_350_784_928}

# FP64 operations of one (receiver, source) pair in tde_pair_traction_sd
# (unit strike and dip slip, both projected), counted operator by operator
# from pyquake3d_amd/csrc/tde_device.h: per angular dislocation 191 (brackets)
# + 8 (side frame) + 2 x (55 prefactors/combination + 34 side rotation) + 4
# (slip rotation), three of them, plus 18 (TD transform) + 28 (configuration)
# + 2 x (129 global rotation + 20 Hooke + 30 projection) + 24 (summing the
# three angular dislocations); divisions and square roots count as one.
FLOPS_PER_PAIR = 1571.0
# displacement of one (point, triangle) pair, counted per operation in the
# device source (+ - * / sqrt log atan atan2 each 1): full space 570 (three
# angular dislocations 3 x 127, their frame rotations 3 x 25, point transform
# and configuration 46, Burgers function and output ~68); half space on a
# vertical ruled fault 3,370 (main + image 2 x 570, two corrected sides
# 2 x 1,105, slip vector and sums 21)
DISP_FLOPS_PER_PAIR_HS = 3370.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


# ---------------------------------------------------------------- hstep
def build_hmatrix(problem, dtype):
    import torch

    from pyquake3d_amd.hmatrix import HMatrix
    t0 = time.perf_counter()
    H = HMatrix.from_mesh(problem.mesh, problem.mu, problem.lambda_, leaf=64, eta=2.0, eps=1e-4,
                          dtype=dtype, device=torch.device("cuda:0"))
    torch.cuda.synchronize()
    return H, time.perf_counter() - t0


def run_variant(problems, H, steps, warmup, name, rounds=8, graph=False):
    """``problems`` (NB friction draws on one mesh) stepped by DeviceEnsemble:
    one NB-RHS H-matvec, ``rounds`` RK45 attempt rounds with the reference's
    accept/reject decision on the device, and the commit per iteration, with
    no host synchronisation inside the timed region. The rate is the sum of
    the draws' simulated years per wall-hour."""
    import torch

    from pyquake3d_amd.qdsim import DeviceEnsemble

    ens = DeviceEnsemble(problems, H, device=H.device, rounds=rounds)
    for _ in range(warmup):
        ens.step()
    torch.cuda.synchronize()
    ens.sync()
    t_sim0 = ens.time.copy()
    acc0 = ens.naccepted.copy()
    if graph:
        # every timed step replays one recorded HIP graph of the step
        ens.capture()
    else:
        ens.timing = True
    ens.matvec_events = []
    w0 = time.perf_counter()
    for _ in range(steps):
        ens.step()
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    ens.sync()
    mv = ens.matvec_ms()
    yrs = (ens.time - t_sim0) / YEAR
    nb = len(problems)
    acc = ens.naccepted - acc0
    res = dict(name=name, dtype=str(H.dtype).replace("torch.", ""), nb=nb, rounds=rounds,
               sim_years=float(yrs.sum()), wall_s=wall, ms_per_step=1e3 * wall / steps,
               accepted_steps_per_draw=float(acc.mean()),
               years_per_hour=float(yrs.sum()) / (wall / 3600.0),
               matvec_ms=sum(mv) / len(mv) if mv else None,
               matvec_bytes=H.matvec_bytes + 5 * 8 * H.n * (nb - 1),
               t_end_years=float((ens.time / YEAR).mean()),
               max_slipv=float(ens.max_slipv().max()))
    log(f"{name}: {json.dumps(res)}")
    return res, ens


def cpu_baseline_hstep(problem, H, gpu_res, nsteps=4):
    """Reference-style CPU step (per-block numpy FP64 matvec + numpy RK45) on
    the same blocks, timed over ``nsteps`` steps from the initial state; the
    rate is expressed for the GPU run's simulated interval."""
    import numpy as np

    from oracle.hmat_cpu import CpuHMatrix
    from oracle.qdsim_oracle import OracleSim
    t0 = time.perf_counter()
    C = CpuHMatrix(H)
    setup = time.perf_counter() - t0
    perm = H.perm
    init = [np.asarray(x)[perm] for x in problem.initial_state()]
    g = lambda x: np.asarray(x)[perm]
    ora = OracleSim(None, g(problem.a), g(problem.b), g(problem.dc), g(problem.f0), problem.V0,
                    problem.mu, problem.Cs, g(problem.Vpl), g(problem.rake), init,
                    matvec=C.matvec)
    dt = 1e-3
    t0 = time.perf_counter()
    for _ in range(nsteps):
        _, dt = ora.simu_forward(dt)
    per_step = (time.perf_counter() - t0) / nsteps
    steps = gpu_res["wall_s"] / (gpu_res["ms_per_step"] / 1e3)
    cpu_hours = per_step * steps / 3600.0
    return {"value": gpu_res["sim_years"] / cpu_hours, "unit": "sim-years/wall-hour",
            "cores": 1, "kind": "port",
            "sample": f"{nsteps} steps of the reference-style CPU step (per-block numpy FP64 "
                      f"matvec over the same {len(C.dense)} dense + {len(C.lowrank)} low-rank "
                      f"blocks, numpy RK45) at {per_step:.2f} s/step, applied to the GPU run's "
                      f"{steps:.0f} steps / {gpu_res['sim_years']:.2f} simulated years; "
                      f"host setup {setup:.1f} s"}


def bench_hstep(args):
    import torch

    from pyquake3d_amd.problem import undulating_c3

    problem = undulating_c3()
    log(f"hstep: N={problem.n}")
    H64, b64 = build_hmatrix(problem, torch.float64)
    plain, _ = run_variant([problem], H64, args.steps, args.warmup, "plain_fp64")
    plain["build_s"] = b64
    del H64
    torch.cuda.empty_cache()
    H32, b32 = build_hmatrix(problem, torch.float32)
    fast, _ = run_variant([problem], H32, args.steps, args.warmup, "fast_fp32")
    fast["build_s"] = b32
    # friction draws 0..NB-1 on the same mesh; draw 0 is the single-run problem
    draws = [problem] + [undulating_c3(friction_seed=k, mesh=problem.mesh)
                         for k in range(1, args.nb)]
    batched, _ = run_variant(draws, H32, args.steps, args.warmup, f"batched_fp32_nb{args.nb}")
    head = batched
    graphed = None
    if not args.no_graph:
        # the same ensemble, each step replaying one recorded HIP graph
        graphed, _ = run_variant(draws, H32, args.steps, args.warmup,
                                 f"batched_fp32_nb{args.nb}_graph", graph=True)
        graphed["matvec_ms"] = batched["matvec_ms"]    # same kernels, timed in the eager run
        head = graphed
    batched_mfma = None
    if args.nb == 16:
        # the same ensemble with the 16-RHS block products on FP32 matrix cores
        H32.accumulate = "mfma"
        batched_mfma, _ = run_variant(draws, H32, args.steps, args.warmup, "batched_fp32_mfma")
        H32.accumulate = "f64"
        if args.accumulate == "mfma":
            head = batched_mfma
    mf = head is batched_mfma
    gr = head is graphed
    achieved = head["matvec_bytes"] / (head["matvec_ms"] * 1e-3) / 1e9
    res = {
        "metric": "sim-years/wall-hour at N=1e5 (fp32+batched+learned dt vs plain fp64)",
        "value": head["years_per_hour"],
        "unit": "sim-years/wall-hour",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("f32 operator storage and matrix-core products (f32 sums per item group), "
                  "f64 low-rank V products, state and RK45" if mf else
                  "f32 operator storage, f64 state and accumulation"),
        "data": "synthetic undulating fault mesh and seeded friction fields",
        "config": {"workload": f"H-matrix RK45 time stepping, undulating fault 100x25 km, "
                               f"N={problem.n}, leaf 64, eta 2, ACA eps 1e-4, BP5-like friction, "
                               f"{args.steps} steps after {args.warmup} from the seeded initial "
                               f"state (a step: one accepted RK45 step of every draw; a draw "
                               f"that needs more than 8 attempts finishes in the next step)",
                   "variant": (("batched_fp32_mfma" if mf else "batched_fp32") +
                               f": {args.nb} friction draws sharing the operator, "
                               f"one {args.nb}-RHS H-matvec per step"
                               + (" (dense and U products on FP32 matrix cores)" if mf else "")
                               + (", each step one replay of a recorded HIP graph" if gr else "")
                               + ", reference step controller per draw run on the device "
                                 "(learned step size not in this line)")},
        "speedup_vs_plain_fp64": head["years_per_hour"] / plain["years_per_hour"],
        "plain_fp64": plain,
        "fast_fp32": fast,
        "batched_fp32": batched,
        "batched_fp32_mfma": batched_mfma,
        "batched_fp32_graph": graphed,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": MI355X_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / MI355X_HBM_GBS,
                     "traffic": MEASURED_TRAFFIC.get(("f32-mfma" if mf else "f32", args.nb)),
                     "kernel": f"H-matvec, {args.nb} right-hand sides (k_hmv_phase1 + "
                               f"k_hmv_phase1b + " + ("k_hmv_phase2_mfma16)" if mf else "k_hmv_phase2)"),
                     "kernel_ms": head["matvec_ms"], "algorithmic_bytes": head["matvec_bytes"]},
    }
    if not args.no_cpu_baseline:
        res["cpu_baseline"] = cpu_baseline_hstep(problem, H32, fast)
    return res


# ------------------------------------------------------------- matvec
def bench_matvec(args):
    """One H-matvec of --nb right-hand sides per step on the hstep operator,
    timed with HIP events on the launch stream; reports the effective HBM
    bandwidth over the algorithmic bytes (stored operator once + five FP64
    vectors) and the per-phase split used to read hardware counters."""
    import torch

    from pyquake3d_amd.problem import undulating_c3

    # a plain 2 GiB device copy (read 2 GiB, write 2 GiB): the known-size
    # reference for the hardware byte counters and the attainable bandwidth
    src = torch.ones(2 ** 28, dtype=torch.float64, device="cuda:0")
    dst = torch.empty_like(src)
    c0, c1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    dst.copy_(src)
    c0.record()
    for _ in range(3):
        dst.copy_(src)
    c1.record()
    torch.cuda.synchronize()
    copy_gbs = 2 * src.numel() * 8 * 3 / (c0.elapsed_time(c1) * 1e-3) / 1e9
    del src, dst
    torch.cuda.empty_cache()

    problem = undulating_c3()
    dtype = torch.float32 if args.dtype == "f32" else torch.float64
    H, build_s = build_hmatrix(problem, dtype)
    H.accumulate = args.accumulate
    n, nb, dev = H.n, args.nb, H.device
    gen = torch.Generator(device=dev)
    gen.manual_seed(0)
    v1 = torch.randn(n * nb, dtype=torch.float64, device=dev, generator=gen)
    v2 = torch.randn(n * nb, dtype=torch.float64, device=dev, generator=gen)
    out = torch.empty((3, n * nb), dtype=torch.float64, device=dev)
    for _ in range(args.warmup):
        H.matvec(v1, v2, out=out, nrhs=nb)
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(args.steps)]
    w0 = time.perf_counter()
    for e0, e1 in evs:
        e0.record()
        H.matvec(v1, v2, out=out, nrhs=nb)
        e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    ms = sum(e0.elapsed_time(e1) for e0, e1 in evs) / len(evs)
    B = H.blocks
    lr, k = B["lr_idx"], B["ranks"]
    v_bytes = int((2 * B["n"][lr] * k).sum()) * H.esize
    u_bytes = int((3 * B["m"][lr] * k).sum()) * H.esize
    d_bytes = int(H.stats.dense_entries) * H.esize
    vec_bytes = 5 * 8 * n * nb
    algo = H.matvec_bytes + 5 * 8 * n * (nb - 1)
    achieved = algo / (ms * 1e-3) / 1e9
    return {
        "metric": "H-matvec effective HBM bandwidth",
        "value": achieved, "unit": "GB/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": (f"{args.dtype} operator storage, f32 matrix-core products for the dense and U "
                  f"blocks, f64 V products and vectors"
                  if (H.accumulate == "mfma" and nb == 16 and args.dtype == "f32") else
                  f"{args.dtype} operator storage, f64 vectors and accumulation"),
        "data": "synthetic undulating fault mesh, seeded random slip-rate vectors",
        "config": {"workload": f"H-matvec, undulating fault 100x25 km, N={n}, leaf 64, eta 2, "
                               f"ACA eps 1e-4, {nb} right-hand sides"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": MI355X_HBM_GBS,
                     "unit": "GB/s", "frac": achieved / MI355X_HBM_GBS,
                     "traffic": MEASURED_TRAFFIC.get((args.dtype, nb)),
                     "kernel_ms": ms, "algorithmic_bytes": algo},
        "bytes": {"phase1_lowrank_V": v_bytes, "phase2_lowrank_U": u_bytes,
                  "phase2_dense": d_bytes, "vectors": vec_bytes, "total": algo},
        "copy_2GiB_GBs": copy_gbs,
        "build_s": build_s,
    }


# ------------------------------------------------------------- learned
def _controller_run(ens, steps):
    import torch
    a0, s0, r0, t0 = ens.nattempts, ens.nsteps, int(ens.nrejects.sum()), ens.time.copy()
    w0 = time.perf_counter()
    for _ in range(steps):
        ens.step()
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    attempts = ens.nattempts - a0
    rejected = int(ens.nrejects.sum()) - r0
    accepted = (ens.nsteps - s0) * ens.nb
    yrs = float(((ens.time - t0) / YEAR).sum())
    return {"steps": ens.nsteps - s0, "attempt_rounds": attempts,
            "accepted_steps": accepted, "rejected_steps": rejected,
            "rejected_fraction": rejected / (accepted + rejected),
            "attempt_rounds_per_step": attempts / max(ens.nsteps - s0, 1),
            "wall_s": wall, "sim_years": yrs, "years_per_hour": yrs / (wall / 3600.0)}


def bench_learned(args):
    """Learned first-attempt step size against the reference controller at
    equal rtol (1e-4) on the hstep problem: --nb friction draws, FP32
    H-matrix at N = 100,352. The reference controller runs --record steps
    and logs (features, log10 h_accepted/dtnext) pairs; the MLP (hand-written
    forward/backward/Adam kernels) trains on them; then both controllers run
    --steps steps from the same saved state. Both loops take their accept or
    reject decision on the host after every attempt round, so the wall times
    compare the controllers, not the drivers."""
    import torch

    from pyquake3d_amd.learned_dt import FeatureEnsemble, StepMLP
    from pyquake3d_amd.problem import undulating_c3

    problem = undulating_c3()
    draws = [problem] + [undulating_c3(friction_seed=k, mesh=problem.mesh)
                         for k in range(1, args.nb)]
    H, _ = build_hmatrix(problem, torch.float32)
    rec = FeatureEnsemble(draws, H, device=H.device)
    rec.record = True
    for _ in range(args.record):
        rec.step()
    X, T = rec.dataset()
    mlp = StepMLP(device=H.device, seed=0)
    t0 = time.perf_counter()
    hist = mlp.fit(X, T, epochs=40, batch=1024, lr=2e-3)
    train_s = time.perf_counter() - t0
    saved = rec.save_state()
    rec.record = False
    plain = _controller_run(rec, args.steps)
    le = FeatureEnsemble(draws, H, mlp=mlp, device=H.device)
    le.load_state(saved)
    learned = _controller_run(le, args.steps)
    learned.update(first_attempt_rejected=le.nfirst_rejected, pi_fallback_attempts=le.nfallback,
                   max_accepted_error_ratio=le.max_accepted_ratio)
    log(f"learned: plain {json.dumps(plain)} learned {json.dumps(learned)}")
    return {
        "metric": "wall-time ratio, plain / learned step-size controller at equal rtol "
                  "(N=1e5, rejected-step fractions alongside)",
        "value": plain["wall_s"] / learned["wall_s"], "unit": "x",
        "n_gpus": 1, "steps": args.steps, "warmup": args.record,
        "ms_per_step": 1e3 * learned["wall_s"] / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 operator storage, f64 state; f32 MLP",
        "data": "synthetic undulating fault mesh and seeded friction fields",
        "config": {"workload": f"learned step size, undulating fault 100x25 km, N={problem.n}, "
                               f"{args.nb} friction draws, rtol 1e-4, {args.record} recorded "
                               f"steps ({len(T)} samples), then {args.steps} steps per "
                               f"controller from the same state"},
        "plain": plain, "learned": learned,
        "training": {"samples": int(len(T)), "epochs": 40, "loss_first": float(hist[0]),
                     "loss_last": float(hist[-1]), "train_s": train_s},
    }


# ------------------------------------------------------------- ragged
def bench_ragged(args):
    """Ragged ensemble: --count simulations (default 64) on meshes of uneven
    size (N log-uniform in [2000, 20000], planar or undulating 40 x 20 km
    faults, own friction draws), one block-diagonal FP32 H-matrix
    (MultiHMatrix: every matvec phase is one launch for all instances), the
    device controller, --steps iterations after --warmup. Accuracy: every
    8th instance is re-run alone in FP64 (and in FP32) to the same accepted
    step; the report gives its event times (max V through 1e-3 m/s) and
    final simulated time against the FP64 run."""
    import numpy as np
    import torch

    from pyquake3d_amd.hmatrix import HMatrix, MultiHMatrix
    from pyquake3d_amd.problem import ensemble_c4
    from pyquake3d_amd.qdsim import DeviceEnsemble

    dev = torch.device("cuda:0")
    probs = ensemble_c4(args.count, seed=0)
    t0 = time.perf_counter()
    hs = [HMatrix.from_mesh(p.mesh, p.mu, p.lambda_, leaf=64, eta=2.0, eps=1e-4,
                            dtype=torch.float32, device=dev) for p in probs]
    M = MultiHMatrix(hs)
    torch.cuda.synchronize()
    build_s = time.perf_counter() - t0
    log(f"ragged: {len(probs)} instances, N total {M.n}, built in {build_s:.1f} s")
    ens = DeviceEnsemble(probs, M, device=dev, rounds=8)
    for _ in range(args.warmup):
        ens.step()
    ens.sync()
    t_sim0 = ens.time.copy()
    ens.timing = True
    ens.matvec_events = []
    w0 = time.perf_counter()
    for _ in range(args.steps):
        ens.step()
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    ens.sync()
    mv = ens.matvec_ms()
    yrs = (ens.time - t_sim0) / YEAR

    def solo(k, H):
        """Instance k alone to the ensemble's accepted-step count; event
        times = upward crossings of max V through 1e-3 m/s (re-armed below
        1e-4 m/s)."""
        one = DeviceEnsemble([probs[k]], H, device=dev, rounds=8)
        events, armed = [], True
        while one.naccepted[0] < ens.naccepted[k]:
            one.step()
            one.sync()
            vmax = float(one.max_slipv()[0])
            if armed and vmax > 1e-3:
                events.append(float(one.time[0]))
                armed = False
            elif vmax < 1e-4:
                armed = True
        return one.time[0], one.instance("slipv", 0), events

    acc = []
    for k in range(0, len(probs), 8):
        v = ens.instance("slipv", k)
        # the FP32 solo run equals the ensemble's instance bit for bit (checked
        # on the final state), so its event times are the ensemble's
        t32, v32, ev32 = solo(k, hs[k])
        H64 = HMatrix.from_mesh(probs[k].mesh, probs[k].mu, probs[k].lambda_, leaf=64, eta=2.0,
                                eps=1e-4, dtype=torch.float64, device=dev)
        t64, v64, ev64 = solo(k, H64)
        del H64
        ne = min(len(ev32), len(ev64))
        acc.append({"instance": k, "n": probs[k].n, "accepted_steps": int(ens.naccepted[k]),
                    "equals_fp32_solo": bool(t32 == ens.time[k] and np.array_equal(v32, v)),
                    "events_fp32": len(ev32), "events_fp64": len(ev64),
                    "event_time_rel_err_vs_fp64": [abs(a - b) / b for a, b in
                                                   zip(ev32[:ne], ev64[:ne])],
                    "final_time_rel_err_vs_fp64": abs(ens.time[k] - t64) / t64})
    algo = M.matvec_bytes
    kern_ms = sum(mv) / len(mv)
    achieved = algo / (kern_ms * 1e-3) / 1e9
    return {
        "metric": "sim-years/wall-hour, ragged ensemble (FP32 H-matrix, device controller)",
        "value": float(yrs.sum()) / (wall / 3600.0), "unit": "sim-years/wall-hour",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 operator storage, f64 state and accumulation",
        "data": "synthetic planar/undulating meshes and seeded friction draws",
        "config": {"workload": f"ragged ensemble of {len(probs)} simulations, N log-uniform in "
                               f"[2000, 20000] (total {M.n}), leaf 64, eta 2, ACA eps 1e-4"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": MI355X_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / MI355X_HBM_GBS, "traffic": None, "kernel_ms": kern_ms,
                     "algorithmic_bytes": algo},
        "sizes": [int(p.n) for p in probs], "build_s": build_s,
        "accepted_steps_min_max": [int(ens.naccepted.min()), int(ens.naccepted.max())],
        "accuracy_sampled": acc,
    }


# ------------------------------------------------------------- accuracy
def bench_accuracy(args):
    """Event times of the hstep problem (draw 0, N = 100,352; Dc scaled by
    --dc-scale) with FP32 and with FP64 operator storage: the device
    controller, one accepted step per iteration, max slip rate read after
    every step; an event starts when it rises through --vup (armed once it
    has been below --vdown). Runs until --events events (0: no cap), --tmax years,
    --steps steps, or the reference's abort rule (dt below 1e-15 s or more
    than 20 rejections, src/QDsim.py:956-958), whichever comes first. Every
    FP64 event is paired with the nearest FP32 event.

    With Dc as given (0.14 m mean) this fault produces slow-slip transients
    (max V ~1e-8..2e-8 m/s) and no seismic rates within 600 years: use
    --vup 1e-8 --vdown 3e-9 (the defaults). With --dc-scale 0.7 a seismic
    event nucleates after ~2.4 years (--vup 1e-3 --vdown 1e-4); the
    reference controller then aborts during the dynamic phase in either
    precision."""
    import numpy as np
    import torch
    from dataclasses import replace

    from pyquake3d_amd.problem import undulating_c3
    from pyquake3d_amd.qdsim import DeviceEnsemble, StepAborted

    problem = undulating_c3()
    if args.dc_scale != 1.0:
        problem = replace(problem, dc=problem.dc * args.dc_scale)
    runs = {}
    for name, dtype in (("fp64", torch.float64), ("fp32", torch.float32)):
        H, _ = build_hmatrix(problem, dtype)
        ens = DeviceEnsemble([problem], H, device=H.device, rounds=8)
        # armed once max V is below --vdown (the initial nucleation patch
        # starts above a slow-slip threshold)
        events, armed, peak, status = [], False, 0.0, "ok"
        w0 = time.perf_counter()
        for i in range(args.steps):
            ens.step()
            try:
                ens.sync()
            except StepAborted:
                status = f"reference abort rule at {float(ens.time[0]) / YEAR:.4f} yr"
                log(f"accuracy {name}: {status}")
                break
            vmax = float(ens.max_slipv()[0])
            peak = max(peak, vmax)
            if armed and vmax > args.vup:
                events.append(float(ens.time[0]) / YEAR)
                armed = False
                log(f"accuracy {name}: event {len(events)} at {events[-1]:.4f} yr, step {i}")
                if 0 < args.events <= len(events):
                    break
            elif vmax < args.vdown:
                armed = True
            if ens.time[0] / YEAR > args.tmax:
                break
            if i % 5000 == 4999:
                log(f"accuracy {name}: step {i + 1}, t = {ens.time[0] / YEAR:.2f} yr, "
                    f"max V {vmax:.3e}")
        wall = time.perf_counter() - w0
        runs[name] = {"event_years": events, "steps": int(ens.naccepted[0]),
                      "sim_years": float(ens.time[0]) / YEAR, "wall_s": wall,
                      "peak_slip_rate": peak, "status": status}
        del H, ens
        torch.cuda.empty_cache()
    e64, e32 = runs["fp64"]["event_years"], runs["fp32"]["event_years"]
    pairs = []
    for t64 in e64:
        if e32:
            t32 = min(e32, key=lambda t: abs(t - t64))
            pairs.append({"fp64_yr": t64, "fp32_yr": t32, "rel_err": abs(t32 - t64) / t64})
    unmatched32 = [t for t in e32 if all(abs(t - p["fp32_yr"]) > 0 for p in pairs)]
    return {
        "metric": "event-time relative error, FP32 vs FP64 operator storage (N=1e5)",
        "value": max(p["rel_err"] for p in pairs) if pairs else None, "unit": "relative",
        "n_gpus": 1, "steps": args.steps, "warmup": 0,
        "ms_per_step": 1e3 * runs["fp32"]["wall_s"] / max(runs["fp32"]["steps"], 1),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 vs f64 operator storage, f64 state",
        "data": "synthetic undulating fault mesh and seeded friction fields",
        "config": {"workload": f"event times, undulating fault 100x25 km, N={problem.n}, leaf 64, "
                               f"eta 2, ACA eps 1e-4, rtol 1e-4, Dc x {args.dc_scale:g}, event = "
                               f"max V up through {args.vup:g} m/s (armed below "
                               f"{args.vdown:g}), up to {args.events or 'any number of'} events / "
                               f"{args.tmax} yr / "
                               f"{args.steps} steps (host reads max V every step)"},
        "pairs": pairs, "fp32_events_without_fp64_partner": unmatched32,
        "fp64": runs["fp64"], "fp32": runs["fp32"],
    }


# ------------------------------------------------------------- example
def bench_example(args):
    """The reference's BP5-QD example model (examples/BP5-QD: 9,250
    triangles, half space, nucleation disc at 3e-2 m/s; committed as
    tests/golden/bp5_example.npz) run from its initial state for --steps
    accepted steps with the device controller, three ways: dense FP64
    operators (the reference's plain method), FP64 and FP32 H-matrices (leaf
    32, eta 2, ACA eps 1e-6). Max slip rate and time are copied on the
    device after every step (no host synchronisation in the timed loop);
    events are up-crossings of 1e-3 m/s (re-armed below 1e-4 m/s)."""
    import numpy as np
    import torch

    from pyquake3d_amd.hmatrix import HMatrix
    from pyquake3d_amd.problem import load_problem
    from pyquake3d_amd.qdsim import DenseOperator, DeviceEnsemble, StepAborted

    p, hs = load_problem(str(REPO / "tests" / "golden" / "bp5_example.npz"))
    dev = torch.device("cuda:0")
    builds = (("dense_fp64", lambda: DenseOperator(p.mesh, p.mu, p.lambda_, device=dev,
                                                   halfspace=hs)),
              ("hmatrix_fp64", lambda: HMatrix.from_mesh(p.mesh, p.mu, p.lambda_, leaf=32,
                                                         eta=2.0, eps=1e-6, dtype=torch.float64,
                                                         device=dev, halfspace=hs)),
              ("hmatrix_fp32", lambda: HMatrix.from_mesh(p.mesh, p.mu, p.lambda_, leaf=32,
                                                         eta=2.0, eps=1e-6, dtype=torch.float32,
                                                         device=dev, halfspace=hs)))
    runs = {}
    for name, build in builds:
        t0 = time.perf_counter()
        op = build()
        torch.cuda.synchronize()
        build_s = time.perf_counter() - t0
        ens = DeviceEnsemble([p], op, device=dev, rounds=8)
        vh = torch.zeros(args.steps, dtype=torch.int64, device=dev)
        th = torch.zeros(args.steps, dtype=torch.float64, device=dev)
        w0 = time.perf_counter()
        for i in range(args.steps):
            ens.step()
            vh[i].copy_(ens.vmax[0])
            th[i].copy_(ens.tdev[0])
        torch.cuda.synchronize()
        wall = time.perf_counter() - w0
        status = "ok"
        try:
            ens.sync()
        except StepAborted as e:
            status = str(e)
        v = vh.cpu().numpy().view(np.float64)
        t = th.cpu().numpy() / YEAR
        events, armed = [], True
        for vi, ti in zip(v, t):
            if armed and vi > 1e-3:
                events.append(float(ti))
                armed = False
            elif vi < 1e-4:
                armed = True
        yrs = float(ens.time[0]) / YEAR
        runs[name] = {"ms_per_step": 1e3 * wall / args.steps, "wall_s": wall, "sim_years": yrs,
                      "years_per_hour": yrs / (wall / 3600.0), "accepted_steps":
                          int(ens.naccepted[0]), "event_years": events,
                      "peak_slip_rate": float(v.max()), "build_s": build_s, "status": status}
        log(f"example {name}: {json.dumps(runs[name])}")
        del op, ens
        torch.cuda.empty_cache()
    head = runs["hmatrix_fp32"]
    return {
        "metric": "sim-years/wall-hour, reference BP5-QD example (N=9250, half space)",
        "value": head["years_per_hour"], "unit": "sim-years/wall-hour",
        "n_gpus": 1, "steps": args.steps, "warmup": 0, "ms_per_step": head["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 H-matrix storage, f64 state (dense f64 and H f64 alongside)",
        "data": "the reference's examples/BP5-QD model files (mesh, parameters), as archived",
        "config": {"workload": f"BP5-QD example, N={p.n}, half space, {args.steps} accepted "
                               f"steps from the model's initial state, device controller"},
        "speedup_vs_dense_fp64": head["years_per_hour"] / runs["dense_fp64"]["years_per_hour"],
        **runs,
    }


# ------------------------------------------------------------- assembly
def bench_assembly(args):
    import numpy as np
    import torch

    from pyquake3d_amd import tde
    from pyquake3d_amd.mesh import undulating_fault

    dev = torch.device("cuda:0")
    mesh = undulating_fault(length=100e3, depth=25e3, nx=200, nz=50, amplitude=300.0,
                            wavelength=5e3, seed=0)
    N = mesh.n
    hs = bool(args.halfspace)
    dm = tde.DeviceMesh.from_mesh(mesh, 32e9, 32e9, device=dev, halfspace=hs)
    out = torch.empty((6, N, N), dtype=torch.float64, device=dev)
    log(f"assembly: N={N}, output {out.numel() * 8 / 1e9:.1f} GB, half space {hs}")
    for _ in range(args.warmup):
        dm.traction_block(out=out)
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record()
    for _ in range(args.steps):
        dm.traction_block(out=out)
    ev1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = ev0.elapsed_time(ev1) / args.steps
    pairs = float(N) * N
    sample = out[:, :64, :64].cpu().numpy()
    assert np.isfinite(sample).all()
    if hs:
        res = {
            "metric": "TDE evaluations/s (N x N assembly, FP64, half space)",
            "value": 2.0 * pairs * args.steps / wall, "unit": "evaluations/s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": wall * 1e3 / args.steps,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic undulating fault mesh, seed 0",
            "config": {"workload": "kernel assembly, undulating fault 100x25 km, N=20000, unit "
                                   "Ss and Ds, six traction operators, FP64, half space "
                                   "(free surface z=0)"},
            "roofline": {"bound": "fp64-valu", "achieved": None, "peak": MI355X_FP64_TFLOPS,
                         "unit": "TFLOP/s", "frac": None, "traffic": None,
                         "kernel_ms": kern_ms},
        }
        return res
    achieved = pairs * FLOPS_PER_PAIR / (kern_ms * 1e-3) / 1e12
    res = {
        "metric": "TDE evaluations/s (N x N assembly, FP64)",
        "value": 2.0 * pairs * args.steps / wall,
        "unit": "evaluations/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic undulating fault mesh, seed 0",
        "config": {"workload": "kernel assembly, undulating fault 100x25 km, N=20000, unit Ss "
                               "and Ds, six traction operators, FP64"},
        "roofline": {"bound": "fp64-valu", "achieved": achieved, "peak": MI355X_FP64_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / MI355X_FP64_TFLOPS, "traffic": None,
                     "kernel_ms": kern_ms},
    }
    if not args.no_cpu_baseline:
        res["cpu_baseline"] = cpu_baseline_assembly(mesh, dm)
    return res


def bench_disp(args):
    import numpy as np
    import torch

    from pyquake3d_amd import tde
    from pyquake3d_amd.mesh import undulating_fault

    dev = torch.device("cuda:0")
    mesh = undulating_fault(length=100e3, depth=25e3, nx=200, nz=50, amplitude=300.0,
                            wavelength=5e3, seed=0)
    N = mesh.n
    dm = tde.DeviceMesh.from_mesh(mesh, 32e9, 32e9, device=dev, halfspace=True)
    gx, gy = np.meshgrid(np.linspace(-60e3, 60e3, 256), np.linspace(-20e3, 20e3, 256) + 37.0)
    obs_np = np.stack([gx.ravel(), gy.ravel(), np.zeros(gx.size)], 1)
    rng = np.random.default_rng(0)
    ss_np, ds_np = rng.normal(size=N), rng.normal(size=N)
    obs = torch.from_numpy(obs_np).to(dev)
    ss, ds = torch.from_numpy(ss_np).to(dev), torch.from_numpy(ds_np).to(dev)
    out = torch.empty((len(obs_np), 3), dtype=torch.float64, device=dev)
    log(f"disp: {len(obs_np)} surface points x {N} elements, half space")
    for _ in range(args.warmup):
        dm.displacement(obs, ss, ds, out=out)
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record()
    for _ in range(args.steps):
        dm.displacement(obs, ss, ds, out=out)
    ev1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = ev0.elapsed_time(ev1) / args.steps
    u = out.cpu().numpy()
    assert np.isfinite(u).all()
    pairs = float(len(obs_np)) * N
    achieved = pairs * DISP_FLOPS_PER_PAIR_HS / (kern_ms * 1e-3) / 1e12
    res = {
        "metric": "displacement pair evaluations/s (surface grid x elements, half space, FP64)",
        "value": pairs * args.steps / wall,
        "unit": "pairs/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic undulating fault mesh and seeded slip",
        "config": {"workload": "surface displacement, 256x256 grid at z=0 from the undulating "
                               "fault 100x25 km, N=20000, half space, FP64"},
        "roofline": {"bound": "fp64-valu", "achieved": achieved, "peak": MI355X_FP64_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / MI355X_FP64_TFLOPS, "traffic": None,
                     "kernel_ms": kern_ms},
    }
    if not args.no_cpu_baseline:
        res["cpu_baseline"] = cpu_baseline_disp(mesh, obs_np, ss_np, ds_np)
    return res


def cpu_baseline_disp(mesh, obs, ss, ds):
    """The displacement oracle (numpy, one thread per call, vectorised over
    points) on the first elements of the workload until ~10 s have passed."""
    from oracle.tde_disp_oracle import td_disp_hs
    tri = mesh.triangles()
    t = time.perf_counter()
    j = 0
    while time.perf_counter() - t < 10.0 and j < mesh.n:
        td_disp_hs(obs, tri[j, 0:3], tri[j, 3:6], tri[j, 6:9], ss[j], ds[j], 0.0, 0.25)
        j += 1
    dt = time.perf_counter() - t
    return {"value": len(obs) * j / dt, "unit": "pairs/s", "cores": 1, "kind": "port",
            "sample": f"numpy displacement oracle, {j} elements x {len(obs)} surface points, "
                      f"{dt:.1f} s"}


def cpu_baseline_assembly(mesh, dm):
    import numpy as np

    from oracle.tde_oracle import Oracle
    o = Oracle()
    tri = mesh.triangles()
    rows = np.arange(mesh.n)
    t = time.perf_counter()
    o.traction_block(dm.xg, dm.eleVec, tri, rows, np.arange(4), 32e9, 32e9)
    per_col = (time.perf_counter() - t) / 4
    ncol = int(max(8, min(mesh.n, 10.0 / max(per_col, 1e-6))))
    t = time.perf_counter()
    o.traction_block(dm.xg, dm.eleVec, tri, rows, np.arange(ncol), 32e9, 32e9)
    dt = time.perf_counter() - t
    return {"value": 2.0 * mesh.n * ncol / dt, "unit": "evaluations/s",
            "cores": o.num_threads(), "kind": "port",
            "sample": f"oracle C restatement (OpenMP), {ncol} source columns x {mesh.n} "
                      f"receivers x (Ss, Ds), {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--workload", default="hstep", choices=["hstep", "assembly", "matvec", "learned", "ragged",
                                                   "accuracy", "example", "disp"])
    ap.add_argument("--dtype", default="f32", choices=["f32", "f64"],
                    help="operator storage of the matvec workload")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-graph", action="store_true",
                    help="hstep: time the headline ensemble with eager launches only")
    ap.add_argument("--accumulate", default="f64", choices=["f64", "mfma"],
                    help="16 right-hand sides over FP32 storage: 'mfma' runs the dense and U "
                         "products on FP32 matrix cores (hstep headline and matvec)")
    ap.add_argument("--halfspace", action="store_true",
                    help="assembly: half-space operators (free surface z=0)")
    ap.add_argument("--nb", type=int, default=None, choices=[1, 2, 4, 8, 16],
                    help="friction draws in the batched variant (default 16; learned: 8)")
    ap.add_argument("--count", type=int, default=64, help="ragged: number of simulations")
    ap.add_argument("--events", type=int, default=0,
                    help="accuracy: stop after this many events (0: run to --tmax)")
    ap.add_argument("--vup", type=float, default=1e-8, help="accuracy: event onset max V (m/s)")
    ap.add_argument("--vdown", type=float, default=3e-9, help="accuracy: re-arm max V (m/s)")
    ap.add_argument("--tmax", type=float, default=150.0, help="accuracy: simulated years cap")
    ap.add_argument("--dc-scale", type=float, default=1.0, help="accuracy: factor on Dc")
    ap.add_argument("--record", type=int, default=800,
                    help="learned: steps of the reference controller logged for training")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (the "
                                   "multi-GPU design is documented, not built)"}))
        sys.exit(2)
    if args.workload != "learned":
        args.nb = 16 if args.nb is None else args.nb
    if args.workload == "hstep":
        args.steps = 3000 if args.steps is None else args.steps
        args.warmup = 200 if args.warmup is None else args.warmup
        res = bench_hstep(args)
    elif args.workload == "example":
        args.steps = 20000 if args.steps is None else args.steps
        res = bench_example(args)
    elif args.workload == "accuracy":
        args.steps = 60000 if args.steps is None else args.steps
        res = bench_accuracy(args)
    elif args.workload == "ragged":
        args.steps = 1000 if args.steps is None else args.steps
        args.warmup = 100 if args.warmup is None else args.warmup
        res = bench_ragged(args)
    elif args.workload == "learned":
        args.steps = 1500 if args.steps is None else args.steps
        args.nb = 8 if args.nb is None else args.nb
        res = bench_learned(args)
    elif args.workload == "disp":
        args.steps = 3 if args.steps is None else args.steps
        args.warmup = 1 if args.warmup is None else args.warmup
        res = bench_disp(args)
    elif args.workload == "matvec":
        args.steps = 20 if args.steps is None else args.steps
        args.warmup = 3 if args.warmup is None else args.warmup
        res = bench_matvec(args)
    else:
        args.steps = 3 if args.steps is None else args.steps
        args.warmup = 1 if args.warmup is None else args.warmup
        res = bench_assembly(args)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
