"""Benchmark of the PyQuake3D hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

Workloads
  hstep     (default) H-matrix time stepping at N = 100,352 triangles on the
            undulating 100 x 25 km fault (leaf 64, eta 2, ACA eps 1e-4),
            BP5-like heterogeneous friction with Dc x 0.7 and the normal
            traction fixed. Every draw is integrated from its seeded initial
            state over [0, 380] years (>= 3 whole seismic cycles) whatever
            --steps is; the metric is simulated years per wall-hour over that
            interval:
              plain_fp64   : FP64 operator storage, one simulation, the
                             reference controller (the baseline)
              fp32_batched : FP32 operator storage, --nb friction draws
                             (default 16) sharing one multi-RHS H-matvec per
                             step, learned first-attempt step size; rate =
                             sum over the draws (``value``)
              fp64_nb16    : every draw in FP64 with the reference
                             controller: the event-onset reference of each
                             draw (reported per draw)
  example   The reference's BP5-QD example model (N = 9,250, half space)
            from its initial state: dense FP64 vs FP64 / FP32 H-matrices,
            event times and simulated years per wall-hour.
  accuracy  Event times of the hstep problem with FP32 and with FP64
            operator storage (one draw, device controller): the relative
            event-time error behind "matched accuracy".
  ragged    --count simulations (default 64) on meshes of uneven size as
            one block-diagonal FP32 H-matrix with the device controller;
            per-instance accuracy of every 8th against its FP64 solo run.
  learned   The step-size network at its configured scale: ~1e6 training
            rows from the reference controller on the N = 20,000 mesh,
            training on the GPU, then ten seismic cycles at N = 100,352 with
            the reference controller and with the learned proposal:
            rejected-step fraction and wall time at equal rtol.
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

# HBM bytes per H-matvec from two rocprofv3 --pmc passes (FETCH_SIZE, then
# WRITE_SIZE) of the matvec workload, summed over the product's kernels,
# FETCH scaled by 2.0 and WRITE by 1.0: the factors the same counters read
# for a 2 GiB device-to-device copy (2 GiB read as 1,048,608 KiB fetched,
# 2 GiB written as 2,097,152 KiB). Key: (storage, right-hand sides, row
# planes, accumulation); N = 100,352, the tuned configuration 4,2,cost,2,4.
MEASURED_TRAFFIC = {("f32", 16, 2, "mfma"): 5_169_650_432}

# FP64 operations of one (receiver, source) pair in tde_pair_traction_sd
# (unit strike and dip slip, both projected), counted operator by operator
# from pyquake3d_amd/csrc/tde_device.h: per angular dislocation 191 (brackets)
# + 8 (side frame) + 2 x (55 prefactors/combination + 34 side rotation) + 4
# (slip rotation), three of them, plus 18 (TD transform) + 28 (configuration)
# + 2 x (129 global rotation + 20 Hooke + 30 projection) + 24 (summing the
# three angular dislocations); divisions and square roots count as one.
FLOPS_PER_PAIR = 1571.0
# What the GPU executes per pair of the N = 20,000 assembly, from the
# SQ_INSTS_VALU and SQ_INSTS_VALU_{FMA,MUL,ADD,TRANS}_F64 counters of one
# launch (x 64 lanes / 4e8 pairs; an FMA counts 2): full space 3,309 VALU
# instructions and 3,830 FP64 operations per pair (the correctly rounded
# divisions and square roots expand into ~10 instructions each), half space
# 23,919 and 24,526.
VALU_INSTS_PER_PAIR = {False: 3309.0, True: 23919.0}
FP64_OPS_PER_PAIR = {False: 3830.0, True: 24526.0}
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


# The config-3 workload of the headline: BP5-like friction draws on the
# undulating fault with Dc scaled by 0.7 and the normal traction fixed
# (the reference's Fix_Tn switch, src/QDsim.py:906-909). With the normal
# traction coupled, every rupture on this geometry (300 m undulation over
# 5 km) drives the normal traction through zero within the first event; the
# rate-and-state law then overflows (V ~ sinh(tau / (a sigma))) and the
# reference's abort rule ends the run. With it fixed, every draw completes
# seismic cycles of ~110 years (peak slip rates 1-2 m/s) under the
# reference controller.
C3_DC_SCALE = 0.7
C3_T_END_YEARS = 380.0       # every draw has had >= 4 events (3 whole cycles) by then
EVENT_VUP, EVENT_VDOWN = 1e-2, 1e-4   # seismic onset and re-arm thresholds (m/s)
EVENT_MERGE_YEARS = 1.0      # onsets closer than this belong to one rupture sequence
STEP_MLP = REPO / "pyquake3d_amd" / "data" / "step_mlp_c2.npz"
# first attempt = dtnext 10^min(prediction - margin, 0): a negative margin
# leans the proposal towards the reference controller's dtnext. On the
# device a rejected attempt costs one cell-local attempt round while every
# accepted step costs an operator product, so leaning towards larger steps
# pays: over [0, 380] yr of the 16 draws (plain FP32 matrix-core sums)
# margin 0 took 57,300 iterations, -0.03 56,100, -0.06 54,500; the reference
# controller 57,300.
LEARNED_MARGIN = -0.06
# attempt rounds per iteration of the device controller in the cycle runs: a
# draw that needs more attempts continues in the next iteration (same
# trajectory), an empty round still costs two launches. Over [0, 120] yr of
# the 16 draws (learned proposal, matrix-core product): 2 rounds 24,800
# iterations / 31.4 s, 3: 24,000 / 30.8 s, 4: 23,800 / 30.8 s, 6: 23,600 /
# 31.0 s, 8: 23,600 / 31.2 s.
CYCLE_ROUNDS = 4
# the one-draw FP64 baseline rejects less often per iteration and measured
# faster with 8 rounds (41,000 iterations / 65.6 s over [0, 380] yr against
# 43,200 / 68.6 s with 4), so it keeps 8
PLAIN_ROUNDS = 8
RAGGED_GRAPH_STEPS = 10   # ragged: iterations per captured HIP graph


def c3_draws(nb, nx=448, nz=112, dc_scale=C3_DC_SCALE):
    """NB friction draws (seeds 0..NB-1) on one undulating 100 x 25 km mesh
    (448 x 112 rectangles -> N = 100,352; 200 x 50 -> the N = 20,000 mesh
    of the assembly workload), Dc scaled by ``dc_scale``, normal traction
    fixed."""
    from dataclasses import replace

    from pyquake3d_amd.problem import undulating_c3
    p0 = undulating_c3(nx=nx, nz=nz)
    draws = [p0] + [undulating_c3(nx=nx, nz=nz, friction_seed=k, mesh=p0.mesh)
                    for k in range(1, nb)]
    return [replace(p, dc=p.dc * dc_scale, fix_Tn=True) for p in draws]


def merge_events(ev_s):
    """Event onsets (s) -> onsets in years with sequences closer than
    EVENT_MERGE_YEARS merged into their first onset."""
    out = []
    for t in sorted(float(x) / YEAR for x in ev_s):
        if not out or t - out[-1] > EVENT_MERGE_YEARS:
            out.append(t)
    return out


def run_cycles(problems, H, name, t_end_years, warmup, rounds=CYCLE_ROUNDS, mlp=None,
               timing_every=0,
               check_every=100, max_wall=900.0, margin=0.0, hi=0.0, dt0=1e-3):
    """``problems`` (friction draws on the operator's mesh) stepped by the
    device controller from their initial state until every draw reaches
    ``t_end_years`` (a draw that gets there first stops taking steps; the
    shared product still runs for all). ``warmup`` iterations are run first
    and the state restored, so the timed region starts from t = 0. Events
    are detected on the device. Returns the result dict and the ensemble."""
    import numpy as np
    import torch

    from pyquake3d_amd.qdsim import DeviceEnsemble, StepAborted

    ens = DeviceEnsemble(problems, H, device=H.device, rounds=rounds, t_end=t_end_years * YEAR,
                         vup=EVENT_VUP, vdown=EVENT_VDOWN, mlp=mlp, margin=margin, hi=hi,
                         dt0=dt0)
    s0 = ens.save_state()
    for _ in range(warmup):
        ens.step()
    torch.cuda.synchronize()
    ens.load_state(s0)
    ens.iterations = 0
    if timing_every:
        ens.timing = True
        ens.timing_every = timing_every
    ens.matvec_events = []
    it = 0
    w0 = time.perf_counter()
    last = w0
    while True:
        ens.step()
        it += 1
        if it % check_every == 0:
            fl = ens.flag.cpu().numpy()
            now = time.perf_counter()
            if np.all(fl >= 2) or now - w0 > max_wall:
                break
            if now - last > 30.0:
                last = now
                tt = ens.tdev.cpu().numpy() / YEAR
                log(f"{name}: {it} iterations, t = {tt.min():.1f}..{tt.max():.1f} yr, "
                    f"{now - w0:.0f} s")
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    try:
        ens.sync()
        status = "ok"
    except StepAborted as e:
        status = f"abort: {e}"
    mv = ens.matvec_ms() if timing_every else []
    nb = len(problems)
    yrs = ens.time / YEAR
    ev = [merge_events(e) for e in ens.events()]
    rej = ens.rejected()
    acc = ens.naccepted
    planes = 2 if problems[0].fix_Tn else 3
    res = dict(name=name, dtype=str(H.dtype).replace("torch.", ""), nb=nb, rounds=rounds,
               learned=mlp is not None, iterations=it, wall_s=wall,
               ms_per_iteration=1e3 * wall / it, sim_years=float(yrs.sum()),
               years_per_hour=float(yrs.sum()) / (wall / 3600.0),
               t_end_years=[float(x) for x in yrs], finished=ens.finished().tolist(),
               aborted=ens.aborted().tolist(), status=status,
               accepted_steps=[int(x) for x in acc], rejected_attempts=[int(x) for x in rej],
               rejected_fraction=float(rej.sum() / max(rej.sum() + acc.sum(), 1)),
               events_years=ev, events_per_draw=[len(e) for e in ev],
               peak_slip_rate=[float(x) for x in ens.peak_slipv()],
               matvec_ms=float(np.mean(mv)) if mv else None, matvec_samples=len(mv),
               matvec_bytes=H.algorithmic_bytes(ens.nrhs, planes))
    log(f"{name}: {json.dumps({k: v for k, v in res.items() if k not in ('events_years', 't_end_years', 'accepted_steps', 'rejected_attempts', 'peak_slip_rate', 'finished', 'aborted')})}")
    return res, ens


def event_errors(fast, ref):
    """Per draw: paired event onsets (k-th with k-th) of a run against the
    FP64 reference run, relative errors of the onset times and of the
    recurrence intervals, peak slip rates."""
    import numpy as np
    out = []
    for b, (ef, er) in enumerate(zip(fast["events_years"], ref["events_years"])):
        k = min(len(ef), len(er))
        rel = [abs(ef[i] - er[i]) / er[i] for i in range(k)]
        rf, rr = np.diff(ef[:k]), np.diff(er[:k])
        out.append({"draw": b, "events": len(ef), "events_fp64": len(er),
                    "onsets_yr": [round(x, 4) for x in ef[:k]],
                    "onsets_fp64_yr": [round(x, 4) for x in er[:k]],
                    "onset_rel_err": [float(x) for x in rel],
                    "max_onset_rel_err_after_first": float(max(rel[1:])) if k > 1 else None,
                    "max_recurrence_rel_err": (float(np.max(np.abs(rf - rr) / rr))
                                               if k > 1 else None),
                    "peak_log10_slip_rate": float(np.log10(fast["peak_slip_rate"][b])),
                    "peak_log10_slip_rate_fp64": float(np.log10(ref["peak_slip_rate"][b]))})
    return out


# host cores one GPU's job may use on the box (its worker-pool share)
CPU_BASELINE_PROCS = (8, 16)


def cpu_baseline_hstep(problem, H, gpu_res, nsteps=3, nprocs=CPU_BASELINE_PROCS):
    """Reference-style CPU step on the same blocks, timed over ``nsteps``
    steps from the initial state: the per-block numpy FP64 matvec of the
    four A operators (as the reference applies them with a fixed normal
    traction) and the numpy RK45 step, once on one process (one BLAS
    thread) and once for each count in ``nprocs`` with the blocks scattered
    over that many processes as the reference's MPI ranks hold them
    (partial products summed); the rates are expressed for the plain GPU
    run's simulated interval and accepted steps. ``value`` is the rate of
    the largest process count."""
    import numpy as np
    from threadpoolctl import threadpool_limits

    from oracle.hmat_cpu import CpuHMatrix, ScatterMatvec
    from oracle.qdsim_oracle import OracleSim
    t0 = time.perf_counter()
    C = CpuHMatrix(H)
    setup = time.perf_counter() - t0
    perm = H.perm
    init = [np.asarray(x)[perm] for x in problem.initial_state()]
    g = lambda x: np.asarray(x)[perm]
    planes = 2 if problem.fix_Tn else 3
    steps = gpu_res["accepted_steps"][0]

    def timed(matvec):
        ora = OracleSim(None, g(problem.a), g(problem.b), g(problem.dc), g(problem.f0),
                        problem.V0, problem.mu, problem.Cs, g(problem.Vpl), g(problem.rake), init,
                        normal_loading=g(problem.normal_loading), fix_Tn=problem.fix_Tn,
                        matvec=matvec)
        dt = 1e-3
        t0 = time.perf_counter()
        for _ in range(nsteps):
            _, dt = ora.simu_forward(dt)
        return (time.perf_counter() - t0) / nsteps

    per = {}
    with threadpool_limits(limits=1):
        one = timed(lambda v1, v2: C.matvec(v1, v2, planes=planes))
        for q in nprocs:
            S = ScatterMatvec(C, q)
            try:
                per[q] = timed(lambda v1, v2: S.matvec(v1, v2, planes=planes))
            finally:
                S.close()
    nproc = max(nprocs)
    many = per[nproc]
    rate = lambda per_step: gpu_res["sim_years"] / (per_step * steps / 3600.0)
    blocks = f"{len(C.dense)} dense + {len(C.lowrank)} low-rank blocks"
    return {"value": rate(many), "unit": "sim-years/wall-hour", "cores": nproc, "kind": "port",
            "sample": f"{nsteps} steps of the reference-style CPU step (per-block numpy FP64 "
                      f"matvec of the {2 * planes} operators over the same {blocks}, scattered "
                      f"over {nproc} processes with one BLAS thread each and reduced, numpy "
                      f"RK45) at {many:.3f} s/step, applied to the plain GPU run's {steps} "
                      f"accepted steps / {gpu_res['sim_years']:.1f} simulated years; host "
                      f"setup {setup:.1f} s",
            "single_process": {"value": rate(one), "cores": 1, "s_per_step": one},
            "by_processes": {str(q): {"value": rate(t), "s_per_step": t} for q, t in per.items()},
            "s_per_step": many}


def tune_product(H, args, planes):
    """The autotuner's choice for the matrix-core product (HMatrix.tune), or
    the configuration given with --tuning; None for other products."""
    if H.accumulate != "mfma" or H.esize != 4:
        return None
    if args.tuning:
        d, w, o, p1d, p1w = args.tuning.split(",")
        H.tuning = {"depth": int(d), "waves": int(w), "order": o, "p1_depth": int(p1d),
                    "p1_waves": int(p1w)}
        return {"choice": H.tuning, "planes": planes, "candidates": []}
    return H.tune(planes=planes)


def load_step_mlp(dev):
    from pyquake3d_amd.learned_dt import StepMLP
    if not STEP_MLP.exists():
        raise RuntimeError(f"{STEP_MLP} is missing: train it with `python bench.py --workload "
                           f"learned` (it writes the network there)")
    return StepMLP.load(STEP_MLP, device=dev)


def bench_hstep(args):
    """The headline: sim-years per wall-hour over whole seismic cycles at
    N = 100,352, plain FP64 (one simulation, FP64 operator storage, the
    reference controller) against the metric's variant (FP32 storage,
    --nb draws batched into one multi-RHS product per step, learned
    step-size proposal), each integrating [0, C3_T_END_YEARS] of every
    draw whatever --steps is; event onsets of every draw of the variant are
    compared with the same draw's FP64 run."""
    import torch

    draws = c3_draws(args.nb)
    problem = draws[0]
    t_end = args.t_end_years
    log(f"hstep: N={problem.n}, {args.nb} draws, [0, {t_end} yr]")
    dev = torch.device("cuda:0")
    mlp = None if args.no_learned else load_step_mlp(dev)
    H64, b64 = build_hmatrix(problem, torch.float64)
    plain, _ = run_cycles([problem], H64, "plain_fp64", t_end, args.warmup, timing_every=20,
                          rounds=PLAIN_ROUNDS)
    plain["build_s"] = b64
    ref64 = floor = None
    if not args.no_accuracy:
        # FP64 reference of every draw: the batched FP64 product is bitwise the
        # single-draw product, so draw b here is draw b's FP64 solo run
        ref64, _ = run_cycles(draws, H64, f"fp64_nb{args.nb}", t_end, args.warmup)
        if args.noise_floor:
            # the same FP64 runs started from dt = 2e-3 s instead of 1e-3 s:
            # how far the reference method's own event onsets move when only
            # its step sequence changes
            alt, _ = run_cycles(draws, H64, f"fp64_nb{args.nb}_dt0x2", t_end, args.warmup,
                                dt0=2e-3)
            floor = event_errors(alt, ref64)
    del H64
    torch.cuda.empty_cache()
    H32, b32 = build_hmatrix(problem, torch.float32)
    H32.accumulate = args.accumulate
    tuning = tune_product(H32, args, 2)
    if tuning and tuning["candidates"]:
        log(f"hstep: autotuner chose {tuning['choice']} "
            f"({min(ms for _, ms in tuning['candidates']):.4f} ms)")
    fast, _ = run_cycles(draws, H32, f"fp32_nb{args.nb}" + ("_learned" if mlp else ""), t_end,
                         args.warmup, mlp=mlp, timing_every=20, margin=LEARNED_MARGIN)
    fast["accumulate"] = args.accumulate
    fast["tuning"] = tuning
    fast["learned_margin"] = LEARNED_MARGIN if mlp else None
    fast["build_s"] = b32
    achieved = fast["matvec_bytes"] / (fast["matvec_ms"] * 1e-3) / 1e9
    acc = event_errors(fast, ref64) if ref64 else None
    worst = None
    if acc:
        vals = [a["max_onset_rel_err_after_first"] for a in acc
                if a["max_onset_rel_err_after_first"] is not None]
        worst = max(vals) if vals else None
    res = {
        "metric": "sim-years/wall-hour at N=1e5 (fp32+batched+learned dt vs plain fp64)",
        "value": fast["years_per_hour"],
        "unit": "sim-years/wall-hour",
        "n_gpus": 1,
        "steps": fast["iterations"],
        "warmup": args.warmup,
        "ms_per_step": fast["ms_per_iteration"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("f32 operator storage and f32 matrix-core block products (compensated f32 "
                  "sums across items, f64 across item groups), f64 state and RK45; f32 "
                  "step-size network" if args.accumulate == "mfma_comp" else
                  "f32 operator storage and f32 matrix-core block products, f64 state and RK45; "
                  "f32 step-size network" if args.accumulate == "mfma" else
                  "f32 operator storage, f64 state and accumulation; f32 step-size network"),
        "data": "synthetic undulating fault mesh and seeded friction fields",
        "config": {"workload": f"H-matrix RK45 time stepping over whole seismic cycles, "
                               f"undulating fault 100x25 km, N={problem.n}, leaf 64, eta 2, ACA "
                               f"eps 1e-4, BP5-like friction draws with Dc x {C3_DC_SCALE}, "
                               f"normal traction fixed (reference Fix_Tn), every draw "
                               f"integrated over [0, {t_end:g}] yr (>= 3 whole cycles) "
                               f"whatever --steps is; a step: one iteration of the device "
                               f"controller (one product, up to {CYCLE_ROUNDS} attempt rounds, commit)",
                   "variant": f"fp32+batched+learned: {args.nb} friction draws sharing the "
                              f"operator, one {args.nb}-RHS H-matvec per step, learned "
                              f"first-attempt step size (RK45 estimator accepts or rejects, PI "
                              f"fallback)" if mlp else
                              f"fp32+batched: {args.nb} draws, reference controller",
                   "baseline": "plain_fp64: one simulation (draw 0), FP64 operator storage, "
                               "reference controller"},
        "speedup_vs_plain_fp64": fast["years_per_hour"] / plain["years_per_hour"],
        "event_onset_max_rel_err_after_first": worst,
        "plain_fp64": plain,
        "fp32_batched": fast,
        "fp64_batched_reference": ref64,
        "event_accuracy_vs_fp64": acc,
        "fp64_noise_floor_dt0x2": floor,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": MI355X_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / MI355X_HBM_GBS,
                     "traffic": MEASURED_TRAFFIC.get(("f32", args.nb, 2, args.accumulate)),
                     "kernel": (f"H-matvec, {args.nb} right-hand sides, two row planes "
                                f"(k_inputs_f32 + k_hmv_phase1_mfp + k_hmv_phase1b + "
                                f"k_hmv_phase2_mfs)" if args.accumulate == "mfma" else
                                f"H-matvec, {args.nb} right-hand sides, two row planes "
                                f"(k_inputs_f32 + k_hmv_phase1_mf16 + k_hmv_phase1b + "
                                f"k_hmv_phase2_mf16)" if args.accumulate.startswith("mfma")
                                else f"H-matvec, {args.nb} right-hand sides, two row planes "
                                     f"(k_hmv_phase1 + k_hmv_phase1b + k_hmv_phase2)"),
                     "kernel_ms": fast["matvec_ms"], "algorithmic_bytes": fast["matvec_bytes"]},
    }
    if not args.no_cpu_baseline:
        res["cpu_baseline"] = cpu_baseline_hstep(problem, H32, plain)
    return res


def per_block_errors(H, mesh, nsample=1000, seed=0, max_entries=4_000_000):
    """Relative Frobenius error of the stored blocks of H against the same
    blocks assembled in FP64 by the TDE kernel: up to ``nsample`` dense
    (near-field) and ``nsample`` low-rank blocks drawn with ``seed``, blocks
    with m n <= ``max_entries``. A dense block's error is the rounding of its
    storage type; a low-rank block's adds the ACA truncation (eps) and the
    rounding of U and V. Returns quantiles per kind."""
    import numpy as np
    import torch

    from pyquake3d_amd.tde import DeviceMesh
    dm = DeviceMesh.from_mesh(mesh, H.mu, H.lambda_, device=H.device)
    B = H.blocks
    rng = np.random.default_rng(seed)
    m_all, n_all = np.asarray(B["m"]), np.asarray(B["n"])
    perm = H.perm

    def exact(bi):
        r = perm[B["r0"][bi]:B["r0"][bi] + B["m"][bi]]
        c = perm[B["c0"][bi]:B["c0"][bi] + B["n"][bi]]
        T = dm.traction_block(r, c).cpu().numpy()
        return T

    out = {}
    di = np.asarray(B["dense_idx"])
    ok = np.where(m_all[di] * n_all[di] <= max_entries)[0]
    pick = rng.choice(ok, size=min(nsample, len(ok)), replace=False)
    errs = []
    for q in pick:
        bi, base = int(di[q]), int(B["d_base"][q])
        m, n = int(m_all[bi]), int(n_all[bi])
        T = exact(bi)
        St = H.pool[base:base + 6 * m * n].double().cpu().numpy().reshape(6, n, m).transpose(0, 2, 1)
        errs.append(np.linalg.norm(St - T) / np.linalg.norm(T))
    out["dense"] = errs
    li = np.asarray(B["lr_idx"])
    ok = np.where(m_all[li] * n_all[li] <= max_entries)[0]
    pick = rng.choice(ok, size=min(nsample, len(ok)), replace=False)
    errs = []
    for q in pick:
        bi = int(li[q])
        T = exact(bi)
        S = np.block([[T[0], T[3]], [T[1], T[4]], [T[2], T[5]]])
        U, V = H.lowrank_factors(int(q))
        errs.append(np.linalg.norm(U @ V - S) / np.linalg.norm(S))
    out["lowrank"] = errs
    res = {}
    for kind, e in out.items():
        e = np.asarray(e)
        res[kind] = {"blocks": int(len(e)), "max": float(e.max()) if len(e) else None,
                     "p50": float(np.median(e)) if len(e) else None,
                     "p99": float(np.quantile(e, 0.99)) if len(e) else None}
    res["storage"] = str(H.dtype).replace("torch.", "")
    res["note"] = ("relative Frobenius error per block against the FP64 TDE assembly of the "
                   "same block; sampled blocks")
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
    tuning = tune_product(H, args, args.planes) if args.nb == 16 else None
    n, nb, dev = H.n, args.nb, H.device
    gen = torch.Generator(device=dev)
    gen.manual_seed(0)
    v1 = torch.randn(n * nb, dtype=torch.float64, device=dev, generator=gen)
    v2 = torch.randn(n * nb, dtype=torch.float64, device=dev, generator=gen)
    out = torch.empty((3, n * nb), dtype=torch.float64, device=dev)
    planes = args.planes
    for _ in range(args.warmup):
        H.matvec(v1, v2, out=out, nrhs=nb, planes=planes)
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(args.steps)]
    w0 = time.perf_counter()
    for e0, e1 in evs:
        e0.record()
        H.matvec(v1, v2, out=out, nrhs=nb, planes=planes)
        e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    ms = sum(e0.elapsed_time(e1) for e0, e1 in evs) / len(evs)
    mf_err = None
    if H.accumulate.startswith("mfma") and nb == 16 and args.dtype == "f32":
        # the matrix-core product against the FP64-accumulated product of the
        # same FP32 storage: max |difference| / max |FP64-accumulated value|
        # per output row plane and right-hand side, worst case reported
        ref = torch.empty_like(out)
        H.accumulate = "f64"
        H.matvec(v1, v2, out=ref, nrhs=nb, planes=planes)
        H.accumulate = args.accumulate
        H.matvec(v1, v2, out=out, nrhs=nb, planes=planes)
        d = (out[:planes] - ref[:planes]).abs().view(planes, n, nb).amax(dim=1)
        r = ref[:planes].abs().view(planes, n, nb).amax(dim=1)
        mf_err = float((d / r).max())
    B = H.blocks
    lr, k = B["lr_idx"], B["ranks"]
    v_bytes = int((2 * B["n"][lr] * k).sum()) * H.esize
    u_bytes = int((planes * B["m"][lr] * k).sum()) * H.esize
    d_bytes = int(H.stats.dense_entries) * H.esize * planes // 3
    vec_bytes = (2 + planes) * 8 * n * nb
    algo = H.algorithmic_bytes(nb, planes)
    achieved = algo / (ms * 1e-3) / 1e9
    return {
        "metric": "H-matvec effective HBM bandwidth",
        "value": achieved, "unit": "GB/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": (f"{args.dtype} operator storage, f32 matrix-core products for the dense and U "
                  f"blocks, f64 V products and vectors"
                  if (H.accumulate.startswith("mfma") and nb == 16 and args.dtype == "f32") else
                  f"{args.dtype} operator storage, f64 vectors and accumulation"),
        "data": "synthetic undulating fault mesh, seeded random slip-rate vectors",
        "config": {"workload": f"H-matvec, undulating fault 100x25 km, N={n}, leaf 64, eta 2, "
                               f"ACA eps 1e-4, {nb} right-hand sides, {planes} row planes"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": MI355X_HBM_GBS,
                     "unit": "GB/s", "frac": achieved / MI355X_HBM_GBS,
                     "traffic": MEASURED_TRAFFIC.get((args.dtype, nb, planes, args.accumulate)),
                     "kernel_ms": ms, "algorithmic_bytes": algo},
        "bytes": {"phase1_lowrank_V": v_bytes, "phase2_lowrank_U": u_bytes,
                  "phase2_dense": d_bytes, "vectors": vec_bytes, "total": algo},
        "copy_2GiB_GBs": copy_gbs,
        "build_s": build_s,
        "mfma_rel_err_vs_f64_accumulation": mf_err,
        "tuning": tuning,
        "per_block_error": (per_block_errors(H, problem.mesh, nsample=args.block_sample)
                            if args.block_sample > 0 else None),
    }


# ------------------------------------------------------------- learned
def bench_learned(args):
    """The learned step-size proposal at its configured scale.

    1. Training rows from the reference controller on the device: --nb
       friction draws of the config-3 friction on the N = 20,000 mesh of the
       same undulating geometry (200 x 50 rectangles), FP32 H-matrix,
       --record iterations; every accepted step logs (six features of its
       start, log10 h_accepted / dtnext) -> about --nb x --record rows
       through interseismic, nucleation and rupture phases.
    2. Training on the GPU with the library's forward/backward/Adam kernels;
       the network is written to --mlp-out.
    3. Ten seismic cycles at N = 100,352 (the hstep draws, [0, --cycle-years]
       yr), plain reference controller against the learned proposal (RK45
       estimator accepts or rejects every attempt, PI fallback), both on the
       device controller at the same rtol (1e-4) and attempt rounds, and the
       reference again at the single run's rounds (the ratio is taken
       against its faster run): rejected-step fraction, wall time, accepted
       steps, event onsets of every draw."""
    import numpy as np
    import torch

    from pyquake3d_amd.learned_dt import StepMLP
    from pyquake3d_amd.qdsim import DeviceEnsemble

    dev = torch.device("cuda:0")
    small = c3_draws(args.nb, nx=200, nz=50)
    Hs, _ = build_hmatrix(small[0], torch.float32)
    rec = DeviceEnsemble(small, Hs, device=dev, rounds=8, record=args.nb * args.record)
    w0 = time.perf_counter()
    last = w0
    for i in range(args.record):
        rec.step()
        if i % 2000 == 1999 and time.perf_counter() - last > 30:
            last = time.perf_counter()
            rec.sync()
            log(f"learned: recording {i + 1}/{args.record}, t = "
                f"{rec.time.min() / YEAR:.1f}..{rec.time.max() / YEAR:.1f} yr")
    rec.sync()
    rec_s = time.perf_counter() - w0
    X, T = rec.samples()
    rec_years = float(rec.time.mean() / YEAR)
    rec_events = [len(merge_events(e)) for e in rec.events()]
    del rec, Hs
    torch.cuda.empty_cache()
    log(f"learned: {len(T)} rows from {args.nb} draws over {rec_years:.0f} yr "
        f"({rec_s:.0f} s), events per draw {rec_events}")
    mlp = StepMLP(device=dev, seed=0)
    t0 = time.perf_counter()
    hist = mlp.fit(X, T, epochs=args.epochs, batch=4096, lr=2e-3)
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    pred = mlp.predict_log_dt(X[:200000])
    fit = {"rows": int(len(T)), "epochs": args.epochs, "loss_first": float(hist[0]),
           "loss_last": float(hist[-1]), "train_s": train_s,
           "target_mean": float(T.mean()), "target_sd": float(T.std()),
           "rms_error_log10": float(np.sqrt(np.mean((pred - T[:200000]) ** 2))),
           "recorded_years_per_draw": rec_years, "recorded_events_per_draw": rec_events}
    out = Path(args.mlp_out)
    out.parent.mkdir(parents=True, exist_ok=True)
    mlp.save(out, rows=len(T), epochs=args.epochs, source=f"{args.nb} draws, N=20000 "
             f"undulating mesh, Dc x {C3_DC_SCALE}, fixed normal traction, {args.record} "
             f"iterations of the reference controller")
    log(f"learned: trained ({json.dumps(fit)}), written to {out}")
    draws = c3_draws(args.nb)
    H, _ = build_hmatrix(draws[0], torch.float32)
    H.accumulate = args.accumulate
    tune_product(H, args, 2)
    # the reference controller at the ensemble's attempt rounds and at the
    # single run's (its trajectory is the same, only the iterations differ);
    # the ratio is taken against the faster of the two
    plain, _ = run_cycles(draws, H, "plain_controller", args.cycle_years, args.warmup,
                          max_wall=args.max_wall)
    plain8, _ = run_cycles(draws, H, f"plain_controller_r{PLAIN_ROUNDS}", args.cycle_years,
                           args.warmup, max_wall=args.max_wall, rounds=PLAIN_ROUNDS)
    learned, _ = run_cycles(draws, H, "learned_controller", args.cycle_years, args.warmup,
                            mlp=mlp, max_wall=args.max_wall, margin=LEARNED_MARGIN)
    acc = event_errors(learned, plain)
    best = min((plain, plain8), key=lambda r: r["wall_s"])
    return {
        "metric": "wall-time ratio over ten seismic cycles, reference / learned step-size "
                  "controller at equal rtol (N=1e5, rejected-step fractions alongside)",
        "value": best["wall_s"] / learned["wall_s"], "unit": "x",
        "value_same_rounds": plain["wall_s"] / learned["wall_s"],
        "n_gpus": 1, "steps": learned["iterations"], "warmup": args.warmup,
        "ms_per_step": learned["ms_per_iteration"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": f"f32 operator storage ({args.accumulate} products), f64 state; f32 network",
        "data": "synthetic undulating fault meshes and seeded friction fields",
        "config": {"workload": f"learned step size: trained on {len(T)} rows of the reference "
                               f"controller ({args.nb} draws, N=20000, {args.record} "
                               f"iterations), evaluated over [0, {args.cycle_years:g}] yr of "
                               f"{args.nb} draws at N={draws[0].n}, rtol 1e-4, FP32 H-matrix, "
                               f"device controller, {CYCLE_ROUNDS} attempt rounds per iteration "
                               f"(the reference also at {PLAIN_ROUNDS}; value against its faster "
                               f"run), proposal margin {LEARNED_MARGIN}"},
        "rejected_fraction": {"reference": plain["rejected_fraction"],
                              "learned": learned["rejected_fraction"]},
        "accepted_steps": {"reference": int(np.sum(plain["accepted_steps"])),
                           "learned": int(np.sum(learned["accepted_steps"]))},
        "sim_years_per_hour": {"reference": plain["years_per_hour"],
                               "learned": learned["years_per_hour"]},
        "training": fit, "reference": plain, "reference_rounds8": plain8, "learned": learned,
        "event_onsets_learned_vs_reference": acc,
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
    s0 = ens.save_state()
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
    # the same steps again as captured HIP graphs of RAGGED_GRAPH_STEPS
    # iterations each (one host launch per graph), from the same state: they
    # must end in the eager run's state bit for bit
    end = {k: getattr(ens, k).clone() for k in ("tdev", "slipv", "state", "nacc")}
    ens.timing = False
    ens.load_state(s0)
    ens.capture(steps=RAGGED_GRAPH_STEPS)
    torch.cuda.synchronize()
    w1 = time.perf_counter()
    for _ in range(args.steps // RAGGED_GRAPH_STEPS):
        ens.step()
    for _ in range(args.steps % RAGGED_GRAPH_STEPS):
        ens.release_graph()
        ens.step()
    torch.cuda.synchronize()
    wall_graph = time.perf_counter() - w1
    ens.release_graph()
    graph_equal = all(torch.equal(getattr(ens, k), v) for k, v in end.items())
    ens.sync()

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
        "captured_graphs": {"iterations_per_graph": RAGGED_GRAPH_STEPS,
                            "ms_per_step": 1e3 * wall_graph / args.steps,
                            "state_equals_eager": graph_equal},
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
    # VALU issue bound: one wave64 instruction occupies a SIMD (16 lanes) for
    # 4 cycles; 1,024 SIMDs at 2.4 GHz
    issue_ms = pairs * VALU_INSTS_PER_PAIR[hs] / 64.0 * 4.0 / 1024 / 2.4e9 * 1e3
    executed = pairs * FP64_OPS_PER_PAIR[hs] / (kern_ms * 1e-3) / 1e12
    roof = {"bound": "fp64-valu", "peak": MI355X_FP64_TFLOPS, "unit": "TFLOP/s",
            "traffic": None, "kernel_ms": kern_ms,
            "executed_fp64_tflops": executed, "executed_frac": executed / MI355X_FP64_TFLOPS,
            "valu_issue_bound_ms": issue_ms, "valu_issue_frac": issue_ms / kern_ms}
    if hs:
        # no operation-by-operation count of the half-space pair: the
        # hardware-counted FP64 operations stand in for the algorithmic ones
        roof.update(achieved=executed, frac=executed / MI355X_FP64_TFLOPS,
                    achieved_basis="hardware-counted FP64 operations per pair")
    else:
        achieved = pairs * FLOPS_PER_PAIR / (kern_ms * 1e-3) / 1e12
        roof.update(achieved=achieved, frac=achieved / MI355X_FP64_TFLOPS,
                    achieved_basis="algorithmic FP64 operations per pair (division = 1)")
    res = {
        "metric": "TDE evaluations/s (N x N assembly, FP64" + (", half space)" if hs else ")"),
        "value": 2.0 * pairs * args.steps / wall,
        "unit": "evaluations/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic undulating fault mesh, seed 0",
        "config": {"workload": "kernel assembly, undulating fault 100x25 km, N=20000, unit Ss "
                               "and Ds, six traction operators, FP64"
                               + (", half space (free surface z=0)" if hs else "")},
        "roofline": roof,
    }
    if hs:
        return res
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
    ap.add_argument("--planes", type=int, default=3, choices=[2, 3],
                    help="matvec: row planes (2: the Bs, Bd operators skipped)")
    ap.add_argument("--no-graph", action="store_true",
                    help="hstep: time the headline ensemble with eager launches only")
    ap.add_argument("--accumulate", default=None, choices=["f64", "mfma", "mfma_comp"],
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
    ap.add_argument("--record", type=int, default=62500,
                    help="learned: iterations of the reference controller logged for training")
    ap.add_argument("--epochs", type=int, default=20, help="learned: training epochs")
    ap.add_argument("--cycle-years", type=float, default=1120.0,
                    help="learned: simulated years of the evaluation (about ten cycles)")
    ap.add_argument("--mlp-out", default=str(STEP_MLP), help="learned: where the network goes")
    ap.add_argument("--max-wall", type=float, default=900.0,
                    help="wall-time cap (s) of one cycle run")
    ap.add_argument("--t-end-years", type=float, default=C3_T_END_YEARS,
                    help="hstep: simulated interval of every draw")
    ap.add_argument("--no-learned", action="store_true",
                    help="hstep: the batched variant with the reference controller")
    ap.add_argument("--no-accuracy", action="store_true",
                    help="hstep: skip the FP64 reference run of every draw")
    ap.add_argument("--block-sample", type=int, default=500,
                    help="matvec: blocks of each kind whose stored entries are compared with "
                         "the FP64 assembly (0: none)")
    ap.add_argument("--tuning", default=None,
                    help="hstep/matvec/learned: fixed matrix-core configuration "
                         "'depth,waves,order,p1_depth,p1_waves' instead of the autotuner")
    ap.add_argument("--noise-floor", action="store_true",
                    help="hstep: also rerun the FP64 draws from dt = 2e-3 s (event-onset "
                         "spread of the reference method itself)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (the "
                                   "multi-GPU design is documented, not built)"}))
        sys.exit(2)
    if args.workload != "learned":
        args.nb = 16 if args.nb is None else args.nb
    if args.accumulate is None:
        # the headline's batched products run on the matrix cores; the other
        # workloads default to FP64 sums
        args.accumulate = "mfma" if args.workload in ("hstep", "learned") else "f64"
    if args.workload == "hstep":
        # the timed region is the whole simulated interval; --steps is ignored
        args.warmup = 20 if args.warmup is None else args.warmup
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
        args.warmup = 20 if args.warmup is None else args.warmup
        args.nb = 16 if args.nb is None else args.nb
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
