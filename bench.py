#!/usr/bin/env python3
"""Benchmark of the particle-filter step on one MI355X.

Workload (the reference's headline configuration that fits one GPU):
Lorenz-96 d=40, F=8, additive noise sd 0.1 per 0.05 observation interval,
DOPRI5 (rtol = atol = 1e-5, h0 = 0.05), observations of the even dimensions
with N(0,1) noise, N = 2^20 particles with FP32 state and FP64 log-weights,
systematic resampling when ESS < 0.5 N.  A "step" is one observation interval
of the bootstrap filter over all particles: device-side resampling decision,
resample + ancestor gather (fused into the transition's read), DOPRI5
transition with in-register Philox noise, observation log-density, and the
ESS / log-likelihood reduction.  Observations are synthetic (a truth
trajectory simulated with the library itself); particle-steps/s counts
N x steps over the timed region.

Other workloads (`--workload`), each printing one JSON line of the same form:
  pz     predator-prey bootstrap PF, N=4096 FP64, RK4 h=0.1, systematic
         resampling every step, T=100 daily observations of log P;
  smc2   batched inner filters of SMC^2 on the predator-prey model: 1024
         parameter particles (mu, sigma, c, e from seeded log-normal priors),
         inner N_j = floor(exp(U(log 256, log 16384))), one persistent launch
         per observation interval replayed as a hipGraph, autotuned workgroup;
  l96-bi the default workload on the Lorenz-96 model compiled from the LibBi
         model file libbi_amd/bi/Lorenz96d40.bi (initial x ~ N(0,1); the
         hand-written model is timed beside it from the same start);
  smc2-bi the same on the predator-prey model compiled from the LibBi model
         file libbi_amd/bi/PZ.bi (theta = mu, sigma; the hand-written model
         is timed beside it with the same theta);
  copy   a plain 1 GiB device-to-device copy: the achievable HBM ceiling, and
         the known-size kernel that scales the FETCH_SIZE / WRITE_SIZE
         counters into bytes;
  valu   the vector-ALU issue-rate probe (FP32 FMA, integer, FP32 mul/add
         mixes): the achievable VALU ceiling, and the known-count kernel that
         scales the SQ_INSTS_VALU counter;
  guided the Lorenz-96 d=200 filter with a learned guided proposal
         (301-256-256-256-400 BF16 MFMA network, trained on the GPU before the
         timed region) and the exact importance-weight correction, N=2^18
         particles for each of 64 forcing values F in [7.5, 8.5], stepped
         together; reports its rate and, against the bootstrap filter with the
         same N, the ESS and the log-likelihood spread over two seeds;
  surrogate  the L96 filter above with the DOPRI5 flow replaced by the
         41-256-256-256-40 BF16 MFMA network, trained on the GPU on pairs from
         the exact integrator before the timed region; reports its rate, the
         exact filter's rate, and the log-likelihood difference over several
         seeds (same observations, same noise streams).

Prints one JSON line.  `--gpus N` with N > 1 exits: the build runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

D = 40
P_DEFAULT = 1 << 20
T_SCHEDULE = 500
DT = 0.05
HBM_PEAK_GBS = 8000.0
# Per-launch figures of k_transition<L96<float,40,4>,DOPRI5> on the command
# `bench.py --gpus 1 --steps 20 --warmup 5` (N = 2^20, the 20 timed launches,
# intervals 55..74, every one resampled), from rocprofv3 --pmc runs of that
# command with --no-cpu-baseline:
#  - HBM bytes: FETCH_SIZE x 2.000 + WRITE_SIZE x 1.000 (the factors the same
#    counters show for a 1 GiB device copy, bench.py --workload copy);
#  - VALU wave-instructions (SQ_INSTS_VALU, which reads exactly the known count
#    of the issue-rate probe) and the share of them in the FP32-rate classes
#    (SQ_INSTS_VALU_FMA_F32, _ADD_F32, _MUL_F32, _TRANS_F32, _CVT, _INT64); the
#    rest (SQ_INSTS_VALU_INT32 and the moves, DPP, compares, selects and bit
#    operations no class counter covers) issue at the integer probe's rate.
L96_TRAFFIC_BYTES = 396_513_232
L96_VALU_PER_LAUNCH = 1.289e8
L96_VALU_FP32_SHARE = 0.735


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", default="l96",
                    choices=["l96", "l96-bi", "pz", "smc2", "smc2-bi", "copy", "surrogate",
                             "guided", "valu", "primitives"])
    ap.add_argument("--thetas", type=int, default=64)
    ap.add_argument("--train-pairs", type=int, default=10_000_000)
    ap.add_argument("--train-steps", type=int, default=24000)
    ap.add_argument("--seeds", type=int, default=3)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--settle", type=int, default=50,
                    help="untimed filter steps before the warm-up (Lorenz-96 workloads): the "
                         "initial cloud's transient")
    ap.add_argument("--particles", type=int, default=P_DEFAULT)
    ap.add_argument("--resampler", default="systematic",
                    choices=["systematic", "stratified", "multinomial"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-particles", type=int, default=1 << 16)
    ap.add_argument("--precision-steps", type=int, default=20,
                    help="timed steps re-run by the FP64 oracle for the logZ error")
    ap.add_argument("--cpu-steps", type=int, default=60)
    return ap.parse_args()


def transition_bytes(P, resampled):
    """Algorithmic HBM bytes of one fused transition launch: the state is read
    (through the ancestor map when resampled) and written once (2 d 4 B); the
    log-weight is written (8 B) and read unless the step resampled (then it
    starts from logLikelihood); a resampling step reads the ancestor (4 B)."""
    per = 2 * D * 4 + 8 + (4 if resampled else 8)
    return P * per


def cpu_model():
    """The host CPU's model name (from /proc/cpuinfo), or None."""
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(args, y, mean0):
    """The FP64 OpenMP restatement of the reference host filter (the oracle),
    timed on a bounded sample of the same workload: the median of five runs."""
    import oracle
    threads = len(os.sched_getaffinity(0))
    m = oracle.make_model("l96", d=D)
    P, T = args.cpu_particles, args.cpu_steps
    times = []
    for rep in range(5):
        t0 = time.perf_counter()
        r = oracle.pf_run(m, y[:T], P, DT, 1234 + rep, resampler=oracle.SYSTEMATIC, ess_rel=0.5,
                          mean0=mean0)
        times.append(time.perf_counter() - t0)
    el = float(np.median(times))
    cpu = {
        "value": P * T / el,
        "unit": "particle-steps/s",
        "cores": int(os.environ.get("OMP_NUM_THREADS", threads)),
        "kind": "port",
        "sample": f"L96 d=40 DOPRI5 FP64 bootstrap PF, N={P}, {T} observation intervals, "
                  f"median of 5 runs ({', '.join(f'{t:.2f}' for t in times)} s), OpenMP over "
                  "particles, serial resampling",
        "cpu_model": cpu_model(),
        "logZ": r["logZ"],
    }
    return cpu


def oracle_model_of(oracle, gm):
    """The oracle's model matching a device model (FP32 kernels draw 24-bit normals)."""
    m = oracle.Model()
    m.model, m.d, m.integrator = gm.model, gm.d, gm.integrator
    m.nsteps, m.host_quirk = gm.nsteps, gm.host_quirk
    from libbi_amd import capi
    m.noise_bits = 24 if gm.precision == capi.F32 else 32
    m.h0, m.atol, m.rtol = gm.h0, gm.atol, gm.rtol
    for i in range(16):
        m.par[i] = gm.par[i]
    return m


def precision_history(args, ctx, gm, y, init_mean, timed):
    """Re-runs the timed filter (untimed, same seed: the device filter is
    deterministic) recording its ancestors, and returns its state at the start
    of the timed steps and its resampling history over the first
    args.precision_steps of them.  The re-run's increments must equal the
    timed run's bit for bit."""
    from libbi_amd import capi
    res = {"systematic": capi.SYSTEMATIC, "stratified": capi.STRATIFIED,
           "multinomial": capi.MULTINOMIAL}[args.resampler]
    P, W = args.particles, args.settle + args.warmup
    K = min(args.steps, args.precision_steps)
    pf = capi.ParticleFilter(ctx, gm, P, T_SCHEDULE, DT, seed=77, resampler=res, ess_rel=0.5,
                             record=True)
    pf.set_observations(y)
    pf.init(init_mean)
    pf.step(W)
    out = pf.outputs()
    X, _, lws = pf.state()
    X = X[:, :P].cpu().numpy()
    lws = lws.cpu().numpy()
    ll_W, ess_prev = out["logLikelihood"], (out["ess"][W - 1] if W > 0 else 0.0)
    pf.step(K)
    out = pf.outputs()
    anc = np.array([pf.ancestors(k) for k in range(W, W + K)])
    pf.close()
    same = bool(np.array_equal(out["logIncrements"][W:W + K], timed["logIncrements"][W:W + K]))
    return dict(model=gm, resampler=res, X=X, lws=lws, ll=ll_W, ess=ess_prev, K=K,
                resampled=out["resampled"][W:W + K], ancestors=anc,
                increments=out["logIncrements"][W:W + K], rerun_identical=same,
                logZ=timed["logLikelihood"])


def precision_check(oracle, args, y, h):
    """The FP64 oracle restarted from the FP32 run's state at the first timed
    step (FP32 values are exact in FP64) and stepped on the FP32 run's
    resampling history: the log-likelihood over those steps, FP32 against
    FP64.  Lorenz-96's chaotic growth (about exp(1.7 t)) of the FP32 rounding
    is part of the measured error."""
    P, W, K = args.particles, args.settle + args.warmup, h["K"]
    om = oracle_model_of(oracle, h["model"])
    of = oracle.StepFilter(om, P, DT, 77, resampler=h["resampler"], ess_rel=0.5)
    t0 = time.perf_counter()
    of.set_state(W, h["X"], h["lws"], h["ll"], h["ess"])
    li = np.zeros(K)
    for j in range(K):
        r = int(h["resampled"][j])
        rc, li[j] = of.step(y[W + j], r, h["ancestors"][j] if r else None)
        if rc != oracle.OK:
            raise RuntimeError(f"oracle step {W + j} failed ({rc})")
    el = time.perf_counter() - t0
    of.close()
    z32 = float(np.sum(h["increments"]))
    z64 = float(np.sum(li))
    return {
        "logZ": h["logZ"],
        "timed_steps_checked": f"{W}..{W + K - 1}",
        "loglik_fp32": z32,
        "loglik_fp64": z64,
        "logZ_rel_err_vs_fp64": abs(z32 - z64) / abs(z64),
        "max_step_increment_diff": float(np.max(np.abs(h["increments"] - li))),
        "rerun_identical": h["rerun_identical"],
        "oracle_seconds": el,
        "method": "FP64 oracle restarted from the FP32 state at the first timed step, on the "
                  "FP32 run's resampling flags and ancestors, same Philox noise",
    }


def _timed(ctx, run):
    import torch
    ctx.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run()
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def _line(value, args, ms, dtype, config, roofline, cpu):
    line = {
        "metric": "particle-steps/s",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic (truth trajectory and observations simulated by the library)",
        "config": config,
        "roofline": roofline,
    }
    if cpu is not None:
        line["cpu_baseline"] = cpu
    return line


PZ_P, PZ_T = 4096, 100


def bench_pz(args):
    """Predator-prey bootstrap PF: N=4096 FP64, RK4 h=0.1, resampling every step."""
    from libbi_amd import capi, data
    ctx = capi.Context(0)
    gm = capi.make_model("pz", precision=capi.F64)
    _, y = data.simulate(ctx, gm, [2.0, 2.0], PZ_T, 1.0, seed=2024)
    steps = min(args.steps, PZ_T - args.warmup)

    def filt(use_graph):
        pf = capi.ParticleFilter(ctx, gm, PZ_P, PZ_T, 1.0, seed=77, ess_rel=1.0,
                                 use_graph=use_graph)
        pf.set_observations(y)
        pf.init()
        pf.step(args.warmup)
        return pf

    args.steps = steps
    pf = filt(True)
    el = _timed(ctx, lambda: pf.step(steps))
    out = pf.outputs()
    pf.close()
    pf = filt(False)  # kernel durations with events on direct launches
    pf.set_timing(True)
    pf.step(steps)
    tr_ms, tr_n = pf.kernel_time(0)
    pf.close()
    per = 2 * 2 * 8 + 8 + 4  # state read+write, log-weight write, ancestor read (every step resamples)
    achieved = PZ_P * per * tr_n / (tr_ms * 1e-3) / 1e9 if tr_ms > 0 else 0.0
    cpu = None
    if not args.no_cpu_baseline:
        import oracle
        m = oracle.make_model("pz")
        t0 = time.perf_counter()
        reps = 0
        while time.perf_counter() - t0 < 10.0:
            oracle.pf_run(m, y, PZ_P, 1.0, 1234, ess_rel=1.0)
            reps += 1
        cel = time.perf_counter() - t0
        cpu = {"value": PZ_P * PZ_T * reps / cel, "unit": "particle-steps/s",
               "cores": int(os.environ.get("OMP_NUM_THREADS", len(os.sched_getaffinity(0)))),
               "kind": "port", "cpu_model": cpu_model(),
               "sample": f"PZ FP64 RK4 bootstrap PF N={PZ_P}, T={PZ_T}, {reps} runs ({cel:.1f} s)"}
    cfg = {"workload": "predator-prey bootstrap PF, N=4096 FP64, RK4 h=0.1, systematic "
                       "resampling every step, T=100 daily observations of log P",
           "particles": PZ_P, "resampled_steps": int(out["resampled"].sum())}
    roof = {"bound": "hbm", "kernel": "k_transition<PZ<double>,RK4>", "achieved": achieved,
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
            "traffic": None, "avg_launch_ms": tr_ms / max(tr_n, 1), "launches": tr_n,
            "note": "launch-latency bound at N=4096"}
    print(json.dumps(_line(PZ_P * steps / el, args, 1e3 * el / steps, "f64", cfg, roof, cpu)),
          flush=True)
    ctx.close()


SMC2_B = 1024


def smc2_population(B=SMC2_B, seed=2024):
    """Inner sizes N_j = floor(exp(U(log 256, log 16384))) and theta_j =
    (mu, sigma, c, e) from log-normal priors around the model's values."""
    rng = np.random.default_rng(seed)
    N = np.floor(np.exp(rng.uniform(np.log(256), np.log(16384), B))).astype(np.int32)
    mu = np.exp(rng.normal(np.log(0.5), 0.2, B))
    sigma = np.exp(rng.normal(np.log(0.2), 0.2, B))
    c = np.exp(rng.normal(np.log(0.25), 0.1, B))
    e = np.exp(rng.normal(np.log(0.3), 0.1, B))
    return N, np.stack([mu, sigma, c, e], axis=1)


def bench_smc2(args, generated=False):
    """One observation interval of all 1024 inner filters per step.  generated:
    the predator-prey model compiled from libbi_amd/bi/PZ.bi, whose theta is
    (mu, sigma) (c and e are constants of the file); the hand-written model is
    then also timed with the same two-parameter theta."""
    from libbi_amd import capi, data
    ctx = capi.Context(0)
    gm = capi.make_model("pz", precision=capi.F64)
    T = args.warmup + args.steps
    _, y = data.simulate(ctx, gm, [2.0, 2.0], T, 1.0, seed=2024)
    N, theta = smc2_population()
    total = int(N.sum())
    if generated:
        theta = np.ascontiguousarray(theta[:, :2])
        hand_rate = _batch_rate(ctx, gm, N, theta, y, args)
        gm = capi.generated_model("PZ", precision=capi.F64, h0=gm.h0)
    bf = capi.BatchFilter(ctx, gm, N, theta, T, 1.0, seed=77, ess_rel=0.5, use_graph=True,
                          block=0)
    bf.set_observations(y)
    bf.init()
    bf.step(args.warmup)
    el = _timed(ctx, lambda: bf.step(args.steps))
    blk, grid, tune_ms = bf.tuning()
    items, oversub, table = bf.tuning_table()
    out = bf.outputs()
    bf.close()
    # kernel duration with events (direct launches at the tuned configuration, same steps)
    bt = capi.BatchFilter(ctx, gm, N, theta, T, 1.0, seed=77, ess_rel=0.5, block=blk,
                          items=items, oversub=oversub)
    bt.set_observations(y)
    bt.init()
    bt.step(args.warmup)
    bt.set_timing(True)
    bt.step(args.steps)
    k_ms, k_n = bt.kernel_time()
    bt.close()
    res = out["resampled"][:, args.warmup:T]
    per_fixed = 2 * 2 * 8 + 2 * 8   # state read + write, log-weight read + write
    per_res = 2 * 4                 # ancestor index written by the resampler, read by the gather
    bytes_total = per_fixed * total * args.steps + per_res * float((res * N[:, None]).sum())
    achieved = bytes_total / (k_ms * 1e-3) / 1e9 if k_ms > 0 else 0.0
    # the same filters one after another through the single-filter path (a sample)
    nseq = 16
    seq_ps, seq_el = 0, 0.0
    for j in range(nseq):
        if generated:
            m = capi.generated_model("PZ", precision=capi.F64, mu=theta[j, 0],
                                     sigma=theta[j, 1])
        else:
            m = capi.make_model("pz", precision=capi.F64, mu=theta[j, 0], sigma=theta[j, 1],
                                c=theta[j, 2], e=theta[j, 3])
        pf = capi.ParticleFilter(ctx, m, int(N[j]), T, 1.0, 77, ess_rel=0.5, batch=j,
                                 use_graph=True)
        pf.set_observations(y)
        pf.init()
        pf.step(args.warmup)
        seq_el += _timed(ctx, lambda: pf.step(args.steps))
        seq_ps += int(N[j]) * args.steps
        pf.close()
    cpu = None
    if not args.no_cpu_baseline:
        import oracle
        om = oracle.make_model("pz")
        t0 = time.perf_counter()
        cps, j = 0, 0
        while time.perf_counter() - t0 < 15.0 and j < SMC2_B:
            for i, v in enumerate(theta[j]):
                om.par[i] = v
            oracle.pf_run(om, y, int(N[j]), 1.0, 77, ess_rel=0.5, batch=j)
            cps += int(N[j]) * T
            j += 1
        cel = time.perf_counter() - t0
        cpu = {"value": cps / cel, "unit": "particle-steps/s",
               "cores": int(os.environ.get("OMP_NUM_THREADS", len(os.sched_getaffinity(0)))),
               "kind": "port", "cpu_model": cpu_model(),
               "sample": f"the first {j} inner filters run one after another over T={T} "
                         f"({cel:.1f} s), OpenMP over particles"}
    cfg = {"workload": ("batched SMC^2 inner filters, predator-prey "
                        + ("generated from libbi_amd/bi/PZ.bi (theta = mu, sigma)"
                           if generated else "(theta = mu, sigma, c, e)")
                        + ", 1024 theta-particles, N_j log-uniform in [256,16384], FP64, RK4 "
                          "h=0.1, systematic at ESS<0.5N_j, one persistent launch per step "
                          "(hipGraph)"),
           "particles": total, "filters": SMC2_B, "block": blk, "grid": grid,
           "tune_ms": [float(v) for v in tune_ms],
           "tile_items": items, "grid_oversub": oversub, "configs_timed": len(table),
           "resampled_fraction": float(res.mean()),
           "sequential_pf_particle_steps_per_s": seq_ps / seq_el,
           "sequential_pf_sample": f"filters 0..{nseq - 1} one after another through bihip_pf "
                                   "(hipGraph per step)"}
    if generated:
        cfg["handwritten_same_theta_particle_steps_per_s"] = hand_rate
    kname = "k_batch_step<Gen_PZ<double>,RK4>" if generated else "k_batch_step<PZ<double>,RK4>"
    roof = {"bound": "hbm", "kernel": kname, "achieved": achieved,
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
            "traffic": None, "avg_launch_ms": k_ms / max(k_n, 1), "launches": k_n}
    print(json.dumps(_line(total * args.steps / el, args, 1e3 * el / args.steps, "f64", cfg,
                           roof, cpu)), flush=True)
    ctx.close()


def _batch_rate(ctx, gm, N, theta, y, args):
    """Particle-steps/s of the batched filters (hipGraph replay, tuned block)."""
    from libbi_amd import capi
    T = args.warmup + args.steps
    bf = capi.BatchFilter(ctx, gm, N, theta, T, 1.0, seed=77, ess_rel=0.5, use_graph=True,
                          block=0)
    bf.set_observations(y)
    bf.init()
    bf.step(args.warmup)
    el = _timed(ctx, lambda: bf.step(args.steps))
    bf.close()
    return int(N.sum()) * args.steps / el


def bench_surrogate(args):
    """Surrogate-flow filter against the exact-integrator filter."""
    import torch
    from libbi_amd import capi, data, surrogate
    ctx = capi.Context(0)
    t0 = time.perf_counter()
    # pairs from the model's prior predictive: an ensemble of 2^15 drawn around
    # the filter's initial mean, recorded at every interval from the first
    # (so the pairs cover the states the filter visits over its run)
    X, Y = surrogate.flow_pairs(ctx, args.train_pairs + (1 << 16), seed=3, burn=0,
                                chunk=1 << 15)
    n = args.train_pairs
    q = args.train_steps // 4
    net, norm, rep = surrogate.train(ctx, steps=(q, q, q, args.train_steps - 3 * q),
                                     lrs=(2e-3, 1e-3, 3e-4, 1e-4), pairs=(X[:n], Y[:n]), seed=3)
    rmse, inc = surrogate.flow_error(ctx, net, norm, X[n:], Y[n:])
    del X, Y
    train_s = time.perf_counter() - t0
    mean0 = data.l96_spinup(ctx, d=D)
    gm = capi.make_model("l96", d=D, precision=capi.F32, F=8.0, noise_sd=0.1, obs_sd=1.0,
                         init_sd=1.0, atol=1e-5, rtol=1e-5, h0=DT)
    T = args.warmup + args.steps
    _, y = data.simulate(ctx, gm, mean0, T, DT, seed=2024)
    P = args.particles

    def run(seed, use_surrogate, timed=False):
        pf = capi.ParticleFilter(ctx, gm, P, T, DT, seed=seed, ess_rel=0.5, use_graph=True)
        if use_surrogate:
            pf.set_surrogate(net, norm)
        pf.set_observations(y)
        pf.init(mean0)
        pf.step(args.warmup)
        el = _timed(ctx, lambda: pf.step(args.steps))
        z = pf.term()
        pf.close()
        return z, el

    res = {"exact": [], "surrogate": []}
    el = {"exact": 0.0, "surrogate": 0.0}
    for seed in range(args.seeds):
        for kind in ("exact", "surrogate"):
            z, e = run(100 + seed, kind == "surrogate")
            res[kind].append(z)
            el[kind] += e
    diff = np.array(res["surrogate"]) - np.array(res["exact"])
    # kernel duration of the fused surrogate step (HIP events on direct launches)
    pf = capi.ParticleFilter(ctx, gm, P, T, DT, seed=100, ess_rel=0.5)
    pf.set_surrogate(net, norm)
    pf.set_observations(y)
    pf.init(mean0)
    pf.step(args.warmup)
    pf.set_timing(True)
    pf.step(args.steps)
    k_ms, k_n = pf.kernel_time(0)
    pf.close()
    # the exact transition's kernel time on the same steps
    pf = capi.ParticleFilter(ctx, gm, P, T, DT, seed=100, ess_rel=0.5)
    pf.set_observations(y)
    pf.init(mean0)
    pf.step(args.warmup)
    pf.set_timing(True, transition_only=True)
    pf.step(args.steps)
    x_ms, x_n = pf.kernel_time(0)
    pf.close()
    local = surrogate_local_error(ctx, gm, net, norm, y, mean0, P, min(T, 40))
    batched = surrogate_batched(ctx, gm, y, mean0, args)
    flops = 2.0 * sum(a * b for a, b in zip(surrogate.DIMS[:-1], surrogate.DIMS[1:])) * P
    achieved = flops * k_n / (k_ms * 1e-3) / 1e12 if k_ms > 0 else 0.0
    v_sur = P * args.steps * args.seeds / el["surrogate"]
    v_ex = P * args.steps * args.seeds / el["exact"]
    cfg = {"workload": "Lorenz-96 d=40 bootstrap PF step with the DOPRI5 flow replaced by a "
                       "41-256-256-256-40 BF16 MFMA surrogate, N=2^20 FP32 state / FP64 "
                       "weights, systematic resampling at ESS<0.5N",
           "particles": P, "training_pairs": rep["pairs"], "training_steps": rep["steps"],
           "training_seconds": train_s, "final_minibatch_loss": rep["final_minibatch_loss"],
           "flow_rmse": rmse, "flow_increment_rms": inc,
           "exact_particle_steps_per_s": v_ex, "speedup_vs_exact": v_sur / v_ex,
           "logZ_exact": res["exact"], "logZ_surrogate": res["surrogate"],
           "logZ_diff_mean": float(diff.mean()),
           "logZ_diff_sd": float(diff.std(ddof=1)) if len(diff) > 1 else None,
           "logZ_exact_sd_over_seeds": float(np.std(res["exact"], ddof=1))
           if len(diff) > 1 else None,
           "loglik_error_per_interval": local,
           # what the network could gain at best: its FLOPs at the full BF16
           # MFMA rate against the exact DOPRI5 transition's measured time
           "exact_transition_ms": x_ms / max(x_n, 1),
           "network_ms_at_mfma_peak": flops / 2.5e15 * 1e3,
           "speedup_bound_at_mfma_peak": (x_ms / max(x_n, 1)) / (flops / 2.5e15 * 1e3),
           "batched_form": batched}
    roof = {"bound": "mfma", "kernel": "k_surrogate<40,256,4,step>", "achieved": achieved,
            "peak": 2500.0, "unit": "TFLOP/s", "frac": achieved / 2500.0, "traffic": None,
            "avg_launch_ms": k_ms / max(k_n, 1), "launches": k_n,
            "flop_per_launch": flops}
    print(json.dumps(_line(v_sur, args, 1e3 * el["surrogate"] / (args.steps * args.seeds),
                           "bf16", cfg, roof, None)), flush=True)
    ctx.close()


def surrogate_batched(ctx, gm, y, mean0, args, B=256, N=4096, lone=16):
    """The batched form: B filters of N particles over forcing values
    F in [7.5, 8.5] (SMC^2's inner filters for the Lorenz-96 forcing) as one
    filter bank, with the exact DOPRI5 flow and with a surrogate trained over
    that forcing range, against the plain GPU filter (bihip_pf, hipGraph
    replay) run for one forcing after another, timed over `lone` of them; then
    the whole SMC^2 over F on exact and on surrogate banks.  The exact bank's
    filters equal the plain filters bit for bit (same log-likelihood error);
    the surrogate's error per interval is measured at two forcings."""
    from libbi_amd import capi, surrogate
    T = args.warmup + args.steps
    Fs = np.linspace(7.5, 8.5, B)
    t0 = time.perf_counter()
    net, norm, rep = surrogate.train_over_forcings(ctx, np.linspace(7.5, 8.5, 5),
                                                   args.train_pairs // 5, seed=13)
    train_s = time.perf_counter() - t0
    local = {}
    for F in (7.75, 8.25):
        mF = capi.Model.from_buffer_copy(gm)
        mF.par[0] = F
        local[str(F)] = surrogate_local_error(ctx, mF, net, norm, y, mean0, 1 << 18, 20, F=F)

    def bank_rate(use_surrogate):
        b = capi.FilterBank(ctx, gm, N, T, DT, 100, Fs, ess_rel=0.5, use_graph=True)
        if use_surrogate:
            b.set_surrogate(net, norm)
        b.set_observations(y)
        b.init(mean0)
        b.step(args.warmup)
        el = _timed(ctx, lambda: b.step(args.steps))
        b.close()
        return B * N * args.steps / el

    fl = []
    for j in range(lone):
        m = capi.Model.from_buffer_copy(gm)
        m.par[0] = Fs[j]
        pf = capi.ParticleFilter(ctx, m, N, T, DT, seed=100, ess_rel=0.5, batch=j, use_graph=True)
        pf.set_observations(y)
        pf.init(mean0)
        pf.step(args.warmup)
        fl.append(pf)

    def run():
        for pf in fl:
            pf.step(args.steps)

    plain = lone * N * args.steps / _timed(ctx, run)
    for pf in fl:
        pf.close()
    exact = bank_rate(False)
    sur = bank_rate(True)

    def smc2(use_surrogate, Ts=20, seed=7):
        # the whole SMC^2 over the forcing (prior log F ~ N(log 8, 0.03^2), one
        # random-walk move after each outer resampling) on banks; every inner
        # particle-step it runs, proposals included, over its wall time
        s = capi.SMC2(ctx, gm, None, Ts, [np.log(8.0)], [0.03], [0.01], dt=DT, seed=seed,
                      ess_rel_outer=0.5, use_graph=True, bank=(B, N), mean0=mean0)
        if use_surrogate:
            s.set_surrogate(net, norm)
        s.set_observations(y[:Ts])
        el = _timed(ctx, s.run)
        out = s.outputs()
        s.close()
        moved = [k + 1 for k in range(Ts) if out["resampled"][k]]
        steps = B * N * (Ts + sum(k for k in moved if k <= Ts))
        return {"logZ": out["logZ"], "outer_resamplings": int(out["resampled"].sum()),
                "accepted_moves": int(out["accepts"].sum()), "seconds": el,
                "inner_particle_steps_per_s": steps / el,
                "F_mean": float(out["theta"].mean())}

    seeds = (7, 8, 9)
    runs_ex = [smc2(False, seed=q) for q in seeds]
    runs_su = [smc2(True, seed=q) for q in seeds]
    s_ex, s_su = runs_ex[0], runs_su[0]
    z_ex = np.array([r["logZ"] for r in runs_ex])
    z_su = np.array([r["logZ"] for r in runs_su])
    return {"filters": B, "particles_per_filter": N,
            "plain_filter_particle_steps_per_s": plain,
            "exact_bank_particle_steps_per_s": exact,
            "surrogate_bank_particle_steps_per_s": sur,
            "exact_bank_vs_plain": exact / plain, "surrogate_bank_vs_plain": sur / plain,
            "error": "exact bank: the plain filter's (bit-identical filters); surrogate bank: "
                     "loglik_error_per_interval",
            "smc2_exact_banks": s_ex, "smc2_surrogate_banks": s_su,
            "smc2_surrogate_vs_plain": s_su["inner_particle_steps_per_s"] / plain,
            "smc2_logZ_exact_by_seed": z_ex.tolist(), "smc2_logZ_surrogate_by_seed": z_su.tolist(),
            "smc2_logZ_seed_sd_exact": float(z_ex.std(ddof=1)),
            "smc2_logZ_mean_surrogate_minus_exact": float(z_su.mean() - z_ex.mean()),
            "smc2_F_mean_exact": float(np.mean([r["F_mean"] for r in runs_ex])),
            "smc2_F_mean_surrogate": float(np.mean([r["F_mean"] for r in runs_su])),
            "forcing_range_network": {"training_seconds": train_s, **rep,
                                      "loglik_error_per_interval": local}}


def surrogate_local_error(ctx, gm, net, norm, y, mean0, P, K, seed=8, F=8.0):
    """Log-likelihood error of the surrogate per observation interval: from
    the exact filter's particles before each of the first K steps (resampled
    to equal weights), one interval with the exact FP32 flow and one with the
    surrogate, the same noise, and the observation density; the difference of
    the two log-mean-weights.  Free of Monte Carlo error between the two runs
    (common particles and noise) and of the chaotic growth of the flow error
    along a lineage."""
    import ctypes
    import torch
    from libbi_amd import capi
    pf = capi.ParticleFilter(ctx, gm, P, K, DT, seed=seed, ess_rel=0.5)
    pf.set_observations(y[:K])
    pf.init(mean0)
    g = torch.Generator(device="cuda")
    g.manual_seed(1000 + seed)
    diffs = []
    for k in range(K):
        if k > 0:
            pf.step(1)
        X, _, lws = pf.state()
        w = torch.exp(lws - lws.max())
        C = X[:, :P][:, torch.multinomial(w, P, replacement=True, generator=g)].contiguous()
        Zx = C.clone()
        capi.check(capi.lib().bihip_flow(ctx.h, ctypes.byref(gm), k * DT, (k + 1) * DT, P,
                                         ctypes.c_void_p(Zx.data_ptr()), P, None, None))
        Zs = net.surrogate_flow(norm, F, C)
        eps = 0.1 * torch.randn(C.shape, generator=g, device="cuda", dtype=torch.float32)
        yk = torch.as_tensor(y[k], device="cuda")

        def ll(Z):
            inc = (-0.5 * (yk[:, None] - (Z + eps).double()[0::2]) ** 2).sum(0)
            return float(torch.logsumexp(inc, 0)) - np.log(P)
        diffs.append(ll(Zs) - ll(Zx))
    pf.close()
    d = np.array(diffs)
    return {"intervals": K, "mean_abs_nats": float(np.abs(d).mean()),
            "max_abs_nats": float(np.abs(d).max()), "signed_sum_nats": float(d.sum()),
            "method": "exact filter's clouds, common particles and noise, exact FP32 flow "
                      "against the surrogate over one interval"}


def bench_guided(args):
    """Guided-proposal filters for 64 forcing values against bootstrap filters."""
    import torch
    from libbi_amd import capi, data, guided
    ctx = capi.Context(0)
    t0 = time.perf_counter()
    Fs_train = [7.5, 7.8333, 8.1667, 8.5]
    # triples from each forcing value's prior predictive (an ensemble drawn
    # around the initial mean, recorded at every interval from the first)
    chunks = [guided.triples(ctx, (1 << 21) // len(Fs_train), F=F, seed=11 + i, burn=0,
                             chunk=1 << 12)
              for i, F in enumerate(Fs_train)]
    Phi = torch.cat([c[0] for c in chunks])
    norm = guided.normalisation(Phi)
    norm.par_scale = 0.5
    Us, Ts = [], []
    for (ph, xn, yy), F in zip(chunks, Fs_train):
        U, T = guided.training_tensors(ph, xn, yy, norm, F=F)
        Us.append(U)
        Ts.append(T)
    U, T = torch.cat(Us), torch.cat(Ts)
    del chunks, Phi, Us, Ts
    net = capi.MLP(ctx, guided.DIMS, seed=5)
    net.set_loss(capi.LOSS_GAUSS)
    losses = [net.train(U, T, batch=4096, steps=s_, lr=lr, seed=9 + k)
              for k, (s_, lr) in enumerate(zip((3000, 2000, 1000), (1e-3, 3e-4, 1e-4)))]
    del U, T
    train_s = time.perf_counter() - t0
    P = 1 << 18
    Tn = args.warmup + args.steps
    Fs = np.linspace(7.5, 8.5, args.thetas)
    mean0 = data.l96_spinup(ctx, d=guided.D)
    _, y = data.simulate(ctx, guided.model(8.0), mean0, Tn, DT, seed=2024)

    def bank(use, seed, use_graph=True):
        # every forcing value's filter in one bank: each launch covers all of them
        b = capi.FilterBank(ctx, guided.model(8.0), P, Tn, DT, seed, Fs, ess_rel=0.5,
                            use_graph=use_graph)
        b.set_observations(y)
        b.init(mean0)
        if use:
            b.set_guided(net, norm)
        return b

    res = {}
    for use in (False, True):
        for seed in (1, 2):
            b = bank(use, seed)
            b.step(args.warmup)
            el = _timed(ctx, lambda: b.step(args.steps))
            out = b.outputs()
            b.close()
            res[(use, seed)] = (el, out["logLikelihood"], out["ess"][:, args.warmup:Tn])
    # the same bootstrap filters one after another (bihip_pf, hipGraph replay)
    lone = []
    for j, F in enumerate(Fs):
        pf = capi.ParticleFilter(ctx, guided.model(F), P, Tn, DT, seed=1, ess_rel=0.5, batch=j,
                                 use_graph=True)
        pf.set_observations(y)
        pf.init(mean0)
        pf.step(args.warmup)
        lone.append(pf)

    def run_lone():
        for pf in lone:
            pf.step(args.steps)

    el_lone = _timed(ctx, run_lone)
    for pf in lone:
        pf.close()
    # kernel durations of the bank's guided steps (events on direct launches)
    b = bank(True, 1, use_graph=False)
    b.set_timing(True)
    b.step(args.steps)
    k_ms, k_n = b.kernel_time(3)
    step_ms, _ = b.kernel_time(0)
    b.close()
    flops = 2.0 * sum(a * b_ for a, b_ in zip(guided.DIMS[:-1], guided.DIMS[1:])) * P * len(Fs)
    ps = P * args.thetas * args.steps
    v_g = 2 * ps / (res[(True, 1)][0] + res[(True, 2)][0])
    v_b = 2 * ps / (res[(False, 1)][0] + res[(False, 2)][0])
    spread = {u: float(np.std(res[(u, 1)][1] - res[(u, 2)][1]) / np.sqrt(2)) for u in (False, True)}
    ess = {u: float(np.mean([res[(u, s_)][2].mean() for s_ in (1, 2)])) / P for u in (False, True)}
    cfg = {"workload": "Lorenz-96 d=200 filters with a learned guided proposal (301-256-256-256-400 "
                       "BF16 MFMA) and exact weight correction, N=2^18 FP32 state / FP64 weights "
                       f"for each of {args.thetas} forcing values F in [7.5, 8.5] as one filter "
                       "bank (shared launches), DOPRI5 rtol=atol=1e-5, systematic resampling at "
                       "ESS<0.5N",
           "particles_per_filter": P, "filters": args.thetas, "training_seconds": train_s,
           "final_minibatch_loss": float(losses[-1][-100:].mean()),
           "bootstrap_particle_steps_per_s": v_b,
           "bootstrap_lone_filters_particle_steps_per_s": ps / el_lone,
           "bank_vs_lone_filters": v_b / (ps / el_lone),
           "mean_ess_fraction_guided": ess[True], "mean_ess_fraction_bootstrap": ess[False],
           "logZ_seed_sd_guided": spread[True], "logZ_seed_sd_bootstrap": spread[False],
           "logZ_mean_guided_minus_bootstrap": float(np.mean(
               (res[(True, 1)][1] + res[(True, 2)][1]) / 2 - (res[(False, 1)][1] + res[(False, 2)][1]) / 2))}
    roof = {"bound": "mfma", "kernel": "k_guided<filter> (one launch: every filter of the bank)",
            "achieved": flops / (k_ms / k_n * 1e-3) / 1e12
            if k_ms > 0 else 0.0, "peak": 2500.0, "unit": "TFLOP/s", "traffic": None,
            "avg_launch_ms": k_ms / max(k_n, 1), "launches": k_n, "flop_per_launch": flops,
            "flow_plus_proposal_ms_per_step": step_ms / max(k_n, 1)}
    roof["frac"] = roof["achieved"] / roof["peak"]
    print(json.dumps(_line(v_g, args, 1e3 * (res[(True, 1)][0] + res[(True, 2)][0]) / (2 * args.steps),
                           "f32", cfg, roof, None)), flush=True)
    ctx.close()


def bench_copy(args):
    """Device-to-device copy of 2^28 floats (1 GiB read + 1 GiB written)."""
    import torch
    n = 1 << 28
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(max(args.warmup, 1)):
        dst.copy_(src)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(args.steps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    gbs = 2 * 4 * n / (ms * 1e-3) / 1e9
    print(json.dumps({"metric": "copy GB/s", "value": gbs, "unit": "GB/s", "n_gpus": 1,
                      "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
                      "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                      "dtype": "f32", "data": "synthetic",
                      "config": {"workload": "device copy of 2^28 floats (1 GiB each way)",
                                 "bytes_read": 4 * n, "bytes_written": 4 * n},
                      "frac_of_peak": gbs / HBM_PEAK_GBS}), flush=True)


def bench_primitives(args, N=1 << 24, reps=20):
    """Every standalone primitive of the path at N = 2^24, timed by HIP events
    on the library's stream (the context is bound to torch's current stream):
    calls per second and algorithmic GB/s against the HBM peak.  Entry points
    that return a host value (the reductions) synchronise inside each call, so
    their event time includes that round trip; rocprofv3 --kernel-trace --stats
    of this workload gives every kernel's own duration."""
    import ctypes
    import torch
    from libbi_amd import capi
    ctx = capi.Context(0)
    L = capi.lib()
    c = lambda t: ctypes.c_void_p(t.data_ptr())
    dev = "cuda"
    g = torch.Generator(device=dev)
    g.manual_seed(1)
    lws = 2.0 * torch.randn(N, dtype=torch.float64, device=dev, generator=g)
    Ws = torch.empty(N, dtype=torch.float64, device=dev)
    Os = torch.empty(N, dtype=torch.int32, device=dev)
    os_ = torch.empty(N, dtype=torch.int32, device=dev)
    as_ = torch.empty(N, dtype=torch.int32, device=dev)
    U = torch.rand(N, dtype=torch.float64, device=dev, generator=g)
    alphas = torch.rand(N, dtype=torch.float64, device=dev, generator=g)
    zf = torch.empty(N, dtype=torch.float32, device=dev)
    ud = torch.empty(N, dtype=torch.float64, device=dev)
    d_cols = 40
    X = torch.randn((d_cols, N // 16), dtype=torch.float32, device=dev, generator=g)
    out1, out2 = ctypes.c_double(), ctypes.c_double()
    capi.check(L.bihip_sumexpu_inclusive_scan(ctx.h, c(lws), N, c(Ws), None))
    capi.check(L.bihip_systematic_cumulative_offspring(ctx.h, c(Ws), N, N, 0.37, c(Os)))
    capi.check(L.bihip_cumulative_offspring_to_ancestors_permute(ctx.h, c(Os), N, c(as_)))
    # ancestors of a 2^20-particle resampling (for the gather of 40 FP32 columns)
    Ng = N // 16
    Wg = torch.empty(Ng, dtype=torch.float64, device=dev)
    Og = torch.empty(Ng, dtype=torch.int32, device=dev)
    ag = torch.empty(Ng, dtype=torch.int32, device=dev)
    capi.check(L.bihip_sumexpu_inclusive_scan(ctx.h, c(lws[:Ng]), Ng, c(Wg), None))
    capi.check(L.bihip_systematic_cumulative_offspring(ctx.h, c(Wg), Ng, Ng, 0.37, c(Og)))
    capi.check(L.bihip_cumulative_offspring_to_ancestors_permute(ctx.h, c(Og), Ng, c(ag)))
    moved = int((ag != torch.arange(Ng, device=dev, dtype=torch.int32)).sum())
    Xg = X.clone()
    items = [
        ("ess_reduce", lambda: L.bihip_ess_reduce(ctx.h, c(lws), N, ctypes.byref(out1),
                                                  ctypes.byref(out2)), 8 * N, True),
        ("logsumexp_reduce", lambda: L.bihip_logsumexp_reduce(ctx.h, c(lws), N,
                                                              ctypes.byref(out1)), 8 * N, True),
        ("sumexpu_inclusive_scan", lambda: L.bihip_sumexpu_inclusive_scan(
            ctx.h, c(lws), N, c(Ws), None), 16 * N, False),
        ("systematic_cumulative_offspring", lambda: L.bihip_systematic_cumulative_offspring(
            ctx.h, c(Ws), N, N, 0.37, c(Os)), 12 * N, False),
        ("stratified_cumulative_offspring", lambda: L.bihip_stratified_cumulative_offspring(
            ctx.h, c(Ws), N, N, c(alphas), c(Os)), 20 * N, False),
        ("multinomial_ancestors", lambda: L.bihip_multinomial_ancestors(
            ctx.h, c(Ws), N, N, c(U), c(as_)), 20 * N, False),
        ("cumulative_offspring_to_ancestors_permute",
         lambda: L.bihip_cumulative_offspring_to_ancestors_permute(ctx.h, c(Os), N, c(as_)),
         8 * N, False),
        ("gather_rows (2^20 x 40 FP32, in place)", lambda: L.bihip_gather_rows(
            ctx.h, c(ag), Ng, d_cols, c(Xg), Ng, 4), 4 * Ng + 2 * 4 * d_cols * moved, False),
        ("normals (FP32)", lambda: L.bihip_normals(ctx.h, 5, 0, 1, 2, 0, N, capi.F32, c(zf)),
         4 * N, False),
        ("uniforms", lambda: L.bihip_uniforms(ctx.h, 5, 1, 3, 0, N, c(ud)), 8 * N, False),
    ]
    res = {}
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    for name, fn, nbytes, host in items:
        for _ in range(3):
            capi.check(fn())
        torch.cuda.synchronize()
        e0.record()
        for _ in range(reps):
            capi.check(fn())
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        res[name] = {"ms_per_call": ms, "algorithmic_bytes": nbytes,
                     "GB_s": nbytes / (ms * 1e-3) / 1e9,
                     "frac_of_peak": nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "host_round_trip_in_call": host}
    ctx.close()
    key = "systematic_cumulative_offspring"
    print(json.dumps({"metric": "GB/s of the standalone primitives", "value": res[key]["GB_s"],
                      "unit": "GB/s", "n_gpus": 1, "steps": reps, "warmup": 3,
                      "ms_per_step": res[key]["ms_per_call"], "higher_is_better": True,
                      "scaling": "weak", "vs_baseline": None, "dtype": "f64",
                      "data": "synthetic (log-weights 2 N(0,1), uniforms, normals)",
                      "config": {"workload": f"standalone primitives at N = 2^24 "
                                             f"(value: {key})", "N": N,
                                 "gather_moved_rows": moved},
                      "primitives": res}), flush=True)


def valu_probe(kinds=(0, 1, 2), trips=4096, reps=3):
    """Wave-level VALU instructions per second of the issue-rate probe
    (libbi_amd/lib/libbiprobe.so): eight independent chains per lane, eight
    waves per SIMD, no memory traffic in the loop.  Returns {kind: rate}."""
    import ctypes
    import torch
    L = ctypes.CDLL(os.path.join(HERE, "libbi_amd", "lib", "libbiprobe.so"))
    L.biprobe_valu.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                               ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_double)]
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    blocks = cus * 8  # 4 waves per block: 8 waves per SIMD
    out = torch.empty(blocks * 256, dtype=torch.float32, device="cuda")
    rates = {}
    for k in kinds:
        best = 0.0
        for r in range(reps + 1):
            ms, n = ctypes.c_float(), ctypes.c_double()
            rc = L.biprobe_valu(k, blocks, trips, ctypes.c_void_p(out.data_ptr()),
                                ctypes.byref(ms), ctypes.byref(n))
            if rc != 0:
                raise RuntimeError(f"VALU probe kind {k} failed ({rc})")
            if r > 0:
                best = max(best, n.value / (ms.value * 1e-3))
        rates[k] = best
    return rates, blocks, trips


def bench_valu(args):
    import torch
    rates, blocks, trips = valu_probe()
    props = torch.cuda.get_device_properties(0)
    print(json.dumps({"metric": "VALU wave-instructions/s", "value": rates[0],
                      "unit": "wave-instr/s", "n_gpus": 1, "steps": 3, "warmup": 1,
                      "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                      "dtype": "f32", "data": "synthetic",
                      "config": {"workload": "VALU issue-rate probe: 8 independent chains per "
                                             "lane, 8 waves per SIMD, 64 VALU instructions per "
                                             "loop trip, no memory traffic",
                                 "blocks": blocks, "trips": trips, "cus": props.multi_processor_count},
                      "rates": {"v_fma_f32": rates[0], "v_xor_b32+v_add_u32": rates[1],
                                "v_mul_f32+v_add_f32": rates[2]}}), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the filter is not "
              "sharded across GPUs in this build.", file=sys.stderr)
        sys.exit(2)
    if args.workload == "pz":
        return bench_pz(args)
    if args.workload == "smc2":
        return bench_smc2(args)
    if args.workload == "smc2-bi":
        return bench_smc2(args, generated=True)
    if args.workload == "copy":
        return bench_copy(args)
    if args.workload == "surrogate":
        return bench_surrogate(args)
    if args.workload == "guided":
        return bench_guided(args)
    if args.workload == "primitives":
        return bench_primitives(args)
    if args.workload == "valu":
        return bench_valu(args)
    import torch
    from libbi_amd import capi, data

    if args.settle + args.warmup + args.steps > T_SCHEDULE:
        sys.exit(f"settle + warmup + steps must be at most {T_SCHEDULE}")
    ctx = capi.Context(0)
    mean0 = data.l96_spinup(ctx, d=D)
    gm = capi.make_model("l96", d=D, precision=capi.F32, F=8.0, noise_sd=0.1, obs_sd=1.0,
                         init_sd=1.0, atol=1e-5, rtol=1e-5, h0=DT)
    _, y = data.simulate(ctx, gm, mean0, T_SCHEDULE, DT, seed=2024)
    P = args.particles
    generated = args.workload == "l96-bi"
    if generated:
        # the same filter on the functor compiled from the model file; its
        # initial block draws x ~ N(0, init_sd), so the hand-written model is
        # timed beside it from the same initial distribution (mean 0)
        bm = capi.generated_model("Lorenz96d40", precision=capi.F32, F=8.0, noise_sd=0.1,
                                  obs_sd=1.0, init_sd=1.0, atol=1e-5, rtol=1e-5, h0=DT)
        hand = run_l96(args, ctx, gm, y, np.zeros(D))
        fig = run_l96(args, ctx, bm, y, None)
    else:
        fig = run_l96(args, ctx, gm, y, mean0)
    elapsed, trans_ms, trans_n, res_ms, res_n, red_ms, red_n, resampled, first, timed_out = fig
    bytes_total = sum(transition_bytes(P, bool(r)) for r in resampled)
    achieved = bytes_total / (trans_ms * 1e-3) / 1e9 if trans_ms > 0 else 0.0

    value = P * args.steps / elapsed
    line = {
        "metric": "particle-steps/s",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (truth trajectory and observations simulated by the library)",
        "config": {
            "workload": ("Lorenz-96 d=40 bootstrap PF step" +
                         (" on the model compiled from libbi_amd/bi/Lorenz96d40.bi (four lanes "
                          "per particle), initial x ~ N(0,1)" if generated else "") +
                         ", N=2^20 FP32 state / FP64 weights, "
                         "DOPRI5 rtol=atol=1e-5, obs every 0.05 on even dims N(0,1), "
                         f"{args.resampler} resampling at ESS<0.5N"),
            "particles": P,
            "resampled_steps": int(resampled.sum()),
            "log_weights": "f64",
            "timed_intervals": f"{args.settle + args.warmup}..{args.settle + args.warmup + args.steps - 1} "
                               f"of the T={T_SCHEDULE} schedule, after {args.settle} untimed settling "
                               f"and {args.warmup} warm-up intervals",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": ("k_transition<Gen_Lorenz96d40<float>,DOPRI5,ONE>" if generated
                       else "k_transition<L96<float,40,4>,DOPRI5>"),
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": L96_TRAFFIC_BYTES if P == P_DEFAULT and not generated else None,
            "traffic_source": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE per launch of this "
                              "command's timed window (bench.py --gpus 1 --steps 20 --warmup 5), "
                              "scaled by a 1 GiB copy (factors 2.0 / 1.0)",
            "algorithmic_bytes_per_launch": bytes_total / max(len(resampled), 1),
            "avg_launch_ms": trans_ms / max(trans_n, 1),
            "launches": trans_n,
            "resample_chain_ms_per_step": res_ms / max(res_n, 1),
            "reduce_ms_per_step": red_ms / max(red_n, 1),
            "phase_timing_sample": f"warm-up steps {args.settle + first}..{args.settle + args.warmup - 1}",
        },
    }
    if not generated and P == P_DEFAULT:
        # what bounds the kernel: its VALU issue rate against the ceiling that
        # the issue-rate probe measures now, on this GPU, for its instruction mix
        rates, _, _ = valu_probe(kinds=(0, 1))
        launch_s = trans_ms * 1e-3 / max(trans_n, 1)
        rate = L96_VALU_PER_LAUNCH / launch_s
        ceiling = 1.0 / (L96_VALU_FP32_SHARE / rates[0] + (1.0 - L96_VALU_FP32_SHARE) / rates[1])
        line["roofline"]["limiter"] = "valu"
        line["roofline"]["valu_frac"] = rate / ceiling
        line["roofline"]["valu"] = {
            "wave_instr_per_launch": L96_VALU_PER_LAUNCH,
            "issue_rate": rate,
            "ceiling": ceiling,
            "unit": "wave-instr/s",
            "probe_fp32_fma": rates[0],
            "probe_int32": rates[1],
            "fp32_rate_share": L96_VALU_FP32_SHARE,
            "source": "SQ_INSTS_VALU and its class counters per launch (rocprofv3 --pmc, this "
                      "command's timed window) / this run's average launch time; ceiling = the "
                      "issue-rate probes (bench.py --workload valu) weighted by the mix",
        }
    if generated:
        line["config"]["handwritten_same_init_particle_steps_per_s"] = P * args.steps / hand[0]
        line["config"]["handwritten_same_init_transition_ms"] = hand[1] / max(hand[2], 1)
    line["logZ"] = timed_out["logLikelihood"]
    if not args.no_cpu_baseline:
        check = None
        if not generated and args.precision_steps > 0:
            check = precision_history(args, ctx, gm, y, mean0, timed_out)
        line["cpu_baseline"] = cpu_baseline(args, y, mean0)
        if check is not None:
            # the same oracle re-runs the first timed steps in FP64 at full N
            # on the timed run's resampling history
            import oracle
            line["precision"] = precision_check(oracle, args, y, check)
            line["logZ_rel_err_vs_fp64"] = line["precision"]["logZ_rel_err_vs_fp64"]
    print(json.dumps(line), flush=True)
    ctx.close()


def run_l96(args, ctx, m, y, init_mean):
    """The timed filter of the Lorenz-96 workloads: warm-up steps (the later
    ones with events around every phase), then args.steps timed steps with
    events around the transition only, since each event costs stream time."""
    import torch
    from libbi_amd import capi
    res = {"systematic": capi.SYSTEMATIC, "stratified": capi.STRATIFIED,
           "multinomial": capi.MULTINOMIAL}[args.resampler]
    P = args.particles
    pf = capi.ParticleFilter(ctx, m, P, T_SCHEDULE, DT, seed=77, resampler=res, ess_rel=0.5)
    pf.set_observations(y)
    pf.init(init_mean)

    # settling: the initial cloud (x0 ~ N(mean0, I)) is wider than the
    # filtering distribution for the first intervals, and DOPRI5 takes more
    # steps on it (about 245 us per transition over intervals 0-24, 190 us
    # from interval 50 on, 193 us averaged over the T = 500 run); the timed
    # window starts after it, in the phase that makes up most of the run
    pf.step(args.settle)
    # warm-up: the first steps untimed, the rest with events around every phase
    # (resampling chain and reduction figures); the timed steps carry events
    # around the transition only, since each event costs stream time
    first = min(2, args.warmup)
    pf.step(first)
    pf.set_timing(True)
    pf.step(args.warmup - first)
    ctx.sync()
    res_ms, res_n = pf.kernel_time(1)
    red_ms, red_n = pf.kernel_time(2)
    pf.set_timing(True, transition_only=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    pf.step(args.steps)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    pf.status()
    trans_ms, trans_n = pf.kernel_time(0)
    out = pf.outputs()
    k0, k1 = args.settle + args.warmup, args.settle + args.warmup + args.steps
    resampled = out["resampled"][k0:k1]
    pf.close()
    return elapsed, trans_ms, trans_n, res_ms, res_n, red_ms, red_n, resampled, first, out


if __name__ == "__main__":
    main()
