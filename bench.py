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


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--particles", type=int, default=P_DEFAULT)
    ap.add_argument("--resampler", default="systematic",
                    choices=["systematic", "stratified", "multinomial"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-particles", type=int, default=1 << 18)
    ap.add_argument("--cpu-steps", type=int, default=60)
    return ap.parse_args()


def transition_bytes(P, resampled):
    """Algorithmic HBM bytes of one fused transition launch: the state is read
    (through the ancestor map when resampled) and written once (2 d 4 B); the
    log-weight is written (8 B) and read unless the step resampled (then it
    starts from logLikelihood); a resampling step reads the ancestor (4 B)."""
    per = 2 * D * 4 + 8 + (4 if resampled else 8)
    return P * per


def cpu_baseline(args, y, mean0):
    """The FP64 OpenMP restatement of the reference host filter (the oracle),
    timed on a bounded sample of the same workload."""
    import oracle
    threads = len(os.sched_getaffinity(0))
    m = oracle.make_model("l96", d=D)
    P, T = args.cpu_particles, args.cpu_steps
    t0 = time.perf_counter()
    r = oracle.pf_run(m, y[:T], P, DT, 1234, resampler=oracle.SYSTEMATIC, ess_rel=0.5,
                      mean0=mean0)
    el = time.perf_counter() - t0
    return {
        "value": P * T / el,
        "unit": "particle-steps/s",
        "cores": int(os.environ.get("OMP_NUM_THREADS", threads)),
        "kind": "port",
        "sample": f"L96 d=40 DOPRI5 FP64 bootstrap PF, N={P}, {T} observation intervals "
                  f"({el:.1f} s), OpenMP over particles, serial resampling",
        "logZ": r["logZ"],
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU "
              "sharding of the filter is described in DESIGN.md but not built.",
              file=sys.stderr)
        sys.exit(2)
    import torch
    from libbi_amd import capi, data

    if args.warmup + args.steps > T_SCHEDULE:
        sys.exit(f"warmup + steps must be at most {T_SCHEDULE}")
    ctx = capi.Context(0)
    mean0 = data.l96_spinup(ctx, d=D)
    gm = capi.make_model("l96", d=D, precision=capi.F32, F=8.0, noise_sd=0.1, obs_sd=1.0,
                         init_sd=1.0, atol=1e-5, rtol=1e-5, h0=DT)
    _, y = data.simulate(ctx, gm, mean0, T_SCHEDULE, DT, seed=2024)
    res = {"systematic": capi.SYSTEMATIC, "stratified": capi.STRATIFIED,
           "multinomial": capi.MULTINOMIAL}[args.resampler]
    P = args.particles
    pf = capi.ParticleFilter(ctx, gm, P, T_SCHEDULE, DT, seed=77, resampler=res, ess_rel=0.5)
    pf.set_observations(y)
    pf.init(mean0)

    pf.step(args.warmup)
    ctx.sync()
    pf.set_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    pf.step(args.steps)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    pf.status()
    trans_ms, trans_n = pf.kernel_time(0)
    res_ms, _ = pf.kernel_time(1)
    red_ms, _ = pf.kernel_time(2)
    out = pf.outputs()
    k0, k1 = args.warmup, args.warmup + args.steps
    resampled = out["resampled"][k0:k1]
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
            "workload": "Lorenz-96 d=40 bootstrap PF step, N=2^20 FP32 state / FP64 weights, "
                        "DOPRI5 rtol=atol=1e-5, obs every 0.05 on even dims N(0,1), "
                        f"{args.resampler} resampling at ESS<0.5N",
            "particles": P,
            "resampled_steps": int(resampled.sum()),
            "log_weights": "f64",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_transition<L96<float,40,4>,DOPRI5>",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "avg_launch_ms": trans_ms / max(trans_n, 1),
            "launches": trans_n,
            "resample_chain_ms_per_step": res_ms / max(args.steps, 1),
            "reduce_ms_per_step": red_ms / max(args.steps, 1),
        },
    }
    pf.close()
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(args, y, mean0)
    print(json.dumps(line), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
