"""Benchmark: Mgridpts/s per AB2 step of the boundary-layer LES timestep on one
MI355X, 256^3 surface boundary layer (stable, cooling flux -0.005, Lagrangian
scale-dependent SGS, rough log-law wall, stress-free top), synthetic seeded
fields resident in HBM.

Prints one JSON line.  `value` is the whole-job throughput
nxt*nyt*(nzt-1) / t_step / 1e6 of the FP64 step (flosol, 103_flosol.f90:4-290,
minus I/O and the ke/dudt diagnostics), timed over K consecutive steps that
include the dynamic-SGS update every cs_count = 5 steps.  `mixed` reports the
same K steps in the mixed-precision mode (FP32 FFTs, nonlinear terms, SGS and
scalar RHS; FP64 state, TDMA, mean mode and statistics), its speed-up over the
FP64 step, and its error against the FP64 run from the same initial fields
after the W + K steps (relative L2 of u, v, w, theta and of the plane-mean u
and theta profiles).

roofline: the dominant kernel family (by device time in the step) timed live
with HIP events on the library's stream; `achieved` = algorithmic bytes per
launch / launch time.  cpu_baseline: the FP64 CPU restatement (oracle) on the
host cores, two steps of the same 256^3 workload (one plain, one dynamic-SGS
update) weighted 4:1 as in a cs_count = 5 cycle.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E datasheet


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="sbl256")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--profile-only", action="store_true", help="run the timed steps only (for rocprofv3)")
    ap.add_argument("--precision", choices=("both", "fp64", "mixed"), default="both",
                    help="which step(s) to time (profile runs usually pick one)")
    return ap.parse_args()


def traffic_from_profiles(kernel_name: str):
    """HBM bytes per launch measured with rocprofv3 --pmc (committed summary)."""
    path = os.path.join(HERE, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        data = json.load(f)
    return data.get(kernel_name)


def cpu_baseline(workload: str):
    from oracle import oracle  # checker / baseline only
    from overlappingbls_amd import config
    nl = config.workload(workload)
    p = nl.params()
    o = oracle.create(p)
    config.load_state(o, config.initial_fields(nl))
    jt_plain, jt_upd = nl.cs_count - 1, nl.cs_count
    t0 = time.perf_counter(); o.step(jt_plain); t_plain = time.perf_counter() - t0
    t0 = time.perf_counter(); o.step(jt_upd); t_upd = time.perf_counter() - t0
    per_step = ((nl.cs_count - 1) * t_plain + t_upd) / nl.cs_count
    n = p.nxt * p.nyt * (p.nzt - 1)
    return {
        "value": n / per_step / 1e6,
        "unit": "Mgridpts/s",
        "cores": oracle.threads(),
        "kind": "port",
        "sample": f"{workload}: steps {jt_plain} (plain, {t_plain:.2f} s) and {jt_upd} (dynamic-SGS update, "
                  f"{t_upd:.2f} s), weighted {nl.cs_count - 1}:1",
    }


def rel(a, b):
    import numpy as np
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def time_steps(ctx, warmup, steps):
    """W untimed steps, then K timed steps (les_run enqueues them all and
    synchronises once at the end); returns seconds for the K steps."""
    jt = 1
    if warmup:
        ctx.run(jt, warmup)
        jt += warmup
    ctx.sync()
    t0 = time.perf_counter()
    ctx.run(jt, steps)
    ctx.sync()
    return time.perf_counter() - t0


def kernel_table(ctx, fams):
    kt = {}
    for name, k in fams.items():
        ms, nbytes = ctx.time_kernel(k, reps=10)
        kt[name] = {"ms": ms, "bytes": nbytes, "GBs": nbytes / (ms * 1e-3) / 1e9}
    return kt


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU z-slab "
              "decomposition is a design note, not part of this build.", file=sys.stderr)
        sys.exit(2)
    from overlappingbls_amd import config, lib
    from overlappingbls_amd.abi import K_TDMA, K_XDERIV, K_YDERIV, K_PAD, K_COPY, LES_MIXED

    nl = config.workload(args.workload)
    p = nl.params()
    n = p.nxt * p.nyt * (p.nzt - 1)
    fields = config.initial_fields(nl, seed=2024)
    runs = {}
    ctxs = {}
    for prec in (("fp64", "mixed") if args.precision == "both" else (args.precision,)):
        ctx = lib.create(p)
        if prec == "mixed":
            ctx.set_precision(LES_MIXED)
        config.load_state(ctx, fields)
        t = time_steps(ctx, args.warmup, args.steps)
        runs[prec] = {"ms_per_step": 1e3 * t / args.steps, "value": n / (t / args.steps) / 1e6}
        ctxs[prec] = ctx
    if args.profile_only:
        print(json.dumps({k: v["ms_per_step"] for k, v in runs.items()}))
        return
    fams = {"tdma": K_TDMA, "xline_deriv": K_XDERIV, "yline_deriv": K_YDERIV, "pad32": K_PAD, "copy": K_COPY}
    head = "fp64" if "fp64" in runs else "mixed"
    kt = kernel_table(ctxs[head], fams)
    value, ms_step = runs[head]["value"], runs[head]["ms_per_step"]
    dom = os.environ.get("LES_ROOFLINE_KERNEL", "yline_deriv")
    d = kt[dom]
    traffic = traffic_from_profiles(dom)
    roofline = {"bound": "hbm", "achieved": d["GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": d["GBs"] / HBM_PEAK_GBS, "traffic": traffic, "kernel": dom,
                "bytes_per_launch": d["bytes"], "ms_per_launch": d["ms"]}
    mixed = None
    if "mixed" in runs:
        mixed = dict(runs["mixed"])
        mixed["kernels"] = kernel_table(ctxs["mixed"], fams)
        if "fp64" in runs:
            mixed["speedup_vs_fp64"] = runs["fp64"]["ms_per_step"] / runs["mixed"]["ms_per_step"]
            err = {"after_steps": args.warmup + args.steps}
            nz = p.nzt
            for name in ("U", "V", "W", "THETA"):
                sl = config.physical_levels(name, nz)
                a = ctxs["mixed"].get(name)[sl]
                b = ctxs["fp64"].get(name)[sl]
                err[name] = rel(a, b)
                if name in ("U", "THETA"):
                    err["mean_" + name] = rel(a.mean(axis=(1, 2)), b.mean(axis=(1, 2)))
            mixed["err_vs_fp64"] = err
    out = {
        "metric": "Mgridpts/s per RK step, FP64 vs mixed @256^3 (err vs FP64); adjoint/forward ratio",
        "value": value,
        "unit": "Mgridpts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if head == "fp64" else "f32+f64",
        "data": "synthetic seeded fields (counter-based RNG), log-law profile + 10% noise below z=0.5",
        "config": {"workload": f"{args.workload}: {p.nxt}x{p.nyt}x{p.nzt - 1} surface BL, stable "
                               f"(wt_s={nl.wt_s / nl.T_scale:+g} u*T), Lagrangian scale-dependent SGS "
                               f"(cs_count={nl.cs_count}), rough log-law wall, stress-free top, "
                               f"dt={p.dt:g}", "precision": head, "parallelism": "1 GPU"},
        "roofline": roofline,
        "kernels": kt,
    }
    if mixed is not None:
        out["mixed"] = mixed
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.workload)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
