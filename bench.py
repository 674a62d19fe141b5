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

roofline: after the timed steps, one plain step and one dynamic-SGS update
step run with every launch of the library's kernel families bracketed by HIP
events on the stream it runs on (les_profile_step); each family's device
time per step is (cs_count - 1) plain + 1 update over cs_count.  The family
with the most device time per step is the roofline kernel.  The timed steps
then run again from the same initial fields with only that family's
launches bracketed (les_profile_run): `achieved` = its algorithmic bytes
per launch / its average launch time over those steps.  `step` adds the
whole step's fraction: the algorithmic bytes of the stage list below (frozen
here) over the timed ms per step.  `traffic` is the HBM bytes per launch of
that family from the committed rocprofv3 --pmc FETCH_SIZE/WRITE_SIZE runs
(profiles/traffic.json), null when none covers it.  cpu_baseline: the FP64
CPU restatement (oracle) on the host cores, three plain steps (median) and
one dynamic-SGS update of the same 256^3 workload, weighted 4:1 as in a
cs_count = 5 cycle.  adjoint: the discrete adjoint of the FP64 step at 128^3
(see adjoint_leg).
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

# Compulsory HBM traffic of one FP64 step with theta on, in units of N*s
# bytes (N = nxt*nyt*(nzt-1) points, s = 8): each stage reads its input 3-D
# fields once and writes its outputs once; a half-spectrum field counts 1 N,
# a 3/2-padded one 2.25 N.  Every cs_count steps the Lagrangian update adds
# LAGRANGE_NS (interpolag_Sdep 14 + lagrange_Sdep 18).
STAGES_NS = {
    "ddx/ddy of u, v, w": 12.0,
    "ddz of u, v, w": 6.0,
    "sgs_stag": 26.0,
    "divstress x3": 12.0,
    "convec (with the padded u, v, w kept for theta)": 18.75,
    "theta: filter_data + 3 gradients": 8.0,
    "scalar_RHS_calc": 12.75,
    "calc_beta + step_scalar": 5.0,
    "RHS assembly": 10.0,
    "Admas_Bashforth": 12.0,
    "press_stag_array I/O": 6.0,
    "dpdz + RHS -= grad p": 10.0,
    "project": 9.0,
    "filter_data of u, v, w": 6.0,
}
LAGRANGE_NS = 32.0


def step_bytes(n_points, cs_count, s=8.0):
    """Algorithmic bytes of one step averaged over a cs_count cycle."""
    return n_points * s * (sum(STAGES_NS.values()) + LAGRANGE_NS / cs_count)


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
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint leg")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="run only the adjoint leg's timed windows (for rocprofv3)")
    ap.add_argument("--kernels-only", action="store_true",
                    help="run only the kernel-family timing loops, FP64 (for rocprofv3 --pmc)")
    ap.add_argument("--pmc-run", action="store_true",
                    help="the workload of the counter summary (run under rocprofv3 --pmc)")
    ap.add_argument("--pmc-summary", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="per-family HBM traffic from two rocprofv3 --pmc runs of --pmc-run")
    ap.add_argument("--trace-summary", nargs=2, metavar=("TRACE_DIR", "BENCH_OUT"),
                    help="per-family kernel durations from a rocprofv3 --kernel-trace run of this bench "
                         "(--precision fp64 or mixed, --no-cpu --no-adjoint) against the live averages "
                         "that the same run printed (its output file)")
    ap.add_argument("--mixed-drift", type=int, default=0, metavar="N",
                    help="only: run FP64 and mixed N steps side by side and print their difference every N/10 steps")
    return ap.parse_args()


ADJ_STATE = ("U", "V", "W", "THETA", "RHSX", "RHSY", "RHSZ", "RHST")


def mixed_drift(workload, nsteps, legs=None):
    """The mixed-precision step against the FP64 step (the FP64 GPU path is
    pinned to the oracle) from the same seeded fields: relative L2 of u, v,
    w, theta and of the plane-mean u and theta profiles every nsteps/10
    steps.  Two legs:
      - sbl256: the 256^3 workload as configured, until its CFL stop.  Both
        precisions stop near step 85 (the synthetic case is more aggressive
        than the reference's shipped namelist: nondimensional dt 2e-4 against
        4e-5, dx/dz 6.3 against 3.5);
      - sbl128: 128^3, which runs the whole nsteps."""
    import numpy as np
    from overlappingbls_amd import config, lib
    from overlappingbls_amd.abi import LES_MIXED
    every = max(1, nsteps // 10)
    out = {"steps": nsteps, "what": "relative L2 of the mixed-precision state against the FP64 state"}
    for leg in (legs or (workload, "sbl128")):
        nl = config.workload(leg)
        p = nl.params()
        fields = config.initial_fields(nl, seed=2024)
        a, b = lib.create(p), lib.create(p)
        b.set_precision(LES_MIXED)
        for c in (a, b):
            config.load_state(c, fields)
        rows = []
        jt = 1
        t0 = time.perf_counter()
        while jt <= nsteps:
            n = min(every, nsteps - jt + 1)
            row = {}
            for tag, c in (("fp64", a), ("mixed", b)):
                try:
                    c.run(jt, n)
                except Exception as e:   # CFL > 0.6 (LES_E_CFL) ends the comparison
                    row["stopped"] = row.get("stopped", "") + f"{tag}: {e}; "
                row["cfl_" + tag] = c.cfl()
            if "stopped" in row:
                rows.append(row)
                break
            jt += n
            row["step"] = jt - 1
            for name in ("U", "V", "W", "THETA"):
                sl = config.physical_levels(name, p.nzt)
                x, y = b.get(name)[sl], a.get(name)[sl]
                row[name] = rel(x, y)
                if name in ("U", "THETA"):
                    row["mean_" + name] = rel(x.mean(axis=(1, 2)), y.mean(axis=(1, 2)))
            rows.append(row)
            print(leg, json.dumps(row), file=sys.stderr, flush=True)
        out[leg] = {"seconds": time.perf_counter() - t0, "rows": rows}
        a.close()
        b.close()
    return out


def adjoint_leg(workload="sbl128", nsteps=20, k_ckpt=5, obj_steps=5, ndir=8, eps=1e-6, timed_only=False):
    """Discrete adjoint at 128^3 (FP64, SGS frozen): a 20-step window,
    J = mean surface momentum flux ustar^2 over the last 5 steps, gradient
    with respect to the whole initial state.  The window is timed with the
    checkpoint spacing chosen from free device memory (k = 0: here the whole
    window is taped in the forward sweep, 11 GB) and with a checkpoint every
    5 steps (each earlier segment recomputed once), both against the plain
    forward window.  Also reports the tangent/adjoint dot-product test over
    the window and central differences of J (eps = 1e-6 relative to each
    field's rms) along 8 seeded Gaussian directions against <grad J, d>."""
    import numpy as np
    from overlappingbls_amd import config, lib
    nl = config.workload(workload)
    p = nl.params()
    nz, jt0 = p.nzt, 3
    g = lib.create(p)
    config.load_state(g, config.initial_fields(nl, seed=2025))
    g.run(1, jt0 - 1)
    g.set_sgs_frozen(True)
    x0 = {n: g.get(n) for n in config.STATE_FIELDS}
    x0["PSI_M_SFC"] = g.get("PSI_M_SFC")
    flags = g.sgs_init_flags()

    def reset(d=None, e=0.0):
        for n, a in x0.items():
            if d is not None and n in d:
                a = a.copy()
                if n == "PSI_M_SFC":
                    a = a + e * d[n]
                else:
                    a[1:nz] += e * d[n]
            g.set(n, a)
        g.set_sgs_init_flags(*flags)

    def direction(seed, scale):
        d = {n: scale[n] * config.gaussian(seed, 60 + i, (nz - 1, p.nyt, p.nxt)) for i, n in enumerate(ADJ_STATE)}
        d["PSI_M_SFC"] = scale["PSI_M_SFC"] * np.ones((p.nyt, p.nxt)) * float(config.gaussian(seed, 70, (1, 1, 1)).flat[0])
        return d

    def lin_load(d):
        g.lin_zero()
        for n, a in d.items():
            if n == "PSI_M_SFC":
                g.lin_set(n, a)
            else:
                g.lin_set(n, a, lev0=1)

    def lin_read():
        out = {n: g.lin_get(n, lev0=1, nlev=nz - 1) for n in ADJ_STATE}
        out["PSI_M_SFC"] = g.lin_get("PSI_M_SFC")
        return out

    def inner(a, b):
        return sum(float(np.sum(a[n] * b[n])) for n in a)

    def timed(fn):
        g.sync()
        t0 = time.perf_counter()
        r = fn()
        g.sync()
        return time.perf_counter() - t0, r

    unit = {n: 1.0 for n in ADJ_STATE}
    unit["PSI_M_SFC"] = 0.1
    # warm: allocates the linear buffers, tapes and checkpoints
    for k in (0, k_ckpt):
        reset(); lin_load(direction(1, unit)); g.adjoint_window(jt0, nsteps, k, obj_steps)
    reset(); t_fwd, _ = timed(lambda: g.run(jt0, nsteps))
    dx, y = direction(2, unit), direction(3, unit)
    reset(); lin_load(dx)
    t_tlm, dJ = timed(lambda: g.tlm_window(jt0, nsteps, obj_steps))
    tl = lin_read()
    reset(); lin_load(y)
    t_ck, _ = timed(lambda: g.adjoint_window(jt0, nsteps, k_ckpt, obj_steps))
    reset(); lin_load(y)
    t_adj, J = timed(lambda: g.adjoint_window(jt0, nsteps, 0, obj_steps))
    grad_y = lin_read()
    res = {
        "workload": f"{workload}: {p.nxt}x{p.nyt}x{nz - 1}, FP64, SGS frozen, window {nsteps} steps from jt={jt0}, "
                    f"J = mean ustar^2 over the last {obj_steps} steps",
        "ms_forward_window": 1e3 * t_fwd,
        "ms_tlm_window": 1e3 * t_tlm,
        "ms_forward_step": 1e3 * t_fwd / nsteps,
        # checkpoint spacing from free device memory (the whole window taped)
        "ms_adjoint_window": 1e3 * t_adj,
        "ratio_window": t_adj / t_fwd,
        # checkpoint every k_ckpt steps: forward with checkpoints, each
        # earlier segment recomputed once with tapes, the adjoint sweep
        "checkpointed": {"k_ckpt": k_ckpt, "ms_adjoint_window": 1e3 * t_ck, "ratio_window": t_ck / t_fwd,
                         # taped recomputation + adjoint sweep per step, against one forward step
                         "ratio_step": (t_ck - t_fwd) / t_fwd},
        "J": J,
    }
    reset()
    times = g.lin_stage_times(jt0)
    res["stage_ms"] = {k: [round(a, 4), round(b, 4)] for k, (a, b) in times.items()}
    res["stage_ms_total"] = [round(sum(a for a, _ in times.values()), 4), round(sum(b for _, b in times.values()), 4)]
    if timed_only:
        return res
    lhs = inner(tl, y) + dJ
    rhs = inner(dx, grad_y)
    res["dot_test"] = {"lhs": lhs, "rhs": rhs, "rel": abs(lhs - rhs) / max(abs(lhs), abs(rhs))}
    # gradient of J alone, and central differences along Gaussian directions
    reset(); g.lin_zero()
    g.adjoint_window(jt0, nsteps, 0, obj_steps)
    gJ = lin_read()
    scale = {n: float(np.sqrt(np.mean(x0[n][1:nz] ** 2))) for n in ADJ_STATE}
    scale["PSI_M_SFC"] = max(abs(float(x0["PSI_M_SFC"].mean())), 1e-3)

    def J_at(d, e):
        reset(d, e)
        s = 0.0
        for i in range(nsteps):
            g.step(jt0 + i)
            if i >= nsteps - obj_steps:
                s += float(np.mean(g.get("USTAR_SFC") ** 2))
        return s / obj_steps

    fd = []
    for s in range(ndir):
        d = direction(100 + s, scale)
        a = inner(d, gJ)
        f = (J_at(d, eps) - J_at(d, -eps)) / (2 * eps)
        fd.append({"adjoint": a, "fd": f, "rel": abs(a - f) / max(abs(a), abs(f))})
    res["fd_check"] = {"eps": eps, "directions": ndir, "max_rel": max(x["rel"] for x in fd), "cases": fd}
    return res


# Kernel families of les_profile_step by kernel name (the library's FamScope
# brackets), for the counter summary.  Line kernels carry their operation as
# the last template argument: 0 derivative, 1 Nyquist filter, 2 pad, 3 unpad.
_LINE_OPS = {"0": "deriv", "1": "filter", "2": "pad", "3": "unpad"}
_POINT_FAMILIES = [
    (("k_obukhov_sums", "k_level_means_seq"), "plane_sums"),
    (("k_rhs_ab2",), "rhs_ab2"),
    (("k_gradp_project", "k_project_bc"), "gradp_project"),
    (("k_strain_stress", "k_strain<", "k_stress<"), "sgs_stress"),
    (("k_scalar_flux", "k_scalar_accum", "k_ddz_w_accum"), "scalar_flux"),
    (("k_ddz_uv", "k_ddz_w<"), "ddz"),
    (("k_divstress",), "divstress"),
    (("k_max3_lag", "k_max3_part", "k_max3_final"), "cfl_max"),
    (("k_press_prep", "k_dpdz", "k_hz_part", "k_hz_fold"), "press_prep"),
    (("k_lag_prep", "k_interpolag", "k_lag_contract<", "k_lag_acc"), "lagrange_point"),
    (("k_theta_prep", "k_step_scalar", "k_calc_beta", "k_scalar_top"), "theta_point"),
    (("k_tdma",), "tdma"),
    (("kr_lag_contract",), "lag_contract_x"),
    (("kr_x_r2c", "k_x_r2c"), "x_r2c"),
    (("kr_x_c2r", "k_x_c2r"), "x_c2r"),
    (("kr_yspec", "k_yline_spec"), "yline.spec"),
    (("k_vort<",), "vort_pad_x"),
]
# the first kernel of every step: obukhov's plane sums, which les_run launches
# before the CFL pass, or the CFL pass where a step has no such sums
_STEP_START = (("k_obukhov_sums",), ("k_max3_lag", "k_max3_part"))
# kernels that follow another one inside the same bracket of les_profile_step,
# so that a family's launches count brackets, as the library's profiler does
_SECONDARY = ("k_project_bc", "k_max3_final", "k_scalar_top", "k_hz_fold")


def kernel_family(name: str) -> str:
    import re
    m = re.search(r"kr_([xy])line<[^>]*?, (\d+)>", name)
    if m:
        return f"{m.group(1)}line.{_LINE_OPS[m.group(2)]}"
    if "kr_fused_x" in name:
        if "VortIn" in name:
            return "vort_pad_x"
        m = re.search(r"kr_fused_x<[^,]+, \d+, \d+, \d+, (\d+),", name)
        return "scalar_unpad_x" if (m and m.group(1) == "1") else "cross_unpad_x"
    for keys, fam in _POINT_FAMILIES:
        if any(k in name for k in keys):
            return fam
    return "other"


def pmc_run(workload: str):
    """The workload the counter summary reads, under rocprofv3 --pmc: a
    device copy of one field (calibration), steps 1-3 (warm-up), then the
    plain step 4 and the dynamic-SGS update step 5 (cs_count = 5)."""
    from overlappingbls_amd import config, lib
    from overlappingbls_amd.abi import K_COPY
    nl = config.workload(workload)
    ctx = lib.create(nl.params())
    config.load_state(ctx, config.initial_fields(nl, seed=2024))
    ctx.time_kernel(K_COPY, 1)
    ctx.run(1, 5)
    ctx.sync()
    print(json.dumps({"pmc_run": workload, "steps": [4, 5], "cs_count": nl.cs_count}))


def _pmc_rows(d: str, counter: str):
    import csv
    import glob
    files = sorted(glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True))
    if not files:
        raise SystemExit(f"no counter_collection.csv under {d}")
    rows = []
    with open(files[0]) as f:
        for r in csv.DictReader(f):
            if r.get("Counter_Name") == counter:
                rows.append((int(r["Dispatch_Id"]), r["Kernel_Name"], float(r["Counter_Value"])))
    rows.sort()
    return rows


def pmc_summary(fetch_dir: str, write_dir: str, workload: str):
    """HBM traffic per family of the plain step 4 and the update step 5 of
    pmc_run, from the FETCH_SIZE and WRITE_SIZE runs (KiB per dispatch),
    scaled so that the calibration copy reads its known size; writes
    profiles/traffic.json and prints a markdown table."""
    from overlappingbls_amd import config
    nl = config.workload(workload)
    p = nl.params()
    field_bytes = 8.0 * p.nxt * p.nyt * (p.nzt + 1)
    out = {}
    for counter, d in (("FETCH_SIZE", fetch_dir), ("WRITE_SIZE", write_dir)):
        rows = _pmc_rows(d, counter)
        copies = [v for _, n, v in rows if "copyBuffer" in n and v > 0.5 * field_bytes / 1024.0]
        if not copies:
            raise SystemExit("no calibration copy in " + d)
        factor = field_bytes / (copies[0] * 1024.0)
        for marks in _STEP_START:
            starts = [i for i, (_, n, _) in enumerate(rows) if any(k in n for k in marks)]
            if len(starts) == 5:
                break
        else:
            raise SystemExit(f"expected 5 steps in {d}, found {len(starts)}")
        steps = {4: rows[starts[3]:starts[4]], 5: rows[starts[4]:]}
        out[counter] = {"factor": factor, "steps": steps}
    fams = {}
    w_plain, w_upd = (nl.cs_count - 1) / nl.cs_count, 1.0 / nl.cs_count
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        fac = out[counter]["factor"]
        for jt, w in ((4, w_plain), (5, w_upd)):
            for _, name, v in out[counter]["steps"][jt]:
                f = fams.setdefault(kernel_family(name), {"FETCH_SIZE": 0.0, "WRITE_SIZE": 0.0, "launches": 0.0})
                f[counter] += w * v * 1024.0 * fac
                if counter == "FETCH_SIZE" and not any(k in name for k in _SECONDARY):
                    f["launches"] += w
    data = {"_unit": "HBM bytes per launch of each kernel family (FETCH_SIZE + WRITE_SIZE, KiB per dispatch, "
                     "scaled by a device copy of one field), averaged over a cs_count cycle",
            "_calibration": {"copy_bytes_each_way": field_bytes, "fetch_factor": out["FETCH_SIZE"]["factor"],
                             "write_factor": out["WRITE_SIZE"]["factor"]},
            "_workload": f"{workload}, FP64, bench.py --pmc-run (plain step 4, update step 5)"}
    print(f"| family | launches/step | fetched MB/step | written MB/step | traffic MB/launch |")
    print("|---|---|---|---|---|")
    for fam, f in sorted(fams.items(), key=lambda kv: -(kv[1]["FETCH_SIZE"] + kv[1]["WRITE_SIZE"])):
        tot = f["FETCH_SIZE"] + f["WRITE_SIZE"]
        per = tot / f["launches"] if f["launches"] else 0.0
        if fam != "other":
            data[fam] = per
            data[fam + "_detail"] = {"fetch_bytes_per_step": f["FETCH_SIZE"], "write_bytes_per_step": f["WRITE_SIZE"],
                                     "launches_per_step": f["launches"]}
        print(f"| {fam} | {f['launches']:.1f} | {f['FETCH_SIZE'] / 1e6:.1f} | {f['WRITE_SIZE'] / 1e6:.1f} | "
              f"{per / 1e6:.1f} |")
    # the name mapping checked against the library's own launch counts per
    # family (les_profile_step, recorded by the bench run in profiles/)
    bench_json = os.path.join(HERE, "profiles", f"bench_{workload}.json")
    if os.path.exists(bench_json):
        with open(bench_json) as fh:
            lib_fams = json.load(fh).get("families", {})
        bad = {f: (round(fams.get(f, {}).get("launches", 0.0), 3), x["launches_per_step"])
               for f, x in lib_fams.items()
               if abs(fams.get(f, {}).get("launches", 0.0) - x["launches_per_step"]) > 1e-6}
        data["_launch_check"] = {"against": os.path.basename(bench_json), "mismatches": bad}
        print(f"launch counts per family against {os.path.basename(bench_json)}: "
              + ("all match" if not bad else f"mismatches {bad}"))
    with open(os.path.join(HERE, "profiles", "traffic.json"), "w") as fh:
        json.dump(data, fh, indent=1)


def trace_summary(trace_dir: str, bench_out: str, cs_count: int = 5):
    """Per-family device time of a traced bench run (rocprofv3
    --kernel-trace of `bench.py --precision fp64|mixed --no-cpu --no-adjoint`)
    beside the live HIP-event averages the same run printed.  The trace holds,
    in order: W warm-up steps, the K timed steps, the two profiled steps
    (plain, then dynamic-SGS update; les_profile_step, every family
    bracketed), the rerun of the W + K steps with only the roofline family
    bracketed (les_profile_run), then the kernel table.  Steps are cut at
    their first kernel (_STEP_START)."""
    import csv
    import glob
    files = sorted(glob.glob(os.path.join(trace_dir, "**", "*kernel_trace.csv"), recursive=True))
    if not files:
        raise SystemExit(f"no kernel_trace.csv under {trace_dir}")
    with open(files[0]) as f:
        ks = sorted((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), r["Kernel_Name"])
                    for r in csv.DictReader(f))
    with open(bench_out) as f:
        line = [x for x in f.read().splitlines() if x.startswith('{"metric"')][-1]
    run = json.loads(line)
    W, K = run["warmup"], run["steps"]
    nst = 2 * (W + K) + 2
    for marks in _STEP_START:
        starts = [i for i, k in enumerate(ks) if any(m in k[2] for m in marks)]
        if len(starts) == nst:
            break
    else:
        raise SystemExit(f"expected {nst} steps in the trace, found {len(starts)}")
    # the kernel table after the rerun starts with back-to-back TDMA launches
    end = next((i for i in range(starts[-1], len(ks) - 1)
                if "k_tdma" in ks[i][2] and "k_tdma" in ks[i + 1][2]), len(ks))

    def fam_time(lo, hi):
        out = {}
        for t0, t1, name in ks[lo:hi]:
            o = out.setdefault(kernel_family(name), [0.0, 0])
            o[0] += (t1 - t0) / 1e3
            if not any(x in name for x in _SECONDARY):  # timed with the kernel it follows
                o[1] += 1
        return out
    timed = fam_time(starts[W], starts[W + K])
    plain = fam_time(starts[W + K], starts[W + K + 1])
    update = fam_time(starts[W + K + 1], starts[W + K + 2])
    rerun = fam_time(starts[2 * W + K + 2], end)
    w = (cs_count - 1) / cs_count
    live = run.get("families", {})
    roof = run["roofline"]
    # wall time of each timed step, first kernel to the next step's first kernel
    walls = {"plain": [], "update": []}
    for j in range(W, W + K):
        jt = j + 1
        walls["update" if jt % cs_count == 0 else "plain"].append((ks[starts[j + 1]][0] - ks[starts[j]][0]) / 1e6)
    med = {k: sorted(v)[len(v) // 2] for k, v in walls.items() if v}
    print("timed steps under the trace, median wall ms: " + ", ".join(f"{k} {v:.3f}" for k, v in med.items()))
    print(f"run: {run['config']['precision']}, {run['ms_per_step']:.3f} ms per timed step; roofline family "
          f"{roof['kernel']}: live {roof['ms_per_launch'] * 1e3:.1f} us per launch over the rerun steps, "
          f"traced {rerun[roof['kernel']][0] / rerun[roof['kernel']][1]:.1f} us "
          f"({rerun[roof['kernel']][1]} launches)")
    print("| family | launches/step | traced us/launch, timed steps | traced us/launch, rerun steps | "
          "traced us/launch, profiled steps | live us/launch, profiled steps | live ms/step, profiled steps |")
    print("|---|---|---|---|---|---|---|")
    fams = sorted(live, key=lambda f: -live[f]["ms_per_step"])
    for fam in fams:
        p, u = plain.get(fam, [0.0, 0]), update.get(fam, [0.0, 0])
        n = w * p[1] + (1 - w) * u[1]
        prof_us = (w * p[0] + (1 - w) * u[0]) / n if n else 0.0
        t, r = timed.get(fam, [0.0, 0]), rerun.get(fam, [0.0, 0])
        lv = live[fam]
        print(f"| {fam} | {n:.1f} | {t[0] / t[1] if t[1] else 0:.1f} | {r[0] / r[1] if r[1] else 0:.1f} | "
              f"{prof_us:.1f} | {lv['ms_per_step'] / lv['launches_per_step'] * 1e3:.1f} | "
              f"{lv['ms_per_step']:.3f} |")


def traffic_from_profiles(family: str, workload: str = "sbl256"):
    """HBM bytes per launch of a kernel family measured with rocprofv3 --pmc
    (committed summary profiles/traffic.json, keyed by family); None for a
    workload the summary was not taken on."""
    path = os.path.join(HERE, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        data = json.load(f)
    if not str(data.get("_workload", "")).startswith(workload + ","):
        return None
    return data.get(family)


def workload_text(name, nl, p):
    if p.ocean_flag:
        return (f"{name}: {p.nxt}x{p.nyt}x{p.nzt - 1} ocean bottom boundary layer (geostrophic current "
                f"{nl.ug_dim:g} m/s, rough wall z0={p.zo:g} z_i at the computational top, stress-free "
                f"surface, no wind stress), one scalar, Lagrangian scale-dependent SGS "
                f"(cs_count={nl.cs_count}), dt={p.dt:g}")
    return (f"{name}: {p.nxt}x{p.nyt}x{p.nzt - 1} surface BL, stable (wt_s={nl.wt_s / nl.T_scale:+g} u*T), "
            f"Lagrangian scale-dependent SGS (cs_count={nl.cs_count}), rough log-law wall, stress-free top, "
            f"dt={p.dt:g}")


def family_profile(ctx, jt, cs_count):
    """Per-family device time, algorithmic bytes and launches per step from
    step jt on: the next plain step and the next dynamic-SGS update step
    profiled (les_profile_step; the steps between run normally), weighted
    (cs_count - 1):1."""
    while jt % cs_count == 0:
        ctx.run(jt, 1)
        jt += 1
    a = ctx.profile_step(jt)
    jt += 1
    while jt % cs_count != 0:
        ctx.run(jt, 1)
        jt += 1
    b = ctx.profile_step(jt)
    w = (cs_count - 1) / cs_count
    out = {}
    for fam in a:
        cnt = w * a[fam][2] + (1 - w) * b[fam][2]
        if cnt > 0:
            ms = w * a[fam][0] + (1 - w) * b[fam][0]
            nb = w * a[fam][1] + (1 - w) * b[fam][1]
            out[fam] = {"ms_per_step": ms, "bytes_per_step": nb, "launches_per_step": cnt,
                        "GBs": nb / (ms * 1e-3) / 1e9 if ms > 0 else None}
    return out


def cpu_baseline(workload: str):
    from oracle import oracle  # checker / baseline only
    from overlappingbls_amd import config
    import statistics
    nl = config.workload(workload)
    p = nl.params()
    o = oracle.create(p)
    config.load_state(o, config.initial_fields(nl))
    plain = []
    for jt in range(1, nl.cs_count):         # plain steps 1 .. cs_count-1 (at most 3 timed)
        t0 = time.perf_counter(); o.step(jt); dt = time.perf_counter() - t0
        if jt >= nl.cs_count - 3:
            plain.append(dt)
    t0 = time.perf_counter(); o.step(nl.cs_count); t_upd = time.perf_counter() - t0
    t_plain = statistics.median(plain)
    per_step = ((nl.cs_count - 1) * t_plain + t_upd) / nl.cs_count
    n = p.nxt * p.nyt * (p.nzt - 1)
    return {
        "value": n / per_step / 1e6,
        "unit": "Mgridpts/s",
        "cores": oracle.threads(),
        "kind": "port",
        "sample": f"{workload}: median of {len(plain)} plain steps ({t_plain:.2f} s) and the dynamic-SGS update "
                  f"step {nl.cs_count} ({t_upd:.2f} s), weighted {nl.cs_count - 1}:1; the oracle is the C "
                  f"restatement built with gcc -O2, no FMA contraction, its own radix-2/3/4 FFT",
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

    if args.pmc_run:
        pmc_run(args.workload)
        return
    if args.pmc_summary:
        pmc_summary(args.pmc_summary[0], args.pmc_summary[1], args.workload)
        return
    if args.trace_summary:
        trace_summary(args.trace_summary[0], args.trace_summary[1], config.workload(args.workload).cs_count)
        return
    if args.adjoint_only:
        print(json.dumps(adjoint_leg(timed_only=True)))
        return
    if args.mixed_drift:
        print(json.dumps(mixed_drift(args.workload, args.mixed_drift)))
        return
    if args.kernels_only:
        nl = config.workload(args.workload)
        ctx = lib.create(nl.params())
        config.load_state(ctx, config.initial_fields(nl, seed=2024))
        fams = {"tdma": K_TDMA, "xline_deriv": K_XDERIV, "yline_deriv": K_YDERIV, "pad32": K_PAD, "copy": K_COPY}
        print(json.dumps(kernel_table(ctx, fams)))
        return
    nl = config.workload(args.workload)
    p = nl.params()
    n = p.nxt * p.nyt * (p.nzt - 1)
    fields = config.initial_fields(nl, seed=2024)
    fams = {"tdma": K_TDMA, "xline_deriv": K_XDERIV, "yline_deriv": K_YDERIV, "pad32": K_PAD, "copy": K_COPY}
    runs, kts, states, profs, roof_runs = {}, {}, {}, {}, {}
    for prec in (("fp64", "mixed") if args.precision == "both" else (args.precision,)):
        # One live context at a time: a context's main and auxiliary streams
        # overlap only when they get hardware queues of their own, and a
        # second live context takes queues too (the mixed step measured
        # 7.38 ms with the FP64 context still open, 6.57 ms alone).
        ctx = lib.create(p)
        if prec == "mixed":
            ctx.set_precision(LES_MIXED)
        config.load_state(ctx, fields)
        t = time_steps(ctx, args.warmup, args.steps)
        runs[prec] = {"ms_per_step": 1e3 * t / args.steps, "value": n / (t / args.steps) / 1e6}
        if not args.profile_only:
            states[prec] = {name: ctx.get(name)[config.physical_levels(name, p.nzt)]
                            for name in ("U", "V", "W", "THETA")}
            # a plain step and an update step with the kernel families timed
            profs[prec] = family_profile(ctx, args.warmup + args.steps + 1, nl.cs_count)
            if prec == ("fp64" if args.precision in ("both", "fp64") else "mixed"):
                # the timed steps again, in a fresh context from the same
                # initial fields, with only the family of most device time
                # bracketed by events
                ctx.close()
                dom = max(profs[prec], key=lambda f: profs[prec][f]["ms_per_step"])
                ctx = lib.create(p)
                if prec == "mixed":
                    ctx.set_precision(LES_MIXED)
                config.load_state(ctx, fields)
                ctx.run(1, args.warmup)
                r = ctx.profile_run(args.warmup + 1, args.steps, dom)[dom]
                roof_runs[prec] = {"family": dom, "ms": r[0], "bytes": r[1], "launches": r[2]}
            kts[prec] = kernel_table(ctx, fams)
        ctx.close()
    if args.profile_only:
        print(json.dumps({k: v["ms_per_step"] for k, v in runs.items()}))
        return
    head = "fp64" if "fp64" in runs else "mixed"
    kt = kts[head]
    value, ms_step = runs[head]["value"], runs[head]["ms_per_step"]
    prof = profs[head]
    rr = roof_runs[head]
    dom = rr["family"]                      # most device time per step (family_profile)
    achieved = rr["bytes"] / (rr["ms"] * 1e-3) / 1e9
    sb = step_bytes(n, nl.cs_count)
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic_from_profiles(dom, args.workload),
                "traffic_source": "rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs, profiles/traffic.json",
                "kernel": dom,
                "measured_over": f"steps {args.warmup + 1}..{args.warmup + args.steps} rerun from the same "
                                 f"initial fields with only this family bracketed by HIP events",
                "bytes_per_launch": rr["bytes"] / rr["launches"],
                "ms_per_launch": rr["ms"] / rr["launches"],
                "launches": rr["launches"],
                "share_of_step": rr["ms"] / args.steps / ms_step,
                "step": {"bytes": sb, "GBs": sb / (ms_step * 1e-3) / 1e9,
                         "frac": sb / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
                         "stages_Ns": sum(STAGES_NS.values()), "update_Ns_per_cs_count": LAGRANGE_NS}}
    mixed = None
    if "mixed" in runs:
        mixed = dict(runs["mixed"])
        mixed["kernels"] = kts["mixed"]
        mixed["families"] = {f: round(x["ms_per_step"], 5) for f, x in profs["mixed"].items()}
        if "fp64" in runs:
            mixed["speedup_vs_fp64"] = runs["fp64"]["ms_per_step"] / runs["mixed"]["ms_per_step"]
            err = {"after_steps": args.warmup + args.steps}
            for name in ("U", "V", "W", "THETA"):
                a, b = states["mixed"][name], states["fp64"][name]
                err[name] = rel(a, b)
                if name in ("U", "THETA"):
                    err["mean_" + name] = rel(a.mean(axis=(1, 2)), b.mean(axis=(1, 2)))
            mixed["err_vs_fp64"] = err
        states.clear()
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
        "config": {"workload": workload_text(args.workload, nl, p), "precision": head, "parallelism": "1 GPU"},
        "roofline": roofline,
        "families": {f: {k: (round(v, 5) if isinstance(v, float) else v) for k, v in x.items()}
                     for f, x in sorted(prof.items(), key=lambda kv: -kv[1]["ms_per_step"])},
        "kernels": kt,
    }
    if mixed is not None:
        out["mixed"] = mixed
    if not args.no_adjoint:
        out["adjoint"] = adjoint_leg()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.workload)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
