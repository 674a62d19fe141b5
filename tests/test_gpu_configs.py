"""Parity at the benchmark configurations themselves:
  - sbl64 (C0): 64 x 64 x 64, the CPU-runnable FP64 case, 100 AB2 steps (20
    dynamic-SGS updates), free-running on both sides;
  - sbl128 (C1): 128^3 with the passive scalar, surface heat flux and
    geostrophic/Coriolis forcing, 50 steps (ten updates), free-running;
  - sbl256 (C2): one FP64 step re-synced from the oracle at the FP64 bar,
    and the mixed-precision step for 25 steps against the FP64 GPU run
    (mean u and theta profiles within 1e-5, per-field errors printed);
  - sbl128 adjoint (C3): a 20-step window with a checkpoint every 5 steps,
    J over the last 5 (mean ustar^2 under the prescribed surface flux; the
    mean surface heat flux under a prescribed surface temperature):
    dot-product test and central differences.
Free-running bar as the free-running parity test: max(1e-10, NOISE times the
spread of the oracle's rounding variants over the same window)."""
import numpy as np
import pytest

from overlappingbls_amd import config
from overlappingbls_amd.abi import LES_MIXED

from test_gpu_parity import NOISE, SYNC_FIELDS, TENDENCIES, TOL, noise_ctx, pair, rel, sensitivity, sync

pytestmark = pytest.mark.gpu

FIELDS = ("U", "V", "W", "THETA", "CS_OPT2", "F_LM", "F_MM", "F_QN", "F_NN")


# 50 steps at 128^3 take the oracle about a minute per context on 16 cores,
# so that case measures the rounding spread with the FMA variant alone
@pytest.mark.timeout(900)
@pytest.mark.parametrize("case,nsteps,variants", [("sbl64", 100, ("fma", "r2")), ("sbl128", 50, ("fma",))])
def test_config_free_running(case, nsteps, variants, oracle_mod, gpu_lib):
    nl, o, g = pair(case, oracle_mod, gpu_lib)
    p = nl.params()
    of = noise_ctx(oracle_mod, p, variants)
    of.load(config.initial_fields(nl, seed=1234))
    o.run(1, nsteps)
    g.run(1, nsteps)
    for c in of.ctxs:
        c.run(1, nsteps)
    bad, seen = {}, {}
    for n in FIELDS:
        sl = config.physical_levels(n, p.nzt)
        ref = o.get(n)[sl]
        r = rel(g.get(n)[sl], ref)
        bound = max(1e-10, NOISE * of.spread(lambda c: c.get(n)[sl], ref))
        seen[n] = (float("%.2g" % r), float("%.2g" % bound))
        if not r < bound:
            bad[n] = (r, bound)
    print(case, nsteps, seen)
    assert not bad, bad


@pytest.mark.timeout(900)
def test_sbl256_fp64_step_resynced(oracle_mod, gpu_lib):
    """C2 at its full size in FP64: step 5 (a dynamic-SGS update with the F
    initialisation) from the oracle's state after four steps, every state
    field at the FP64 bar (tendencies at the conditioning bound, fields
    with cancelling plane means at NOISE times the oracle's rounding
    spread), as the resynced parity test."""
    nl, o, g = pair("sbl256", oracle_mod, gpu_lib)
    p = nl.params()
    o.run(1, 4)
    jt = 5
    state = {n: o.get(n) for n in SYNC_FIELDS}
    state["_flags"] = o.sgs_init_flags()
    sync(o, g)
    of = noise_ctx(oracle_mod, p)
    of.sync_from(o)
    o.step(jt); g.step(jt); of.step(jt)
    kap = sensitivity(oracle_mod, p, state, jt)
    bad, seen = {}, {}
    for n in config.STATE_FIELDS:
        sl = config.physical_levels(n, p.nzt)
        ref = o.get(n)[sl]
        r = rel(g.get(n)[sl], ref)
        bound = max(TOL, NOISE * of.spread(lambda c: c.get(n)[sl], ref))
        if n in TENDENCIES:
            bound = max(bound, 100 * 2.2e-16 * kap[n])
        seen[n] = (float("%.2g" % r), float("%.2g" % bound))
        if not r < bound:
            bad[n] = (r, bound)
    print("sbl256 step", jt, seen)
    assert not bad, bad


def test_sbl256_mixed_against_fp64(gpu_lib):
    """C2 mixed precision: 25 steps (five dynamic-SGS updates) of LES_MIXED
    against the FP64 GPU step (which the resynced test pins to the oracle)
    from the same seeded fields: plane-mean u and theta profiles within
    1e-5 relative, each field's relative L2 printed."""
    nl = config.workload("sbl256")
    p = nl.params()
    fields = config.initial_fields(nl, seed=2024)
    out = {}
    for prec in ("fp64", "mixed"):
        g = gpu_lib.create(p)
        if prec == "mixed":
            g.set_precision(LES_MIXED)
        config.load_state(g, fields)
        g.run(1, 25)
        out[prec] = {n: g.get(n)[config.physical_levels(n, p.nzt)] for n in ("U", "V", "W", "THETA")}
        g.close()
    errs = {n: rel(out["mixed"][n], out["fp64"][n]) for n in out["fp64"]}
    for n in ("U", "THETA"):
        errs["mean " + n] = rel(out["mixed"][n].mean(axis=(1, 2)), out["fp64"][n].mean(axis=(1, 2)))
    print("sbl256 mixed vs fp64 after 25 steps", {k: float("%.3g" % v) for k, v in errs.items()})
    assert errs["mean U"] < 1e-5 and errs["mean THETA"] < 1e-5, errs
    assert all(v < 1e-3 for v in errs.values()), errs


@pytest.mark.timeout(900)
@pytest.mark.parametrize("case,objective", [("sbl128", "ustar2"), ("sbl128_hf", "t_flux")])
def test_sbl128_adjoint_window(case, objective, gpu_lib):
    """C3: 128^3 FP64, SGS frozen, a 20-step window from jt = 3 with a
    checkpoint every 5 steps (each earlier segment recomputed), J over the
    last 5 steps: mean ustar^2 (sbl128, prescribed surface flux) or the mean
    surface heat flux sgs_t3(:,:,1) (sbl128_hf, prescribed surface
    temperature, so the flux depends on theta(1), ustar and psi_h).
    Dot-product test of the window's tangent and adjoint <= 1e-10; <grad J,
    d> against central differences of J (eps = 1e-6 relative to each
    field's rms) along 2 seeded directions <= 1e-6."""
    import test_gpu_adjoint as ta
    from overlappingbls_amd.abi import LES_OBJ_T_FLUX, LES_OBJ_USTAR2
    kind = LES_OBJ_T_FLUX if objective == "t_flux" else LES_OBJ_USTAR2
    nsteps, k, obj, jt0 = 20, 5, 5, ta.JT0

    def spun():
        nl = config.workload(case)
        g = gpu_lib.create(nl.params())
        config.load_state(g, config.initial_fields(nl, seed=2025))
        g.run(1, jt0 - 1)
        g.set_sgs_frozen(True)
        g.set_objective(kind)
        return g

    g = spun()
    x0 = {n: g.get(n) for n in config.STATE_FIELDS}
    x0["PSI_M_SFC"] = g.get("PSI_M_SFC")
    flags = g.sgs_init_flags()

    def reset(d=None, eps=0.0):
        for n, a in x0.items():
            if d is not None and n in d:
                a = a.copy()
                if n == "PSI_M_SFC":
                    a = a + eps * d[n]
                elif n == "SGS_T3":
                    a[1] += eps * d[n]
                else:
                    a[1:g.nz] += eps * d[n][1:g.nz]
            g.set(n, a)
        g.set_sgs_init_flags(*flags)

    dx, y = ta.direction(g, 7), ta.direction(g, 8)
    reset(); ta.lin_load(g, dx)
    dJ = g.tlm_window(jt0, nsteps, obj_steps=obj)
    tl = ta.lin_read(g)
    reset(); ta.lin_load(g, y)
    g.adjoint_window(jt0, nsteps, k, obj_steps=obj)
    grad_y = ta.lin_read(g)
    lhs = ta.inner(tl, ta.cut(y, g.nz)) + dJ
    rhs = ta.inner(ta.cut(dx, g.nz), grad_y)
    dot = ta.rel_diff(lhs, rhs)
    reset(); g.lin_zero()
    Jw = g.adjoint_window(jt0, nsteps, k, obj_steps=obj)
    gJ = ta.lin_read(g)
    reset()
    sc = ta.state_scale(g)

    def J_at(d, eps):
        reset(d, eps)
        s = 0.0
        for i in range(nsteps):
            g.step(jt0 + i)
            if i >= nsteps - obj:
                s += ta.objective_value(g, kind)
        return s / obj

    assert ta.rel_diff(Jw, J_at(None, 0.0)) < 1e-12
    errs = []
    for s in range(2):
        d = ta.direction(g, 30 + s, sc)
        a = ta.inner(ta.cut(d, g.nz), gJ)
        f = (J_at(d, 1e-6) - J_at(d, -1e-6)) / 2e-6
        errs.append(ta.rel_diff(a, f))
    print(case, objective, "adjoint window: J", Jw, "dot", dot, "fd", errs)
    assert dot < ta.DOT_TOL
    assert all(e < ta.FD_TOL for e in errs), errs
