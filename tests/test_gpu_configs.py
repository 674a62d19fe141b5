"""Parity at the benchmark configurations themselves, free-running from the
same seeded fields on both sides:
  - sbl64: 64 x 64 x 64, the CPU-runnable FP64 case, 100 AB2 steps (20
    dynamic-SGS updates);
  - sbl128: 128^3 with the passive scalar, surface heat flux and
    geostrophic/Coriolis forcing, 2 cs_count + 1 = 11 steps (two updates);
    longer windows at 128^3 take the oracle minutes per variant.
Bar as the free-running parity test: max(1e-10, NOISE times the spread of the
oracle's rounding variants over the same window)."""
import pytest

from overlappingbls_amd import config

from test_gpu_parity import NOISE, noise_ctx, pair, rel

pytestmark = pytest.mark.gpu

FIELDS = ("U", "V", "W", "THETA", "CS_OPT2", "F_LM", "F_MM", "F_QN", "F_NN")


@pytest.mark.parametrize("case,nsteps", [("sbl64", 100), ("sbl128", 11)])
def test_config_free_running(case, nsteps, oracle_mod, gpu_lib):
    nl, o, g = pair(case, oracle_mod, gpu_lib)
    p = nl.params()
    of = noise_ctx(oracle_mod, p)
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
