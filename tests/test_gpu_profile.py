"""The per-family device-time profiler (les_profile_step, les_profile_run)
that bench.py's roofline line reads: it times the step it runs without
changing it, counts each family's launches, and with a family filter
brackets only that family's launches."""
import numpy as np
import pytest

from overlappingbls_amd import config
from overlappingbls_amd.abi import LesError

pytestmark = pytest.mark.gpu

CASE = "mid"          # 64 x 64 x 8, cs_count = 2: steps 2, 4, ... run the dynamic update
FIELDS = ("U", "V", "W", "THETA", "CS_OPT2", "F_LM", "RHSX")


def fresh(gpu_lib):
    nl = config.workload(CASE)
    g = gpu_lib.create(nl.params())
    config.load_state(g, config.initial_fields(nl, seed=31))
    return g


def test_profiled_steps_equal_plain_steps(gpu_lib):
    """Profiling adds events only: the state after profiled steps equals the
    state after the same steps run by les_run, bit for bit."""
    a, b, c = fresh(gpu_lib), fresh(gpu_lib), fresh(gpu_lib)
    a.run(1, 4)
    b.profile_step(1)
    b.profile_step(2)
    b.profile_run(3, 2)
    c.profile_run(1, 4, "yline.pad")
    for n in FIELDS:
        ra = a.get(n)
        assert np.array_equal(ra, b.get(n), equal_nan=True), n
        assert np.array_equal(ra, c.get(n), equal_nan=True), n


def test_family_counts_and_filter(gpu_lib):
    g = fresh(gpu_lib)
    g.run(1, 2)
    plain = g.profile_step(3)             # no dynamic update at jt = 3
    for fam, (ms, nb, cnt) in plain.items():
        assert cnt >= 0 and ms >= 0.0 and nb >= 0.0, fam
        if cnt:
            assert ms > 0.0 and nb > 0.0, fam
    # one CFL pass, one TDMA solve, one RHS assembly and one projection per step
    for fam in ("cfl_max", "tdma", "rhs_ab2", "gradp_project"):
        assert plain[fam][2] == 1, (fam, plain[fam])
    assert plain["lag_contract_x"][2] == 0
    update = g.profile_step(4)            # jt = 4: dynamic update
    assert update["lag_contract_x"][2] == 2   # the 2 delta and 4 delta contractions
    # a filtered run over two plain steps (5 and 7 are odd) counts only its family
    g.run(5, 1)
    g.run(6, 1)
    one = g.profile_run(7, 1, "yline.pad")
    assert one["yline.pad"][2] == plain["yline.pad"][2] > 0
    assert one["yline.pad"][0] > 0.0
    assert all(cnt == 0 for fam, (_, _, cnt) in one.items() if fam != "yline.pad")
    # bytes per launch are fixed by the shapes
    assert one["yline.pad"][1] / one["yline.pad"][2] == plain["yline.pad"][1] / plain["yline.pad"][2]


def test_bad_family_is_refused(gpu_lib):
    g = fresh(gpu_lib)
    n = g._f["nfamilies"]()
    import ctypes as C
    for fam in (-2, n):
        ms, nb = np.zeros(n), np.zeros(n)
        cnt = (C.c_int * n)()
        with pytest.raises(LesError):
            g._call("profile_run", g._h, 1, 1, fam, ms.ctypes.data_as(C.POINTER(C.c_double)),
                    nb.ctypes.data_as(C.POINTER(C.c_double)), cnt)
    with pytest.raises(KeyError):
        g.profile_run(1, 1, "no such family")
