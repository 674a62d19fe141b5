"""Mixed precision (LES_MIXED) against the FP64 oracle.

The mixed mode keeps the prognostic state (u, v, w, theta, RHS history), the
Lagrangian SGS state (Cs_opt2, F_*, u_lag), the wall model, the TDMA and its
mean mode, and the surface statistics in FP64; the FFTs, nonlinear terms, SGS
stresses and scalar RHS run on FP32 work fields.  There is no reference
counterpart: the bar is the error against the FP64 oracle, per field.

Bounds (relative L2, physical levels):
  - one FP32 spectral pass on O(1) data:           MIX_FFT
  - work fields produced routine by routine:       MIX_WORK
  - prognostic state after a window of steps:      MIX_STATE
  - horizontal-mean profiles of u and theta:       MIX_MEAN
"""
import numpy as np
import pytest

from overlappingbls_amd import config
from overlappingbls_amd.abi import LES_FP64, LES_MIXED

from test_gpu_parity import TOL, compare_state, pair, rand_field, rel, sync, PROGNOSTIC

pytestmark = pytest.mark.gpu

CASES = ["tiny_ragged", "small", "mid", "mid_rect", "wide", "wide_rect", "deep16"]
# the ocean branch, lbc = 0 and damping methods 2/3 in the mixed step
STEP_CASES = CASES + ["ocean_sbl", "ocean_deep", "land_lbc0", "land_damp2"]

MIX_FFT = 1e-6     # FP32 unit roundoff 6e-8 times the log2 n growth of a transform pair
MIX_WORK = 2e-5    # products and differences of FP32 fields (Cs_opt2 is a ratio of contractions)
MIX_STATE = 1e-6   # u, v, w, theta: FP64 state advanced by FP32 tendencies
MIX_MEAN = 1e-7    # plane-mean profiles
MIX_CS = 1e-4      # Cs_opt2 after two updates: a clipped ratio of FP32 contractions of
                   # cancelling filtered products (6e-5 at 512 x 128)


def mean_profile(a):
    return a.mean(axis=(1, 2))


@pytest.mark.parametrize("case", CASES)
def test_mixed_line_transforms(case, oracle_mod, gpu_lib):
    nl, o, g = pair(case, oracle_mod, gpu_lib)
    g.set_precision(LES_MIXED)
    a = rand_field(o, 11)
    for ctx in (o, g):
        ctx.set("WORK0", a)
        ctx.ddx("WORK1", "WORK0")
        ctx.ddy("WORK2", "WORK0")
    r = {n: rel(g.get(n), o.get(n)) for n in ("WORK1", "WORK2")}
    for ctx in (o, g):
        ctx.filter_data("WORK0")
    r["filter"] = rel(g.get("WORK0"), o.get("WORK0"))
    for ctx in (o, g):
        ctx.set("WORK0", a)
        ctx.test_filter("WORK0", 2, 2)
    r["test_filter"] = rel(g.get("WORK0", 2, 1), o.get("WORK0", 2, 1))
    # derivative of an FP64 state field into an FP32 work field
    for ctx in (o, g):
        ctx.ddx("DUDX", "U")
    nz = nl.params().nzt
    r["ddx(U)"] = rel(g.get("DUDX")[1:nz], o.get("DUDX")[1:nz])
    print(case, r)
    assert all(v < MIX_FFT for v in r.values()), r


def test_mixed_field_roundtrip(gpu_lib, oracle_mod):
    nl = config.workload("tiny")
    g = gpu_lib.create(nl.params())
    g.set_precision(LES_MIXED)
    a = config.uniform(3, 1, (g.nlev, g.ny, g.nx))
    g.set("WORK0", a)
    assert np.array_equal(g.get("WORK0"), a.astype(np.float32).astype(np.float64))
    g.set("U", a)                                    # state stays FP64
    assert np.array_equal(g.get("U"), a)


@pytest.mark.parametrize("case", CASES)
def test_mixed_routines(case, oracle_mod, gpu_lib):
    """flosol's routines one by one in the mixed mode, from identical inputs;
    each routine's outputs against the oracle's FP64 outputs."""
    nl, o, g = pair(case, oracle_mod, gpu_lib)
    g.set_precision(LES_MIXED)
    P = nl.params()
    nz = P.nzt
    jt = P.cs_count
    for k, n in enumerate(("U", "V", "W", "THETA")):
        a = o.get(n)
        a[1:nz] += (0.05 if n != "THETA" else 0.002) * config.uniform(99, k, a[1:nz].shape)
        for ctx in (o, g):
            ctx.set(n, a)
            ctx.filter_data(n)
    errs, over = {}, {}

    def check(names, sl, bound):
        for n in names:
            errs[n] = rel(g.get(n)[sl], o.get(n)[sl])
            if not errs[n] < bound:
                over[n] = (errs[n], bound)

    def sync_inputs(names):
        # each routine starts from the oracle's inputs, so the bound measures
        # that routine's own FP32 error
        for n in names:
            g.set(n, o.get(n))

    for ctx in (o, g):
        ctx.obukhov(jt)
        ctx.ddx("DUDX", "U"); ctx.ddy("DUDY", "U")
        ctx.ddx("DVDX", "V"); ctx.ddy("DVDY", "V")
        ctx.ddx("DWDX", "W"); ctx.ddy("DWDY", "W")
        ctx.ddz_uv("DUDZ", "U"); ctx.ddz_uv("DVDZ", "V"); ctx.ddz_w("DWDZ", "W")
        ctx.wallstress_bottom(); ctx.wallstress_top()
    check(("DUDX", "DUDY", "DVDX", "DVDY", "DWDX", "DWDY"), slice(1, nz), MIX_FFT)
    check(("DUDZ", "DVDZ", "DWDZ"), slice(1, nz + 1), MIX_WORK)
    check(("USTAR_SFC",), slice(None), TOL)            # FP64 wall model
    grads = ("DUDX", "DUDY", "DUDZ", "DVDX", "DVDY", "DVDZ", "DWDX", "DWDY", "DWDZ", "TXZ", "TYZ")
    sync_inputs(grads)
    for ctx in (o, g):
        ctx.sgs_stag(jt)
    check(("S11", "S12", "S13", "S22", "S23", "S33", "NU_T"), slice(1, nz + 1), MIX_WORK)
    check(("CS_OPT2", "F_LM", "F_MM", "F_QN", "F_NN"), slice(1, nz + 1), MIX_WORK)
    check(("TXX", "TXY", "TYY", "TZZ"), slice(1, nz), MIX_WORK)
    taus = ("TXX", "TXY", "TXZ", "TYY", "TYZ", "TZZ", "NU_T")
    sync_inputs(taus)
    for ctx in (o, g):
        ctx.divstress_uv("DIVTX", "TXX", "TXY", "TXZ")
        ctx.divstress_uv("DIVTY", "TXY", "TYY", "TYZ")
        ctx.divstress_w("DIVTZ", "TXZ", "TYZ", "TZZ")
    check(("DIVTX", "DIVTY"), slice(1, nz), MIX_WORK)
    check(("DIVTZ",), slice(1, nz + 1), MIX_WORK)
    sync_inputs(grads)
    for ctx in (o, g):
        ctx.convec()
    check(("CX", "CY", "CZ"), slice(1, nz), MIX_WORK)
    for ctx in (o, g):
        ctx.theta_all_in_one(jt)
    check(("THETA",), slice(1, nz), MIX_STATE)
    check(("RHST", "SGS_T3"), slice(1, nz), MIX_WORK)
    sync_inputs(("DIVTZ",))
    for ctx in (o, g):
        ctx.press_stag_array(jt)
    check(("P",), slice(0, nz), MIX_WORK)
    check(("DPDX", "DPDY", "DPDZ"), slice(1, nz), MIX_WORK)
    print(case, {k: float("%.3g" % v) for k, v in errs.items()})
    assert not over, over


@pytest.mark.parametrize("case", STEP_CASES)
def test_mixed_steps(case, oracle_mod, gpu_lib):
    """2*cs_count+1 free-running mixed steps (two dynamic-SGS updates)
    against the FP64 oracle: prognostic state and mean profiles."""
    nl, o, g = pair(case, oracle_mod, gpu_lib)
    g.set_precision(LES_MIXED)
    nz = nl.params().nzt
    nsteps = 2 * nl.cs_count + 1
    for jt in range(1, nsteps + 1):
        o.step(jt)
        g.step(jt)
    errs = {}
    # each velocity component's error relative to the velocity's norm (v and
    # w start at zero and stay an order below u)
    vel = {n: (g.get(n)[config.physical_levels(n, nz)], o.get(n)[config.physical_levels(n, nz)])
           for n in ("U", "V", "W")}
    vnorm = np.sqrt(sum(np.linalg.norm(b.ravel()) ** 2 for a, b in vel.values()))
    for n, (a, b) in vel.items():
        errs[n] = float(np.linalg.norm((a - b).ravel()) / vnorm)
    sl = config.physical_levels("THETA", nz)
    errs["THETA"] = rel(g.get("THETA")[sl], o.get("THETA")[sl])
    for n in ("U", "THETA"):
        errs["mean " + n] = rel(mean_profile(g.get(n)[1:nz]), mean_profile(o.get(n)[1:nz]))
    sl = slice(1, nz + 1)
    errs["CS_OPT2"] = rel(g.get("CS_OPT2")[sl], o.get("CS_OPT2")[sl])
    print(case, {k: float("%.3g" % v) for k, v in errs.items()})
    for n in ("U", "V", "W", "THETA"):
        assert errs[n] < MIX_STATE, (n, errs[n])
    for n in ("mean U", "mean THETA"):
        assert errs[n] < MIX_MEAN, (n, errs[n])
    assert errs["CS_OPT2"] < MIX_CS


def test_mixed_then_fp64(oracle_mod, gpu_lib):
    """Switching back to LES_FP64 restores the FP64 path: after a mixed
    step, a re-synced FP64 step matches the oracle at the FP64 bar."""
    nl, o, g = pair("small", oracle_mod, gpu_lib)
    g.set_precision(LES_MIXED)
    g.step(1)
    o.step(1)
    g.set_precision(LES_FP64)
    sync(o, g)
    o.step(2)
    g.step(2)
    bad = compare_state(o, g, ("U", "V", "W", "THETA"))
    assert not bad, bad


WORK_STATE = ("RHSX", "RHSY", "RHSZ", "RHSX_F", "RHSY_F", "RHSZ_F", "U_LAG", "V_LAG", "W_LAG")


def test_work_precision_state_follows_the_mode(gpu_lib):
    """The momentum RHS history and u_lag, v_lag, w_lag are state kept in
    the work precision: FP32 values once LES_MIXED is selected, converted
    back (exactly) by LES_FP64; FP64 state such as RHS_T and F_LM is
    untouched."""
    nl = config.workload("small")
    g = gpu_lib.create(nl.params())
    shape = g.get("RHSX").shape
    vals = {n: 1.0 + 1e-3 * config.uniform(40 + k, 1, shape)
            for k, n in enumerate(WORK_STATE + ("RHST", "F_LM"))}
    for n, a in vals.items():
        g.set(n, a)
    for prec in (LES_MIXED, LES_FP64):
        g.set_precision(prec)
        for n, a in vals.items():
            want = a.astype(np.float32).astype(np.float64) if n in WORK_STATE else a
            assert np.array_equal(g.get(n), want), (prec, n)


@pytest.mark.parametrize("mixed", [False, True])
def test_state_has_no_nyquist_modes_after_a_step(gpu_lib, mixed):
    """filter_data of u, v, w ends the step (103_flosol.f90:242-244): their
    kx = nx/2 and ky = ny/2 modes are zero, also with the mixed step's
    alternating-sum form of the filter."""
    from overlappingbls_amd.abi import LES_MIXED
    nl = config.workload("small")
    g = gpu_lib.create(nl.params())
    if mixed:
        g.set_precision(LES_MIXED)
    config.load_state(g, config.initial_fields(nl, seed=21))
    g.run(1, 2)
    nz = nl.params().nzt
    for name in ("U", "V", "W"):
        a = g.get(name)[1:nz]
        F = np.fft.fft2(a)
        ny, nx = a.shape[1:]
        scale = np.abs(F).max()
        assert np.abs(F[:, ny // 2, :]).max() < 1e-13 * scale, name
        assert np.abs(F[:, :, nx // 2]).max() < 1e-13 * scale, name
