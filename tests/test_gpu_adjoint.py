"""Tangent-linear model and adjoint of the FP64 step (frozen SGS).

The reference has no derivative code, so nothing here is compared with it;
the bars are the mathematical identities the discrete derivative must obey:
  - dot-product test <M dx, y> = <dx, M^T y> per stage, per step and over
    checkpointed windows, relative difference <= DOT_TOL = 1e-10 (FP64
    rounding gives ~1e-14);
  - the tangent against central differences of the forward step (SGS frozen
    on both sides), eps = 1e-6 relative: truncation O(eps^2) plus rounding
    O(1e-16/eps), bar FD_TOL = 1e-6 relative L2;
  - the gradient of J (mean ustar^2, or mean surface heat flux sgs_t3(:,:,1),
    over the last steps) against central differences of J along seeded
    directions, same bar;
  - checkpoint spacing does not change the gradient (bitwise);
  - the adjoint window leaves the forward state where a plain forward run
    leaves it.
Cases: tiny (neutral obukhov branch, wt_s = 0), tiny_ragged (unstable branch,
Coriolis, sponge, 24 x 12 lines), small and mid (stable branch), and with
the surface temperature prescribed (lbc = 0, the bulk surface flux a
function of the state): tiny_lbc0 (cooled surface, stable branch) and
ragged_lbc0 (warmed surface, unstable branch, Coriolis, sponge).
"""
import numpy as np
import pytest

from overlappingbls_amd import config
from overlappingbls_amd.abi import LES_OBJ_T_FLUX, LES_OBJ_USTAR2

pytestmark = pytest.mark.gpu

DOT_TOL = 1e-10
FD_TOL = 1e-6
CASES = ["tiny", "tiny_ragged", "small", "mid", "tiny_lbc0", "ragged_lbc0"]
LBC0 = ("tiny_lbc0", "ragged_lbc0")
STATE = ("U", "V", "W", "THETA", "RHSX", "RHSY", "RHSZ", "RHST")
PLANES = ("PSI_M_SFC", "SGS_T3")   # one-plane state: psi_m, the surface flux sgs_t3(:,:,1)
JT0 = 3   # two forward steps first: RHS history, Lagrangian coefficients


def spun_up(case, gpu_lib, frozen=True, seed=1234, objective=LES_OBJ_USTAR2):
    nl = config.workload(case)
    g = gpu_lib.create(nl.params())
    config.load_state(g, config.initial_fields(nl, seed=seed))
    g.run(1, JT0 - 1)
    g.set_sgs_frozen(frozen)
    g.set_objective(objective)
    return nl, g


def objective_value(g, objective):
    """This step's contribution to J: the plane mean of ustar^2 or of the
    surface flux sgs_t3(:,:,1)."""
    if objective == LES_OBJ_T_FLUX:
        return float(np.mean(g.get("SGS_T3")[1]))
    return float(np.mean(g.get("USTAR_SFC") ** 2))


def cut(d, nz):
    """Levels 1..nz-1 of the 3-D state, the planes as they are."""
    return {n: (a if n in PLANES else a[1:nz]) for n, a in d.items()}


def get_plane(g, n):
    return g.get(n)[1] if n == "SGS_T3" else g.get(n)


def set_plane(g, n, a):
    if n == "SGS_T3":
        x = g.get(n)
        x[1] = a
        g.set(n, x)
    else:
        g.set(n, a)


def rel_diff(a, b):
    return abs(a - b) / max(abs(a), abs(b), 1e-300)


def direction(g, seed, scale=None):
    """Seeded random state direction on levels 1..nz-1, scaled per field by
    the field's rms (scale: dict of rms values) or 1."""
    nz = g.nz
    d = {}
    for k, name in enumerate(STATE):
        a = np.zeros((g.nlev, g.ny, g.nx))
        a[1:nz] = config.uniform(seed, 40 + k, (nz - 1, g.ny, g.nx))
        if scale is not None:
            a *= scale[name]
        d[name] = a
    d["PSI_M_SFC"] = np.full((g.ny, g.nx), 0.1 if scale is None else scale["PSI_M_SFC"])
    d["SGS_T3"] = config.uniform(seed, 60, (1, g.ny, g.nx))[0] * (1.0 if scale is None else scale["SGS_T3"])
    return d


def state_scale(g):
    nz = g.nz
    s = {n: float(np.sqrt(np.mean(g.get(n)[1:nz] ** 2))) for n in STATE}
    s["PSI_M_SFC"] = max(abs(float(g.get("PSI_M_SFC").mean())), 1e-3)
    s["SGS_T3"] = max(float(np.sqrt(np.mean(g.get("SGS_T3")[1] ** 2))), 1e-5)
    return s


def lin_load(g, d):
    g.lin_zero()
    for n, a in d.items():
        g.lin_set(n, a, lev0=1 if n == "SGS_T3" else 0)


def lin_read(g):
    nz = g.nz
    out = {n: g.lin_get(n)[1:nz] for n in STATE}
    out["PSI_M_SFC"] = g.lin_get("PSI_M_SFC")
    out["SGS_T3"] = g.lin_get("SGS_T3", lev0=1, nlev=1)[0]
    return out


def inner(a, b):
    return sum(float(np.sum(a[n] * b[n])) for n in a)


@pytest.mark.parametrize("case", CASES)
def test_stage_dot_products(case, gpu_lib):
    nl, g = spun_up(case, gpu_lib)
    names = g.lin_stages()
    assert len(names) >= 20
    bad = {}
    for k, name in enumerate(names):
        lhs, rhs = g.lin_check(k, JT0 + k, seed=11 + k)
        if (name == "obukhov" and case == "tiny") or (name == "surface_flux" and case not in LBC0):
            # neutral branch (wt_s = 0): psi_m and phi_m do not depend on the
            # state; lbc = 1: the surface flux is prescribed
            assert lhs == 0.0 and rhs == 0.0
            continue
        assert lhs != 0.0, name
        if not rel_diff(lhs, rhs) < DOT_TOL:
            bad[name] = (lhs, rhs, rel_diff(lhs, rhs))
    assert not bad, bad


@pytest.mark.parametrize("case", CASES)
def test_step_dot_product(case, gpu_lib):
    nl, g = spun_up(case, gpu_lib)
    for s in range(3):
        lhs, rhs = g.lin_check(-1, JT0 + s, seed=101 + s)
        print(case, s, lhs, rhs, rel_diff(lhs, rhs))
        assert rel_diff(lhs, rhs) < DOT_TOL


def forward_state(g):
    nz = g.nz
    out = {n: g.get(n)[1:nz] for n in STATE}
    for n in PLANES:
        out[n] = get_plane(g, n)
    return out


def perturbed_run(case, gpu_lib, d, eps, nsteps, obj_steps=0, objective=LES_OBJ_USTAR2):
    """State and J after nsteps forward steps from x0 + eps d (SGS frozen)."""
    nl, g = spun_up(case, gpu_lib)
    for n, a in d.items():
        if n in PLANES:
            set_plane(g, n, get_plane(g, n) + eps * a)
        else:
            x = g.get(n)
            x[1:g.nz] += eps * a[1:g.nz]
            g.set(n, x)
    J = 0.0
    for i in range(nsteps):
        g.step(JT0 + i)
        if i >= nsteps - obj_steps:
            J += objective_value(g, objective) / obj_steps
    return forward_state(g), J


FD_CASES = [("tiny_ragged", LES_OBJ_USTAR2), ("mid", LES_OBJ_USTAR2), ("ragged_lbc0", LES_OBJ_T_FLUX),
            ("tiny_lbc0", LES_OBJ_T_FLUX)]


@pytest.mark.parametrize("case,objective", FD_CASES)
def test_tlm_against_central_differences(case, objective, gpu_lib):
    nl, g = spun_up(case, gpu_lib, objective=objective)
    sc = state_scale(g)
    d = direction(g, 5, sc)
    nsteps, eps = 2, 1e-6
    lin_load(g, d)
    dJ = g.tlm_window(JT0, nsteps, obj_steps=2)
    tl = lin_read(g)
    xp, Jp = perturbed_run(case, gpu_lib, d, eps, nsteps, 2, objective)
    xm, Jm = perturbed_run(case, gpu_lib, d, -eps, nsteps, 2, objective)
    errs = {}
    for n in STATE + PLANES:
        fd = (xp[n] - xm[n]) / (2 * eps)
        errs[n] = float(np.linalg.norm(fd - tl[n]) / max(np.linalg.norm(fd), 1e-300))
    fdJ = (Jp - Jm) / (2 * eps)
    errs["J"] = rel_diff(fdJ, dJ)
    print(case, errs)
    assert all(v < FD_TOL for v in errs.values()), errs


@pytest.mark.parametrize("case,objective", [("tiny_ragged", LES_OBJ_USTAR2), ("small", LES_OBJ_USTAR2),
                                            ("ragged_lbc0", LES_OBJ_T_FLUX), ("tiny_lbc0", LES_OBJ_USTAR2)])
def test_window_dot_product_and_objective(case, objective, gpu_lib):
    nsteps, k, obj = 4, 2, 2
    nl, ga = spun_up(case, gpu_lib, objective=objective)
    dx = direction(ga, 7)
    y = direction(ga, 8)
    lin_load(ga, dx)
    dJ = ga.tlm_window(JT0, nsteps, obj_steps=obj)
    tl = lin_read(ga)
    nl, gb = spun_up(case, gpu_lib, objective=objective)
    lin_load(gb, y)
    J = gb.adjoint_window(JT0, nsteps, k, obj_steps=obj)
    grad = lin_read(gb)
    lhs = inner(tl, cut(y, ga.nz)) + dJ
    rhs = inner(cut(dx, ga.nz), grad)
    print(case, lhs, rhs, rel_diff(lhs, rhs), J)
    assert J != 0.0
    # J itself: the mean of the per-step objective over the last obj steps
    nl, gc = spun_up(case, gpu_lib)
    Jf = 0.0
    for i in range(nsteps):
        gc.step(JT0 + i)
        if i >= nsteps - obj:
            Jf += objective_value(gc, objective) / obj
    assert rel_diff(J, Jf) < 1e-12
    assert rel_diff(lhs, rhs) < DOT_TOL
    # the adjoint window leaves the forward state of a plain run
    nl, gc = spun_up(case, gpu_lib)
    gc.run(JT0, nsteps)
    for n in ("U", "V", "W", "THETA"):
        assert np.array_equal(gb.get(n)[1:gb.nz], gc.get(n)[1:gc.nz]), n


def test_checkpoint_spacing_does_not_change_the_gradient(gpu_lib):
    nsteps = 4
    grads = []
    for k in (1, 2, 4, 0):   # 0: spacing from free device memory (here the whole window)
        nl, g = spun_up("tiny_ragged", gpu_lib)
        lin_load(g, direction(g, 9))
        g.adjoint_window(JT0, nsteps, k, obj_steps=2)
        grads.append(lin_read(g))
    for other in grads[1:]:
        for n in grads[0]:
            assert np.array_equal(grads[0][n], other[n]), n


@pytest.mark.parametrize("case,objective", FD_CASES)
def test_gradient_against_central_differences(case, objective, gpu_lib):
    nsteps, k, obj, eps = 3, 2, 2, 1e-6
    nl, g = spun_up(case, gpu_lib, objective=objective)
    sc = state_scale(g)
    g.lin_zero()
    g.adjoint_window(JT0, nsteps, k, obj_steps=obj)
    grad = lin_read(g)
    errs = []
    for s in range(2):
        d = direction(g, 20 + s, sc)
        adj = inner(cut(d, g.nz), grad)
        _, Jp = perturbed_run(case, gpu_lib, d, eps, nsteps, obj, objective)
        _, Jm = perturbed_run(case, gpu_lib, d, -eps, nsteps, obj, objective)
        fd = (Jp - Jm) / (2 * eps)
        errs.append(rel_diff(fd, adj))
    print(case, errs)
    assert all(e < FD_TOL for e in errs), errs


def test_frozen_forward_matches_oracle(oracle_mod, gpu_lib):
    """The frozen-SGS forward (Cs_opt2 and F_* held) against the oracle's."""
    from test_gpu_parity import TOL, pair, rel
    nl, o, g = pair("tiny_ragged", oracle_mod, gpu_lib)
    for ctx in (o, g):
        ctx.run(1, 2)
        ctx.set_sgs_frozen(True)
        ctx.run(3, 2)
    nz = nl.params().nzt
    for n in ("U", "V", "W", "THETA", "CS_OPT2"):
        assert rel(g.get(n)[1:nz], o.get(n)[1:nz]) < 1e3 * TOL, n
