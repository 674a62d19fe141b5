"""GPU parity of the per-routine drop-ins that stand alone at the boundary:
the pressure TDMA on host spectra (les_tdma_solve), Admas_Bashforth
(les_ab2), and the registration of caller arrays (les_bind_field +
les_upload/les_download).  Bar: FP64 relative L2 <= 1e-12 against the oracle
on the same inputs; the registration round trip is exact."""
import numpy as np
import pytest

from overlappingbls_amd import config

from test_gpu_parity import TOL, pair, rel

pytestmark = pytest.mark.gpu


def spectra(nz, ny, lhx, seed):
    sh = (nz + 1, ny, lhx)
    return [config.uniform(seed, 2 * i, sh) + 1j * config.uniform(seed, 2 * i + 1, sh) for i in range(3)]


@pytest.mark.parametrize("case", ["tiny_ragged", "mid", "wide"])
def test_tdma_solve(case, oracle_mod, gpu_lib):
    nl, o, g = pair(case, oracle_mod, gpu_lib)
    p = nl.params()
    Hx, Hy, Hz = spectra(p.nzt, p.nyt, p.nxt // 2 + 1, 5)
    a = g.tdma_solve(Hx, Hy, Hz)
    b = o.tdma_solve(Hx, Hy, Hz)
    assert rel(a, b) < TOL


@pytest.mark.parametrize("jt", [1, 2])
def test_ab2(jt, oracle_mod, gpu_lib):
    nl, o, g = pair("tiny_ragged", oracle_mod, gpu_lib)
    p = nl.params()
    sh = (p.nzt + 1, p.nyt, p.nxt)
    for ctx in (o, g):
        for i, n in enumerate(("RHSX", "RHSY", "RHSZ", "RHSX_F", "RHSY_F", "RHSZ_F")):
            ctx.set(n, config.uniform(13, i, sh))
        ctx.ab2(jt)
    for n in ("U", "V", "W", "RHSX_F"):
        assert rel(g.get(n), o.get(n)) < TOL, n


def test_bind_upload_download(gpu_lib):
    nl = config.workload("tiny_ragged")
    g = gpu_lib.create(nl.params())
    p = nl.params()
    sh = (p.nzt + 1, p.nyt, p.nxt)
    u = np.ascontiguousarray(config.uniform(3, 0, sh))
    th = np.ascontiguousarray(config.uniform(3, 1, sh))
    g.bind("U", u)
    g.bind("THETA", th)
    g.upload()
    assert np.array_equal(g.get("U"), u) and np.array_equal(g.get("THETA"), th)
    g.set("U", 3.0 * u)
    g.download()
    assert np.array_equal(u, 3.0 * config.uniform(3, 0, sh))


def test_namelist_options_the_step_does_not_support_are_rejected(gpu_lib):
    """param.nml options outside the step's scope reach les_create, which
    refuses them (the ocean branch, prescribed surface temperature, damping
    methods 2 and 3 with ubc = 1)."""
    from overlappingbls_amd import config
    from overlappingbls_amd.abi import LES_E_BADARG, LesError
    base = "&domain_param nxt = 16, nyt = 16, nzt = 9 /\n"
    for extra in ("&flow_param ocean_flag = .true. /", "&temperature_param lbc = 0 /",
                  "&flow_param ubc = 1, damping_method = 2 /"):
        with pytest.raises(LesError) as e:
            gpu_lib.create(config.read_namelist(base + extra).params())
        assert e.value.status == LES_E_BADARG, extra
