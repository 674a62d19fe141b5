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


# deep16 (nzt = 129) runs both the levels the TDMA keeps in LDS and the ones
# whose d' goes through HBM, and the boundary between them
@pytest.mark.parametrize("case", ["tiny_ragged", "mid", "wide", "deep16"])
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


def test_options_the_step_does_not_support_are_rejected(gpu_lib):
    """les_create refuses the physics options it does not implement: other
    SGS models, sgs = .false., the Donelan-Pierson Stokes spectrum, and lbc
    = 0 in the ocean (whose surface ranks never set ustar_avg, psi_h)."""
    from overlappingbls_amd.abi import LES_E_BADARG, LesError
    base = ("&domain_param nxt = 16, nyt = 16, nzt = 9 /\n&canopy_param flag_canopy = .false. /\n"
            "&ocean_param stokes_flag = .false. /\n")
    for extra in ("&flow_param model = 2 /", "&flow_param sgs = .false. /",
                  "&ocean_param stokes_flag = .true., stokes_action = 'donelan.pierson1987' /\n"
                  "&flow_param ubc_mom = 'stress free' /",
                  "&flow_param ocean_flag = .true. /\n&temperature_param lbc = 0 /"):
        with pytest.raises(LesError) as e:
            gpu_lib.create(config.read_namelist(base + extra).params())
        assert e.value.status == LES_E_BADARG, extra


def test_reference_configurations_are_accepted(gpu_lib):
    """The options of the reference's own param.nml files (ocean branch,
    wall top, damping method 3, monochromatic Stokes drift) and lbc = 0 on
    land create a context."""
    base = "&domain_param nxt = 16, nyt = 16, nzt = 9 /\n&canopy_param flag_canopy = .false. /\n"
    for extra in ("&flow_param ocean_flag = .true., lbc_mom = 'stress free', ubc_mom = 'wall', ubc = 0,"
                  " damping_method = 3 /\n&ocean_param stokes_flag = .true., stokes_action = 'mono',"
                  " amp_w = 1.13d0, lambda_w = 60.d0 /",
                  "&flow_param ubc_mom = 'stress free' /\n&ocean_param stokes_flag = .false. /\n"
                  "&temperature_param lbc = 0, theta_flag = .true. /",
                  "&flow_param ubc = 1, ubc_mom = 'stress free', damping_method = 2 /\n"
                  "&ocean_param stokes_flag = .false. /"):
        g = gpu_lib.create(config.read_namelist(base + extra).params())
        g.close()
