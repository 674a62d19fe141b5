"""Restart files through the C-ABI (les_write_restart / les_read_restart)
against the oracle's (orc_write_restart / orc_read_restart, whose layout
tests/test_oracle_restart.py pins to checkpoint_final's), and a run resumed
from a restart against the same run straight through."""
import numpy as np
import pytest

from overlappingbls_amd import config
from overlappingbls_amd.abi import LES_E_IO, LesError

pytestmark = pytest.mark.gpu

STATE = ("U", "V", "W", "THETA", "RHSX", "RHSY", "RHSZ", "RHST", "CS_OPT2", "F_LM", "F_MM", "F_QN", "F_NN")


def test_gpu_and_oracle_write_the_same_bytes(oracle_mod, gpu_lib, tmp_path):
    """The same state (uploaded as is: no filter, whose transforms round
    differently) written by both gives identical files."""
    nl = config.workload("tiny_ragged")
    fields = config.initial_fields(nl, seed=11)
    files = {}
    for tag, mod in (("oracle", oracle_mod), ("gpu", gpu_lib)):
        c = mod.create(nl.params())
        config.load_state(c, fields, filter_velocity=False)
        d = tmp_path / tag
        d.mkdir()
        c.write_restart(str(d) + "/", 3)
        files[tag] = [np.fromfile(d / f"{s}00000003.out", dtype=np.uint8) for s in ("vel_tt", "temp_tt")]
    plane = 8 * nl.params().nxt * nl.params().nyt
    for a, b in zip(files["gpu"], files["oracle"]):
        assert a.size == b.size
        bad = np.nonzero(a != b)[0]
        assert bad.size == 0, f"first differing plane {bad[0] // plane}"


def test_gpu_reads_the_oracle_file(oracle_mod, gpu_lib, tmp_path):
    nl = config.workload("small")
    o = oracle_mod.create(nl.params())
    config.load_state(o, config.initial_fields(nl, seed=12))
    o.run(1, 2)
    path = str(tmp_path) + "/"
    o.write_restart(path, 2)
    o2 = oracle_mod.create(nl.params())
    o2.read_restart(path, 2)
    g = gpu_lib.create(nl.params())
    g.read_restart(path, 2)
    for name in STATE + ("SGS_T3", "PSI_M_SFC"):
        assert np.array_equal(g.get(name), o2.get(name)), name
    with pytest.raises(LesError) as e:
        g.read_restart(path, 3)
    assert e.value.status == LES_E_IO


def test_resumed_run_matches_the_oracle_resumed_run(oracle_mod, gpu_lib, tmp_path):
    """Resume from a restart written after a dynamic-SGS update step, on the
    GPU and in the oracle, and step on: the two runs agree at the parity
    bar.  (A resumed run does not equal the run that never stopped: the
    files hold levels 1..nzt-1 only, as the reference's do, so Cs_opt2 and
    F_* at level nzt start from the new context's values.)"""
    from test_gpu_parity import TOL, rel
    nl = config.workload("mid")
    cs = nl.cs_count
    nz = nl.params().nzt
    a = gpu_lib.create(nl.params())
    config.load_state(a, config.initial_fields(nl, seed=13))
    a.run(1, cs)
    path = str(tmp_path) + "/"
    a.write_restart(path, cs)
    flags = a.sgs_init_flags()
    runs = []
    for mod in (oracle_mod, gpu_lib):
        c = mod.create(nl.params())
        c.read_restart(path, cs)
        c.set_sgs_init_flags(*flags)
        c.run(cs + 1, cs + 1)
        runs.append(c)
    o, g = runs
    for name in ("U", "V", "W", "THETA"):
        assert rel(g.get(name)[1:nz], o.get(name)[1:nz]) < 1e3 * TOL, name
    # Cs^2 is a clipped ratio of small plane sums, the least well-conditioned
    # field (free-running parity at 128^3 leaves it within 1e-8)
    assert rel(g.get("CS_OPT2")[1:nz], o.get("CS_OPT2")[1:nz]) < 1e-6


@pytest.mark.parametrize("dim", [1, 2])
def test_gpu_writes_the_avgslice_records(gpu_lib, tmp_path, dim):
    """write_avgslice's files from the device window sums: every record is
    E14.5 of the window slice the library returns (the formatter is pinned
    in tests/test_oracle_restart.py)."""
    from avgslice_text import NAMES, records
    nl = config.workload("small")
    p = nl.params()
    g = gpu_lib.create(p)
    config.load_state(g, config.initial_fields(nl, seed=14))
    g.set_avgslice(True, 1, 2)
    g.run(1, 2)
    (tmp_path / "result").mkdir()
    path = str(tmp_path) + "/"
    g.write_avgslice(path, 2, dim)
    for name, stem in NAMES.items():
        lines = (tmp_path / "result" / f"aver_{stem}.out").read_text().splitlines()
        assert lines == records(g.avgslice(name, 1), 2 * p.dt, dim), name


@pytest.mark.parametrize("dim", [1, 2])
def test_gpu_avgslice_records_wrap_after_1400_values(gpu_lib, tmp_path, dim):
    """Records of more than 1400 values continue on a new line (fmt_5168's
    format reversion, 011_scalars2.f90:21), as the oracle writes them."""
    from avgslice_text import records
    shape = dict(nxt=1536, nyt=8, nzt=5) if dim == 1 else dict(nxt=8, nyt=8, nzt=1537)
    # grid spacing of the 64^3 workload in x and y, so the one step stays below CFL 0.6
    nl = config.Namelist(cs_count=2, dyn_init=2, lx_tot=shape["nxt"] / 64 * 200.0 * 3.14159,
                         ly_tot=shape["nyt"] / 64 * 200.0 * 3.14159, lz_tot=100.0, z_i=100.0, **shape)
    p = nl.params()
    g = gpu_lib.create(p)
    config.load_state(g, config.initial_fields(nl, seed=3))
    g.set_avgslice(True, 1, 1)
    g.run(1, 1)
    (tmp_path / "result").mkdir()
    g.write_avgslice(str(tmp_path) + "/", 1, dim)
    lines = (tmp_path / "result" / "aver_u.out").read_text().splitlines()
    assert [len(l) for l in lines[:2]] == [14 * 1400, 14 * 137]
    assert lines == records(g.avgslice("U", 1), p.dt, dim)
