"""The C-ABI library loads and exports every entry point include/les_gpu.h
declares; parameters round-trip through the ctypes mirror (CPU only)."""
import ctypes as C
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "les_gpu.h")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|void\*)\s+(les_\w+)\s*\(", text, re.M)))


def test_header_declares_the_routines():
    names = declared()
    for must in ("les_create", "les_step", "les_press_stag_array", "les_sgs_stag", "les_ddx",
                 "les_theta_all_in_one", "les_convec", "les_lagrange_sdep", "les_interpolag_sdep"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from overlappingbls_amd import lib
    if not os.path.exists(lib.LIB):
        pytest.skip("liblesgpu.so not built (run __graft_entry__.build())")
    out = subprocess.run(["nm", "-D", "--defined-only", lib.LIB], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\sT\s(les_\w+)", out))
    missing = [n for n in declared() if n not in exported]
    assert not missing, missing
    L = C.CDLL(lib.LIB)
    for n in declared():
        getattr(L, n)


def test_oracle_exports_the_same_surface(oracle_mod):
    out = subprocess.run(["nm", "-D", "--defined-only", oracle_mod.LIB], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\sT\s(orc_\w+)", out))
    # device-pointer and timing entry points have no CPU counterpart
    skip = {"les_stream", "les_sync", "les_time_kernel", "les_set_sponge", "les_zero_padding_noalloc",
            "les_nfamilies", "les_profile_step", "les_profile_run"}
    # the derivative code has no counterpart in the reference, hence none in the oracle
    lin = {"les_lin_set", "les_lin_get", "les_lin_zero", "les_tlm_window", "les_step_tlm", "les_adjoint_window",
           "les_step_adj", "les_lin_check", "les_lin_nstages", "les_lin_stage_times", "les_lin_set_objective"}
    missing = [n for n in declared() if n not in skip | lin and n.replace("les_", "orc_", 1) not in exported]
    assert not missing, missing


def test_params_struct_layout(tmp_path):
    """The ctypes mirror has the C header's size and field offsets (checked
    against gcc's layout of include/les_gpu.h)."""
    import shutil
    from overlappingbls_amd.abi import LesParams
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    names = [n for n, _ in LesParams._fields_]
    src = tmp_path / "layout.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "les_gpu.h"\nint main(void) {\n'
                   '  printf("%zu\\n", sizeof(les_params));\n'
                   + "".join(f'  printf("%zu\\n", offsetof(les_params, {n}));\n' for n in names)
                   + "  return 0;\n}\n")
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    vals = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert vals[0] == C.sizeof(LesParams)
    for n, off in zip(names, vals[1:]):
        assert getattr(LesParams, n).offset == off, n


def test_config_nondimensionalisation():
    from overlappingbls_amd import config
    nl = config.workload("sbl64")
    p = nl.params()
    assert abs(p.lx_tot - 2 * 3.141592653589793) < 1e-12
    assert abs(p.lz_tot - 1.0) < 1e-15
    assert abs(p.dt - 2e-4) < 1e-18
    assert abs(p.zo - 1e-4) < 1e-18           # zo1 = 0.01 m over z_i = 100 m
    assert p.nzt == 65 and p.lz_npz == 1.0    # lz*npz of step_scalar


def test_fortran_binding_compiles_and_links(tmp_path):
    """les_gpu_mod (ISO_C_BINDING) and a driver program compile with amdflang
    and link against liblesgpu.so; nothing is run here."""
    import shutil
    from overlappingbls_amd import lib
    fc = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
    if not os.path.exists(fc):
        pytest.skip("amdflang not available")
    if not os.path.exists(lib.LIB):
        pytest.skip("liblesgpu.so not built")
    mod = os.path.join(ROOT, "overlappingbls_amd", "fortran", "les_gpu_mod.f90")
    drv = os.path.join(ROOT, "tests", "fortran", "drive_step.f90")
    subprocess.run([fc, "-c", mod, "-o", str(tmp_path / "m.o"), "-module-dir", str(tmp_path)], check=True, cwd=tmp_path)
    subprocess.run([fc, "-c", drv, "-o", str(tmp_path / "d.o"), "-I", str(tmp_path)], check=True, cwd=tmp_path)
    libdir = os.path.dirname(lib.LIB)
    subprocess.run([fc, str(tmp_path / "d.o"), str(tmp_path / "m.o"), "-o", str(tmp_path / "drive"),
                    lib.LIB, f"-Wl,-rpath,{libdir}"], check=True, cwd=tmp_path)
    assert os.path.exists(tmp_path / "drive")
    # the per-routine drop-ins with the reference's interfaces, against
    # stand-ins of its param / sim_param modules, and their flosol-order driver
    for obj, src in (("mm.o", os.path.join(ROOT, "tests", "fortran", "mock_modules.f90")),
                     ("di.o", os.path.join(ROOT, "overlappingbls_amd", "fortran", "les_dropin.f90")),
                     ("dr.o", os.path.join(ROOT, "tests", "fortran", "drive_routines.f90"))):
        subprocess.run([fc, "-c", src, "-o", str(tmp_path / obj), "-I", str(tmp_path), "-module-dir", str(tmp_path)],
                       check=True, cwd=tmp_path)
    subprocess.run([fc, str(tmp_path / "dr.o"), str(tmp_path / "di.o"), str(tmp_path / "mm.o"), str(tmp_path / "m.o"),
                    "-o", str(tmp_path / "drive_routines"), lib.LIB, f"-Wl,-rpath,{libdir}"], check=True, cwd=tmp_path)
    assert os.path.exists(tmp_path / "drive_routines")


def test_counter_summary_maps_every_family():
    """bench.py's counter and trace summaries assign kernels to the
    profiler's families by name: every family the library reports
    (les_family_name, callable without a GPU) is the image of some kernel the
    library contains (its host-side launch symbols)."""
    import bench
    from overlappingbls_amd import lib
    if not os.path.exists(lib.LIB):
        pytest.skip("library not built")
    L = C.CDLL(lib.LIB)
    L.les_family_name.restype = C.c_char_p
    fams = {L.les_family_name(k).decode() for k in range(L.les_nfamilies())}
    out = subprocess.run(["nm", "-C", "--defined-only", lib.LIB], capture_output=True, text=True, check=True).stdout
    kernels = [ln.split(" ", 2)[2] for ln in out.splitlines()
               if re.search(r"\s[DVdv]\s.*les::(\(anonymous namespace\)::)?kr?_", ln)]
    mapped = {bench.kernel_family(k) for k in kernels}
    assert fams <= mapped, sorted(fams - mapped)


def test_bench_line_describes_its_workload():
    """The bench line names the workload it ran, and the counter-based traffic
    (taken on the 256^3 surface layer) is attached to that workload only."""
    import bench
    from overlappingbls_amd import config
    for name, word in (("sbl256", "surface BL"), ("bbl512", "ocean bottom boundary layer")):
        nl = config.workload(name)
        text = bench.workload_text(name, nl, nl.params())
        assert text.startswith(name + ":") and word in text, text
    assert bench.traffic_from_profiles("yline.pad", "sbl256") is not None
    assert bench.traffic_from_profiles("yline.pad", "bbl512") is None
