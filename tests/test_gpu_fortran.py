"""The Fortran boundary end to end on the GPU: tests/fortran/drive_window.f90
(built with amdflang against les_gpu_mod and liblesgpu.so) registers
module-style arrays with ldx = nxt+2 columns, uploads, runs flosol steps,
downloads, and runs a frozen-SGS adjoint window.  Its outputs must equal,
bit for bit, the same calls made through the ctypes binding."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from overlappingbls_amd import config

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def build_driver(tmp):
    from overlappingbls_amd import lib
    fc = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
    if not os.path.exists(fc):
        pytest.skip("amdflang not available")
    mod = os.path.join(ROOT, "overlappingbls_amd", "fortran", "les_gpu_mod.f90")
    drv = os.path.join(ROOT, "tests", "fortran", "drive_window.f90")
    subprocess.run([fc, "-c", mod, "-o", "m.o", "-module-dir", "."], check=True, cwd=tmp)
    subprocess.run([fc, "-c", drv, "-o", "d.o", "-I", "."], check=True, cwd=tmp)
    subprocess.run([fc, "d.o", "m.o", "-o", "drive", lib.LIB, f"-Wl,-rpath,{os.path.dirname(lib.LIB)}"],
                   check=True, cwd=tmp)
    return os.path.join(tmp, "drive")


def test_fortran_window_matches_ctypes(tmp_path, gpu_lib):
    tmp = str(tmp_path)
    exe = build_driver(tmp)
    nl = config.workload("tiny_ragged")
    p = nl.params()
    nsteps = 3
    fields = config.initial_fields(nl, seed=77)
    with open(os.path.join(tmp, "params.bin"), "wb") as f:
        f.write(bytes(p))
        f.write(np.int32(nsteps).tobytes())
    with open(os.path.join(tmp, "fields.bin"), "wb") as f:
        for n in config.STATE_FIELDS:
            f.write(np.ascontiguousarray(fields[n], dtype=np.float64).tobytes())
    r = subprocess.run([exe, tmp], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "fortran window ok" in r.stdout
    nz, ny, nx = p.nzt, p.nyt, p.nxt
    out = np.fromfile(os.path.join(tmp, "out.bin"), dtype=np.float64)
    n3, ng = (nz + 1) * ny * nx, (nz - 1) * ny * nx
    u_f, th_f = out[:n3].reshape(nz + 1, ny, nx), out[n3:2 * n3].reshape(nz + 1, ny, nx)
    J_f = out[2 * n3]
    gth_f = out[2 * n3 + 1:2 * n3 + 1 + ng].reshape(nz - 1, ny, nx)
    gu_f = out[2 * n3 + 1 + ng:].reshape(nz - 1, ny, nx)
    # the same through ctypes
    g = gpu_lib.create(p)
    config.load_state(g, fields)
    g.run(1, nsteps)
    assert np.array_equal(g.get("U"), u_f) and np.array_equal(g.get("THETA"), th_f)
    g.set_sgs_frozen(True)
    g.lin_zero()
    J = g.adjoint_window(nsteps + 1, 3, 2, obj_steps=2)
    assert J == J_f and J > 0.0
    assert np.array_equal(g.lin_get("THETA", 1, nz - 1), gth_f)
    assert np.array_equal(g.lin_get("U", 1, nz - 1), gu_f)
    assert np.abs(gu_f).max() > 0.0 and np.abs(gth_f).max() > 0.0


def build_routines_driver(tmp):
    from overlappingbls_amd import lib
    fc = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
    if not os.path.exists(fc):
        pytest.skip("amdflang not available")
    srcs = [("m.o", os.path.join(ROOT, "overlappingbls_amd", "fortran", "les_gpu_mod.f90")),
            ("mm.o", os.path.join(ROOT, "tests", "fortran", "mock_modules.f90")),
            ("di.o", os.path.join(ROOT, "overlappingbls_amd", "fortran", "les_dropin.f90")),
            ("dr.o", os.path.join(ROOT, "tests", "fortran", "drive_routines.f90"))]
    for obj, src in srcs:
        subprocess.run([fc, "-c", src, "-o", obj, "-I", ".", "-module-dir", "."], check=True, cwd=tmp)
    subprocess.run([fc, "dr.o", "di.o", "mm.o", "m.o", "-o", "drive_routines", lib.LIB,
                    f"-Wl,-rpath,{os.path.dirname(lib.LIB)}"], check=True, cwd=tmp)
    return os.path.join(tmp, "drive_routines")


def test_fortran_dropins_in_flosol_order_match_the_whole_step(tmp_path, gpu_lib):
    """overlappingbls_amd/fortran/les_dropin.f90's routines, with the
    reference's interfaces, called in flosol's order from Fortran on the
    reference's (ldx, nynpy, 0:nz) arrays (tests/fortran/drive_routines.f90,
    theta off), reproduce the library's whole step (les_run) on the same
    fields over three steps, the second a dynamic-SGS update."""
    from test_gpu_parity import rel
    tmp = str(tmp_path)
    exe = build_routines_driver(tmp)
    nl = config.workload("tiny")
    nl.theta_flag = False
    p = nl.params()
    nsteps = 3
    fields = config.initial_fields(nl, seed=91)
    with open(os.path.join(tmp, "params.bin"), "wb") as f:
        f.write(bytes(p))
        f.write(np.int32(nsteps).tobytes())
    with open(os.path.join(tmp, "fields.bin"), "wb") as f:
        for n in ("U", "V", "W"):
            f.write(np.ascontiguousarray(fields[n], dtype=np.float64).tobytes())
    r = subprocess.run([exe, tmp], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "fortran routines ok" in r.stdout
    nz, ny, nx = p.nzt, p.nyt, p.nxt
    out = np.fromfile(os.path.join(tmp, "out.bin"), dtype=np.float64).reshape(5, nz + 1, ny, nx)
    g = gpu_lib.create(p)
    config.load_state(g, fields)
    g.run(1, nsteps)
    errs = {}
    for k, (n, sl) in enumerate((("U", slice(1, nz)), ("V", slice(1, nz)), ("W", slice(1, nz + 1)),
                                 ("RHSX", slice(1, nz)), ("P", slice(0, nz)))):
        errs[n] = rel(out[k][sl], g.get(n)[sl])
    print("drop-ins against les_run", errs)
    assert all(e < 1e-12 for e in errs.values()), errs
