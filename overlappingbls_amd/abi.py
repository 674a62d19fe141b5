"""ctypes mirror of include/les_gpu.h and a thin object wrapper over it.

The same wrapper drives any library that exports the les_gpu.h entry points
under a symbol prefix: the HIP library exports them as ``les_*``.  Method
names follow the reference routines they replace (SBL/code/*.f90), so a test
reads like the reference's call sequence in flosol (103_flosol.f90:4-290).
"""
from __future__ import annotations

import ctypes as C
import os
import numpy as np

LES_OK = 0
LES_E_BADARG = -1
LES_E_HIP = -2
LES_E_CFL = -3
LES_E_NOMEM = -4
LES_E_NODEV = -5
LES_E_IO = -6
LES_E_SIZE = -7

K_TDMA, K_XDERIV, K_YDERIV, K_PAD, K_COPY = 0, 1, 2, 3, 4
LES_OBJ_USTAR2, LES_OBJ_T_FLUX = 0, 1

LES_FP64 = 0
LES_MIXED = 1

WALL_ROUGH, WALL_SMOOTH = 0, 1
BC_WALL, BC_STRESS_FREE = 0, 1
STOKES_DP, STOKES_MONO, STOKES_DYNAMIC = 0, 1, 2

_FIELDS3D = [
    "U", "V", "W", "THETA",
    "RHSX", "RHSY", "RHSZ", "RHST",
    "RHSX_F", "RHSY_F", "RHSZ_F", "RHST_F",
    "DUDX", "DUDY", "DUDZ", "DVDX", "DVDY", "DVDZ", "DWDX", "DWDY", "DWDZ",
    "TXX", "TXY", "TXZ", "TYY", "TYZ", "TZZ",
    "DIVTX", "DIVTY", "DIVTZ",
    "P", "DPDX", "DPDY", "DPDZ",
    "CS_OPT2", "F_LM", "F_MM", "F_QN", "F_NN",
    "NU_T", "U_LAG", "V_LAG", "W_LAG",
    "S11", "S12", "S13", "S22", "S23", "S33",
    "CX", "CY", "CZ",
    "DTDX", "DTDY", "DTDZ", "BETA_SCAL", "SGS_T3",
    "WORK0", "WORK1", "WORK2",
]
FIELD = {name: i for i, name in enumerate(_FIELDS3D)}
N_FIELDS3D = len(_FIELDS3D)
SURFACE = {"USTAR_SFC": 100, "PSI_M_SFC": 101, "PHI_M_SFC": 102, "PSI_H_SFC": 103, "PHI_H_SFC": 104,
           "USTAR_TOP_SFC": 105}
FIELD.update(SURFACE)

# time-averaged slices, enum les_slice (flow_slice then theta_slice)
FLOW_SLICES = [
    "U", "V", "W", "P", "DUDT", "DVDT", "DWDT", "DUDZ", "DVDZ",
    "U2", "V2", "W2", "P2", "UV", "UW", "VW", "WP",
    "TXX", "TXY", "TXZ", "TYY", "TYZ", "TZZ",
    "U3", "V3", "W3",
    "U2_FLUC", "V2_FLUC", "UV_FLUC", "U2W_FLUC", "V2W_FLUC", "DISSIP",
    "UTXZ_FLUC", "VTYZ_FLUC", "WTZZ", "CS",
]
THETA_SLICES = ["THETA", "THETA2", "T2_FLUC", "DTDZ", "SGS_T3", "WT", "NU_T"]
SLICE = {name: i for i, name in enumerate(FLOW_SLICES + THETA_SLICES)}


class LesParams(C.Structure):
    """les_params of include/les_gpu.h (field order and types must match)."""

    _fields_ = [
        ("nxt", C.c_int32), ("nyt", C.c_int32), ("nzt", C.c_int32),
        ("model", C.c_int32), ("cs_count", C.c_int32), ("dyn_init", C.c_int32),
        ("inilag", C.c_int32), ("restart_vel", C.c_int32), ("sgs", C.c_int32),
        ("molec", C.c_int32), ("walltype", C.c_int32),
        ("lbc_mom", C.c_int32), ("ubc_mom", C.c_int32),
        ("ubc", C.c_int32), ("damping_method", C.c_int32),
        ("coriolis_forcing", C.c_int32), ("use_mean_p_force", C.c_int32),
        ("ocean_flag", C.c_int32), ("theta_flag", C.c_int32),
        ("restart_theta", C.c_int32), ("theta_init_time", C.c_int32),
        ("passive_scalar", C.c_int32), ("lbc", C.c_int32), ("stokes_flag", C.c_int32),
        ("lx_tot", C.c_double), ("ly_tot", C.c_double), ("lz_tot", C.c_double),
        ("dt", C.c_double),
        ("z_i", C.c_double), ("u_scale", C.c_double), ("T_scale", C.c_double),
        ("nu_molec", C.c_double), ("zo", C.c_double), ("d0", C.c_double),
        ("Pr", C.c_double), ("wt_s", C.c_double), ("dTdz_top", C.c_double),
        ("coriol", C.c_double), ("ug", C.c_double), ("vg", C.c_double),
        ("mean_p_force", C.c_double), ("lz_npz", C.c_double),
        ("stokes_action", C.c_int32), ("flag_crossvel", C.c_int32),
        ("u_star", C.c_double), ("rad_stress", C.c_double), ("alpha_w", C.c_double),
        ("u_cross", C.c_double), ("v_cross", C.c_double),
        ("amp_w", C.c_double), ("lambda_w", C.c_double), ("agl_ust", C.c_double),
        ("inv_strength", C.c_double), ("T_s", C.c_double), ("Co", C.c_double),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


_PD = C.POINTER(C.c_double)


def _pd(a: np.ndarray):
    return a.ctypes.data_as(_PD)


class LesError(RuntimeError):
    def __init__(self, fn, status):
        super().__init__(f"{fn} returned status {status}")
        self.status = status


def bind(lib: C.CDLL, prefix: str):
    """Declare argument types of every les_gpu.h entry point under prefix."""
    P = C.c_void_p
    i = C.c_int
    sig = {
        "create": [C.POINTER(LesParams), C.POINTER(C.c_void_p)],
        "destroy": [P],
        "set_precision": [P, i],
        "set_stokes": [P, _PD, _PD],
        "set_sponge": [P, _PD],
        "set_wind_stress": [P, C.c_double, C.c_double],
        "get_profiles": [P, _PD, _PD, _PD],
        "set_field": [P, i, _PD, i, i, i],
        "get_field": [P, i, _PD, i, i, i],
        "set_sgs_init_flags": [P, i, i],
        "get_sgs_init_flags": [P, C.POINTER(i), C.POINTER(i)],
        "write_restart": [P, C.c_char_p, i],
        "read_restart": [P, C.c_char_p, i, i],
        "write_avgslice": [P, C.c_char_p, i, i],
        "ddx": [P, i, i], "ddy": [P, i, i], "ddz_uv": [P, i, i], "ddz_w": [P, i, i],
        "filter_data": [P, i],
        "test_filter": [P, i, i, i],
        "fftn": [P, _PD, i, _PD],
        "ifftn": [P, _PD, _PD, i],
        "zero_padding": [P, _PD, i, _PD],
        "zero_unpadding": [P, _PD, _PD, i],
        "zero_padding_noalloc": [P, C.c_void_p, i, C.c_void_p, i],
        "obukhov": [P, i],
        "wallstress_bottom": [P], "wallstress_top": [P],
        "sgs_stag": [P, i],
        "lagrange_sdep": [P, i],
        "interpolag_sdep": [P],
        "divstress_uv": [P, i, i, i, i], "divstress_w": [P, i, i, i, i],
        "convec": [P],
        "theta_all_in_one": [P, i],
        "press_stag_array": [P, i],
        "project": [P],
        "step": [P, i],
        "run": [P, i, i, C.POINTER(i)],
        "get_cfl": [P, C.POINTER(C.c_double)],
        "sync": [P],
        "time_kernel": [P, i, i, C.POINTER(C.c_double), C.POINTER(C.c_double)],
        "set_avgslice": [P, i, i, i],
        "get_avgslice": [P, i, i, _PD],
        "ab2": [P, i],
        "tdma_solve": [P, _PD, _PD, _PD, _PD],
        "bind_field": [P, i, _PD, i, i, i],
        "upload": [P],
        "download": [P],
        "lin_set": [P, i, _PD, i, i, i],
        "lin_get": [P, i, _PD, i, i, i],
        "lin_zero": [P],
        "set_sgs_frozen": [P, i],
        "tlm_window": [P, i, i, i, C.POINTER(C.c_double)],
        "step_tlm": [P, i],
        "adjoint_window": [P, i, i, i, i, C.POINTER(C.c_double)],
        "step_adj": [P, i],
        "lin_check": [P, i, i, C.c_ulonglong, C.POINTER(C.c_double), C.POINTER(C.c_double)],
        "lin_set_objective": [P, i],
        "lin_nstages": [],
        "lin_stage_times": [P, i, _PD, _PD],
        "profile_step": [P, i, _PD, _PD, C.POINTER(i)],
        "profile_run": [P, i, i, i, _PD, _PD, C.POINTER(i)],
        "nfamilies": [],
    }
    fns = {}
    for name, argtypes in sig.items():
        f = getattr(lib, prefix + name, None)
        if f is None:
            continue
        f.argtypes = argtypes
        f.restype = C.c_int
        fns[name] = f
    for name in ("lin_stage_name", "family_name"):
        f = getattr(lib, prefix + name, None)
        if f is not None:
            f.argtypes = [i]
            f.restype = C.c_char_p
            fns[name] = f
    return fns


class LesContext:
    """One domain's state behind a les_gpu.h-style library.

    Arrays at the boundary are numpy float64, C order (levels, nyt, nxt),
    i.e. the reference's (ldx, nynpy, 0:nz) column-major layout without the
    two padding columns.
    """

    def __init__(self, fns: dict, prm: LesParams, label: str):
        self._f = fns
        self.label = label
        self.prm = prm
        h = C.c_void_p()
        self._call("create", C.byref(prm), C.byref(h))
        self._h = h
        self.nx, self.ny, self.nz = prm.nxt, prm.nyt, prm.nzt
        self.nlev = self.nz + 1

    def _call(self, name, *args):
        st = self._f[name](*args)
        if st != LES_OK:
            raise LesError(f"{self.label}:{name}", st)
        return st

    def close(self):
        if getattr(self, "_h", None):
            self._f["destroy"](self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- state transfer -------------------------------------------------
    def set(self, name, arr, lev0=0):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        if name in SURFACE:
            a = a.reshape(1, self.ny, self.nx)
            self._call("set_field", self._h, FIELD[name], _pd(a), self.nx, 0, 1)
            return
        a = a.reshape(-1, self.ny, self.nx)
        self._call("set_field", self._h, FIELD[name], _pd(a), self.nx, lev0, a.shape[0])

    def get(self, name, lev0=0, nlev=None):
        if name in SURFACE:
            out = np.empty((self.ny, self.nx))
            self._call("get_field", self._h, FIELD[name], _pd(out), self.nx, 0, 1)
            return out
        if nlev is None:
            nlev = self.nlev - lev0
        out = np.empty((nlev, self.ny, self.nx))
        self._call("get_field", self._h, FIELD[name], _pd(out), self.nx, lev0, nlev)
        return out

    def set_stokes(self, ust, vst):
        u = np.ascontiguousarray(ust, dtype=np.float64)
        v = np.ascontiguousarray(vst, dtype=np.float64)
        self._call("set_stokes", self._h, _pd(u), _pd(v))

    def set_sponge(self, sponge):
        a = np.ascontiguousarray(sponge, dtype=np.float64)
        self._call("set_sponge", self._h, _pd(a))

    def profiles(self):
        """(ust, vst, sponge) over levels 0..nzt."""
        a = [np.zeros(self.nlev) for _ in range(3)]
        self._call("get_profiles", self._h, *[_pd(x) for x in a])
        return tuple(a)

    def set_wind_stress(self, ustar, agl_deg):
        """Ocean wind stress (friction velocity / u_scale, angle in degrees)
        for the following steps (104_wallstress.f90:118-136)."""
        self._call("set_wind_stress", self._h, float(ustar), float(agl_deg))

    def set_precision(self, p):
        self._call("set_precision", self._h, p)

    def sgs_init_flags(self):
        a, b = C.c_int(), C.c_int()
        self._call("get_sgs_init_flags", self._h, C.byref(a), C.byref(b))
        return a.value, b.value

    def set_sgs_init_flags(self, a, b):
        self._call("set_sgs_init_flags", self._h, int(a), int(b))

    def write_restart(self, path, nums):
        """checkpoint_final (012_io.f90:496-559): <path>vel_tt<nums>.out and,
        with theta_flag, <path>temp_tt<nums>.out; path ends in '/'."""
        self._call("write_restart", self._h, os.fsencode(path), int(nums))

    def write_avgslice(self, path, nums, average_dim_num=2):
        """write_avgslice (011_scalars2.f90:393-442) of the last closed
        window: appends to <path>result/aver_<name>.out."""
        self._call("write_avgslice", self._h, os.fsencode(path), int(nums), int(average_dim_num))

    def read_restart(self, path, nums, theta=True):
        """read_flow_field (102_01_ic.f90:682-734) and, with theta,
        init_temperature_field (102_initial.f90:740-797)."""
        self._call("read_restart", self._h, os.fsencode(path), int(nums), int(bool(theta)))

    def cfl(self):
        x = C.c_double()
        self._call("get_cfl", self._h, C.byref(x))
        return x.value

    # ---- routines -------------------------------------------------------
    def ddx(self, dst, src):
        self._call("ddx", self._h, FIELD[dst], FIELD[src])

    def ddy(self, dst, src):
        self._call("ddy", self._h, FIELD[dst], FIELD[src])

    def ddz_uv(self, dst, src):
        self._call("ddz_uv", self._h, FIELD[dst], FIELD[src])

    def ddz_w(self, dst, src):
        self._call("ddz_w", self._h, FIELD[dst], FIELD[src])

    def filter_data(self, f):
        self._call("filter_data", self._h, FIELD[f])

    def test_filter(self, f, level, which=2):
        self._call("test_filter", self._h, FIELD[f], int(level), int(which))

    def fftn(self, plane):
        a = np.ascontiguousarray(plane, dtype=np.float64)
        out = np.zeros((self.ny, self.nx // 2 + 1), dtype=np.complex128)
        self._call("fftn", self._h, _pd(a), a.shape[-1], out.ctypes.data_as(_PD))
        return out

    def ifftn(self, spec):
        s = np.ascontiguousarray(spec, dtype=np.complex128)
        out = np.zeros((self.ny, self.nx))
        self._call("ifftn", self._h, s.ctypes.data_as(_PD), _pd(out), self.nx)
        return out

    def zero_padding(self, plane):
        a = np.ascontiguousarray(plane, dtype=np.float64)
        out = np.zeros((3 * self.ny // 2, 3 * self.nx // 2))
        self._call("zero_padding", self._h, _pd(a), a.shape[-1], _pd(out))
        return out

    def zero_unpadding(self, big):
        a = np.ascontiguousarray(big, dtype=np.float64)
        out = np.zeros((self.ny, self.nx))
        self._call("zero_unpadding", self._h, _pd(a), _pd(out), self.nx)
        return out

    def obukhov(self, jt):
        self._call("obukhov", self._h, int(jt))

    def wallstress_bottom(self):
        self._call("wallstress_bottom", self._h)

    def wallstress_top(self):
        self._call("wallstress_top", self._h)

    def sgs_stag(self, jt):
        self._call("sgs_stag", self._h, int(jt))

    def lagrange_sdep(self, jt):
        self._call("lagrange_sdep", self._h, int(jt))

    def interpolag_sdep(self):
        self._call("interpolag_sdep", self._h)

    def divstress_uv(self, divt, tx, ty, tz):
        self._call("divstress_uv", self._h, FIELD[divt], FIELD[tx], FIELD[ty], FIELD[tz])

    def divstress_w(self, divt, tx, ty, tz):
        self._call("divstress_w", self._h, FIELD[divt], FIELD[tx], FIELD[ty], FIELD[tz])

    def convec(self):
        self._call("convec", self._h)

    def theta_all_in_one(self, jt):
        self._call("theta_all_in_one", self._h, int(jt))

    def press_stag_array(self, jt):
        self._call("press_stag_array", self._h, int(jt))

    def ab2(self, jt):
        """Admas_Bashforth (103_flosol.f90:319-389) on the context's fields."""
        self._call("ab2", self._h, int(jt))

    def tdma_solve(self, hx, hy, hz):
        """p_hat (nzt, nyt, nxt/2+1) complex from Hx, Hy, Hz (nzt+1, nyt, nxt/2+1)."""
        a = [np.ascontiguousarray(x, dtype=np.complex128) for x in (hx, hy, hz)]
        out = np.zeros((self.nz, self.ny, self.nx // 2 + 1), dtype=np.complex128)
        self._call("tdma_solve", self._h, *[x.ctypes.data_as(_PD) for x in a], out.ctypes.data_as(_PD))
        return out

    def bind(self, name, arr, lev0=0):
        """Register a caller array (kept alive by this object) for upload/download."""
        if not hasattr(self, "_bound"):
            self._bound = []
        a = arr if name in SURFACE else arr.reshape(-1, self.ny, self.nx)
        assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
        nlev = 1 if name in SURFACE else a.shape[0]
        self._call("bind_field", self._h, FIELD[name], _pd(a), self.nx, 0 if name in SURFACE else lev0, nlev)
        self._bound.append(a)

    def upload(self):
        self._call("upload", self._h)

    def download(self):
        self._call("download", self._h)

    def project(self):
        self._call("project", self._h)

    def step(self, jt):
        """flosol (103_flosol.f90:4-290); raises LesError(status=LES_E_CFL)."""
        self._call("step", self._h, int(jt))

    def run(self, jt0, nsteps):
        failed = C.c_int(0)
        st = self._f["run"](self._h, int(jt0), int(nsteps), C.byref(failed))
        if st != LES_OK:
            raise LesError(f"{self.label}:run (step {failed.value})", st)

    def set_avgslice(self, use_avgslice, c_count, p_count):
        """use_avgslice, c_count, p_count (001_param.f90:73,140-141)."""
        self._call("set_avgslice", self._h, int(bool(use_avgslice)), int(c_count), int(p_count))

    def avgslice(self, name, window=0):
        """(nzt-1, nxt) y-averaged slice LES_S_<name>; window 0 running, 1 last closed."""
        out = np.empty((self.nz - 1, self.nx))
        self._call("get_avgslice", self._h, SLICE[name], int(window), _pd(out))
        return out

    def time_kernel(self, kernel, reps=20):
        """(ms per launch, algorithmic bytes per launch) of one kernel family."""
        ms, nb = C.c_double(), C.c_double()
        self._call("time_kernel", self._h, int(kernel), int(reps), C.byref(ms), C.byref(nb))
        return ms.value, nb.value

    def sync(self):
        if "sync" in self._f:
            self._call("sync", self._h)

    def profile_step(self, jt):
        """Run step jt with the kernel families timed (les_profile_step):
        {family: (device ms, algorithmic bytes, launches)}."""
        n = self._f["nfamilies"]()
        ms, nb = np.zeros(n), np.zeros(n)
        cnt = (C.c_int * n)()
        self._call("profile_step", self._h, int(jt), _pd(ms), _pd(nb), cnt)
        return {self._f["family_name"](k).decode(): (float(ms[k]), float(nb[k]), int(cnt[k])) for k in range(n)}

    def family_index(self, name):
        n = self._f["nfamilies"]()
        for k in range(n):
            if self._f["family_name"](k).decode() == name:
                return k
        raise KeyError(name)

    def profile_run(self, jt0, nsteps, family=None):
        """Run nsteps steps from jt0 with one kernel family (by name; None:
        every family) timed (les_profile_run): {family: (device ms,
        algorithmic bytes, launches)} summed over the steps."""
        n = self._f["nfamilies"]()
        ms, nb = np.zeros(n), np.zeros(n)
        cnt = (C.c_int * n)()
        fam = -1 if family is None else self.family_index(family)
        self._call("profile_run", self._h, int(jt0), int(nsteps), fam, _pd(ms), _pd(nb), cnt)
        return {self._f["family_name"](k).decode(): (float(ms[k]), float(nb[k]), int(cnt[k])) for k in range(n)}

    # ---- tangent-linear model and adjoint (FP64, frozen SGS) -------------
    def lin_set(self, name, arr, lev0=0):
        """Set the linear state (tangent or adjoint seed) of U, V, W, THETA,
        RHSX, RHSY, RHSZ, RHST (levels), PSI_M_SFC (one plane) or SGS_T3
        (the surface flux plane: lev0 = 1, one level)."""
        a = np.ascontiguousarray(arr, dtype=np.float64)
        if name in SURFACE:
            a = a.reshape(1, self.ny, self.nx)
            self._call("lin_set", self._h, FIELD[name], _pd(a), self.nx, 0, 1)
            return
        a = a.reshape(-1, self.ny, self.nx)
        self._call("lin_set", self._h, FIELD[name], _pd(a), self.nx, lev0, a.shape[0])

    def lin_get(self, name, lev0=0, nlev=None):
        if name in SURFACE:
            out = np.empty((self.ny, self.nx))
            self._call("lin_get", self._h, FIELD[name], _pd(out), self.nx, 0, 1)
            return out
        if nlev is None:
            nlev = self.nlev - lev0
        out = np.empty((nlev, self.ny, self.nx))
        self._call("lin_get", self._h, FIELD[name], _pd(out), self.nx, lev0, nlev)
        return out

    def set_objective(self, kind):
        """Objective of the derivative windows: LES_OBJ_USTAR2 (mean ustar^2)
        or LES_OBJ_T_FLUX (mean surface flux sgs_t3(:,:,1))."""
        self._call("lin_set_objective", self._h, int(kind))

    def lin_zero(self):
        self._call("lin_zero", self._h)

    def set_sgs_frozen(self, frozen=True):
        self._call("set_sgs_frozen", self._h, int(bool(frozen)))

    def tlm_window(self, jt0, nsteps, obj_steps=0):
        """Advance the state and the tangent nsteps; returns the tangent of J."""
        dj = C.c_double(0.0)
        self._call("tlm_window", self._h, int(jt0), int(nsteps), int(obj_steps), C.byref(dj))
        return dj.value

    def adjoint_window(self, jt0, nsteps, k_ckpt, obj_steps=0):
        """Adjoint of the window (checkpoint every k_ckpt; 0: spacing chosen
        from free device memory, up to the whole window taped); returns J."""
        j = C.c_double(0.0)
        self._call("adjoint_window", self._h, int(jt0), int(nsteps), int(k_ckpt), int(obj_steps), C.byref(j))
        return j.value

    def lin_check(self, stage, jt, seed=1):
        """(<M dx, y>, <dx, M^T y>) of one stage (-1: the whole step)."""
        a, b = C.c_double(), C.c_double()
        self._call("lin_check", self._h, int(stage), int(jt), C.c_ulonglong(seed), C.byref(a), C.byref(b))
        return a.value, b.value

    def lin_stage_times(self, jt):
        """{stage: (tangent ms, adjoint ms)} about forward step jt."""
        names = self.lin_stages()
        a = np.zeros(len(names)); b = np.zeros(len(names))
        self._call("lin_stage_times", self._h, int(jt), _pd(a), _pd(b))
        return {n: (float(x), float(y)) for n, x, y in zip(names, a, b)}

    def lin_stages(self):
        n = self._f["lin_nstages"]()
        return [self._f["lin_stage_name"](k).decode() for k in range(n)]
