"""Run configuration: namelist-equivalent inputs, nondimensionalisation and
seeded synthetic initial fields for the boundary-layer LES step.

The reference reads Fortran namelists (001_param.f90) and nondimensionalises
them in init_nondimensional (102_initial.f90:176-256).  `Namelist` holds the
same inputs, dimensional where the namelist is dimensional, and
`Namelist.params()` returns the les_params block the C-ABI takes.

Synthetic fields come from a counter-based generator keyed by
(seed, field id, i, j, k), so a field does not depend on how the domain is
split, and the same seed gives the same field on every machine.
"""
from __future__ import annotations

import dataclasses
import math
import os
import re

import numpy as np

from .abi import (LesParams, WALL_ROUGH, WALL_SMOOTH, BC_WALL, BC_STRESS_FREE,
                  STOKES_DP, STOKES_MONO, STOKES_DYNAMIC)

BOGUS = -1234567890.0  # 001_param.f90:40
ZO1 = 0.01             # 008_bottombc.f90:18, roughness in metres
DO1 = 0.0              # 008_bottombc.f90:19
THETA_S1 = 288.0       # 008_bottombc.f90:23, surface temperature of the patches (K)
STOKES_ACTIONS = {"donelan.pierson1987": STOKES_DP, "mono": STOKES_MONO, "dynamic": STOKES_DYNAMIC}


class UnsupportedOption(ValueError):
    """A namelist option that selects physics outside the timestep this
    library provides (the reference would run it)."""


@dataclasses.dataclass
class Namelist:
    """Inputs of the namelists in 001_param.f90 that the step reads."""

    nxt: int = 64
    nyt: int = 64
    nzt: int = 65                       # z nodes incl. the top ghost
    lx_tot: float = 200.0 * math.pi     # metres (domain_param)
    ly_tot: float = 200.0 * math.pi
    lz_tot: float = 100.0
    z_i: float = 100.0
    dt: float = 2e-4 * 100.0            # seconds; nondimensional dt * z_i / u_scale
    u_scale: float = 1.0
    cs_count: int = 5
    dyn_init: int = 5
    inilag: bool = True
    restart_vel: bool = False
    model: int = 5                      # flow_param model (5: Lagrangian scale-dependent)
    sgs: bool = True
    Co: float = 0.16                    # flow_param Co (Smagorinsky, model 1)
    walltype: str = "rough"
    lbc_mom: str = "wall"
    ubc_mom: str = "stress free"
    ubc: int = 0
    damping_method: int = 1
    coriolis_forcing: bool = False
    freq_coriolis: float = 1e-4
    ug_dim: float = 0.0
    vg_dim: float = 0.0
    use_mean_p_force: bool = False
    mean_p_force: float | None = None   # None: float_missing, i.e. 1/lz_tot (102_initial.f90:214-219)
    nu_molec: float = 1.14e-5
    molec: bool = False
    Pr: float = 0.4
    theta_flag: bool = True
    restart_theta: bool = False
    theta_init_time: int = 1
    passive_scalar: bool = False
    T_scale: float = 300.0
    wt_s: float = 0.0                   # K m/s
    dTdz_top: float = 0.0               # K/m
    inv_strength: float = 0.0           # K/m (damping method 3)
    npz: int = 2                        # only lz*npz of step_scalar depends on it
    lbc: int = 1                        # 1: prescribed surface flux; 0: surface temperature
    theta_s1: float = THETA_S1          # K, surface temperature for lbc = 0
    # ocean branch (ocean_param, 001_param.f90:304-333)
    ocean_flag: bool = False
    u_star: float = 0.0                 # m/s, wind-stress friction velocity
    agl_stress: float = 0.0             # degrees
    flag_dynStress: bool = False        # per-step wind stress: the driver calls set_wind_stress
    alpha_w: float = 2.0e-4             # 1/K
    flag_crossVel: bool = False
    udim_cross: float = 0.0
    vdim_cross: float = 0.0
    stokes_flag: bool = False
    stokes_action: str = "mono"
    amp_w: float = 0.0                  # m
    lambda_w: float = 1.0               # m
    agl_Ust: float = 0.0                # degrees
    # options outside the library's step: params() refuses them when set
    dns_bc: bool = False
    inflow: bool = False
    flag_canopy: bool = False
    pcon_flag: bool = False
    flag_relaxForce: bool = False
    spatial_flux_flag: bool = False
    flag_dynWT: bool = False
    # roughness length / z_i; None: the reference's patch constant ZO1 / z_i
    # (008_bottombc.f90:18, :119), a module constant there, not a namelist entry
    zo: float | None = None

    def params(self) -> LesParams:
        """les_params as init_nondimensional leaves them (102_initial.f90:176-256);
        raises UnsupportedOption for physics the library does not provide."""
        refused = [name for name in ("dns_bc", "inflow", "flag_canopy", "pcon_flag", "flag_relaxForce",
                                     "spatial_flux_flag", "flag_dynWT") if getattr(self, name)]
        if refused:
            raise UnsupportedOption("namelist selects options outside the timestep: " + ", ".join(refused))
        if self.stokes_action not in STOKES_ACTIONS:
            raise UnsupportedOption(f"stokes_action {self.stokes_action!r} (014_stokes_drift.f90:45-56)")
        p = LesParams()
        p.nxt, p.nyt, p.nzt = self.nxt, self.nyt, self.nzt
        p.model, p.sgs = int(self.model), int(bool(self.sgs))
        p.cs_count, p.dyn_init = self.cs_count, self.dyn_init
        p.inilag, p.restart_vel = int(self.inilag), int(self.restart_vel)
        p.molec = int(self.molec)
        p.walltype = {"rough": WALL_ROUGH, "smooth": WALL_SMOOTH}[self.walltype]
        bc = {"wall": BC_WALL, "stress free": BC_STRESS_FREE}
        p.lbc_mom, p.ubc_mom = bc[self.lbc_mom], bc[self.ubc_mom]
        p.ubc, p.damping_method = self.ubc, self.damping_method
        p.coriolis_forcing = int(self.coriolis_forcing)
        p.use_mean_p_force = int(self.use_mean_p_force)
        p.ocean_flag = int(self.ocean_flag)
        p.theta_flag, p.restart_theta = int(self.theta_flag), int(self.restart_theta)
        p.theta_init_time, p.passive_scalar, p.lbc = self.theta_init_time, int(self.passive_scalar), self.lbc
        # init_nondimensional, 102_initial.f90:186-250
        z_i, us = self.z_i, self.u_scale
        p.lx_tot, p.ly_tot, p.lz_tot = self.lx_tot / z_i, self.ly_tot / z_i, self.lz_tot / z_i
        p.dt = self.dt * us / z_i
        p.z_i, p.u_scale, p.T_scale = z_i, us, self.T_scale
        p.nu_molec = self.nu_molec
        p.zo, p.d0 = ZO1 / z_i, DO1 / z_i          # patches, 008_bottombc.f90:119-120
        if self.zo is not None:
            p.zo = self.zo
        p.Pr, p.wt_s, p.dTdz_top = self.Pr, self.wt_s, self.dTdz_top
        p.Co = self.Co
        if self.coriolis_forcing or self.ocean_flag:   # :208-212, :235-238
            p.coriol = self.freq_coriolis * z_i / us
            p.ug, p.vg = self.ug_dim / us, self.vg_dim / us
        if self.use_mean_p_force:                   # :214-219
            p.mean_p_force = (1.0 / p.lz_tot) if self.mean_p_force is None else self.mean_p_force / us**2 * z_i
        lz = self.lz_tot / self.npz / z_i          # 102_initial.f90:142, 186
        p.lz_npz = lz * self.npz
        p.u_star = self.u_star / us                # :201
        p.alpha_w = self.alpha_w * self.T_scale if (self.ocean_flag and self.theta_flag) else self.alpha_w   # :228-230
        p.rad_stress = self.agl_stress / 180.0 * math.pi if self.ocean_flag else 0.0   # :246
        p.flag_crossvel = int(self.flag_crossVel and self.ocean_flag)
        if p.flag_crossvel:                        # :249-252
            p.u_cross, p.v_cross = self.udim_cross / us, self.vdim_cross / us
        p.stokes_flag = int(self.stokes_flag)
        p.stokes_action = STOKES_ACTIONS[self.stokes_action]
        p.amp_w, p.lambda_w, p.agl_ust = self.amp_w, self.lambda_w, self.agl_Ust
        p.inv_strength = self.inv_strength
        p.T_s = self.theta_s1 / self.T_scale       # patches, 008_bottombc.f90:113-114
        return p


# ---- the reference's namelist file (001_param.f90 namelists, read in
# 102_initial.f90:65-110) -------------------------------------------------------
_NML_TOKEN = re.compile(r"!.*|&\w+|/|'[^']*'|\"[^\"]*\"|=|[^\s,=/!'\"]+")


def _nml_value(tok: str):
    """One namelist value: logical, quoted string, integer or real (d/D
    exponents as Fortran writes double precision)."""
    if tok[0] in "'\"":
        return tok[1:-1]
    low = tok.lower()
    if low in (".true.", ".t.", "t"):
        return True
    if low in (".false.", ".f.", "f"):
        return False
    num = low.replace("d", "e")
    try:
        return int(num)
    except ValueError:
        return float(num)


def parse_namelist(text: str) -> dict:
    """{group: {variable: value or list of values}} of a Fortran namelist
    text; names in lower case; '!' starts a comment; separators are commas
    or white space."""
    groups, cur, name, vals = {}, None, None, []

    def flush():
        if cur is not None and name is not None:
            groups[cur][name] = vals[0] if len(vals) == 1 else list(vals)

    toks = [t for t in _NML_TOKEN.findall(text) if not t.startswith("!")]
    i = 0
    while i < len(toks):
        t = toks[i]
        if t.startswith("&"):
            cur, name, vals = t[1:].lower(), None, []
            groups.setdefault(cur, {})
        elif t == "/":
            flush()
            cur, name, vals = None, None, []
        elif cur is not None and i + 1 < len(toks) and toks[i + 1] == "=":
            flush()
            name, vals = t.lower(), []
            i += 1
        elif cur is not None and name is not None:
            vals.append(_nml_value(t))
        i += 1
    return groups


# namelist variable -> Namelist field (the variables the step reads, and the
# switches of the options it refuses)
_NML_FIELDS = {f.lower(): f for f in (
    "nxt", "nyt", "nzt", "lx_tot", "ly_tot", "lz_tot", "z_i", "dt", "u_scale", "cs_count", "dyn_init",
    "inilag", "restart_vel", "model", "sgs", "Co", "walltype", "lbc_mom", "ubc_mom", "ubc", "damping_method",
    "coriolis_forcing", "freq_coriolis", "ug_dim", "vg_dim", "use_mean_p_force", "mean_p_force", "nu_molec",
    "molec", "Pr", "theta_flag", "restart_theta", "theta_init_time", "passive_scalar", "T_scale", "wt_s",
    "dTdz_top", "inv_strength", "npz", "lbc", "ocean_flag", "u_star", "agl_stress", "flag_dynStress",
    "alpha_w", "flag_crossVel", "udim_cross", "vdim_cross", "stokes_flag", "stokes_action", "amp_w",
    "lambda_w", "agl_Ust", "dns_bc", "inflow", "flag_canopy", "pcon_flag", "flag_relaxForce",
    "spatial_flux_flag", "flag_dynWT")}

# The reference's defaults (001_param.f90) where they differ from the
# Namelist dataclass's: a param.nml that leaves a variable out gets these.
REFERENCE_DEFAULTS = dict(
    ubc=1, ubc_mom="wall", stokes_flag=True, stokes_action="donelan.pierson1987",
    flag_canopy=True, flag_dynStress=True, inv_strength=BOGUS, dTdz_top=BOGUS, wt_s=BOGUS,
    theta_flag=False, theta_init_time=-9999, dyn_init=-9999, amp_w=BOGUS, lambda_w=BOGUS, agl_Ust=BOGUS,
    udim_cross=BOGUS, vdim_cross=BOGUS, u_star=BOGUS, nu_molec=1.14e-5)


def read_namelist(path_or_text: str) -> Namelist:
    """Namelist from the reference's param.nml (a path, or its text): the
    variables of every group that the step reads; the others keep the
    reference's defaults (001_param.f90).  Options outside the step (canopy,
    DNS, inflow, pollen, relaxation forcing, file-driven surface flux)
    make Namelist.params() raise UnsupportedOption; physics options the
    library does not implement reach les_create, which rejects them."""
    text = path_or_text
    if "&" not in path_or_text and os.path.exists(path_or_text):
        with open(path_or_text) as f:
            text = f.read()
    nl = Namelist(**REFERENCE_DEFAULTS)
    for group in parse_namelist(text).values():
        for var, val in group.items():
            field = _NML_FIELDS.get(var)
            if field is None:
                continue
            cur = getattr(nl, field)
            if isinstance(cur, bool):
                val = bool(val)
            elif isinstance(cur, int) and not isinstance(val, bool):
                val = int(val)
            elif isinstance(cur, float) or field == "mean_p_force":
                val = float(val)
            setattr(nl, field, val)
    return nl


def workload(name: str) -> Namelist:
    """Named workloads.  sbl64/sbl128/sbl256 are the 64^3, 128^3 and 256^3
    surface boundary layers (domain 2pi x 2pi x 1 in z_i units, z0 = 1e-4);
    tiny* are small cases for tests."""
    base = dict(lx_tot=200.0 * math.pi, ly_tot=200.0 * math.pi, lz_tot=100.0, z_i=100.0,
                dt=2e-4 * 100.0, u_scale=1.0, dTdz_top=0.01 * 300.0 / 100.0)
    if name == "sbl64":
        return Namelist(nxt=64, nyt=64, nzt=65, **base)
    if name == "sbl128":
        return Namelist(nxt=128, nyt=128, nzt=129, wt_s=0.01 * 300.0, passive_scalar=True,
                        coriolis_forcing=True, freq_coriolis=1e-4, ug_dim=1.0, vg_dim=0.0, **base)
    if name == "sbl256":
        return Namelist(nxt=256, nyt=256, nzt=257, wt_s=-0.005 * 300.0, **base)
    if name == "tiny":
        return Namelist(nxt=16, nyt=16, nzt=9, cs_count=2, dyn_init=2, **base)
    if name == "tiny_ragged":
        return Namelist(nxt=24, nyt=12, nzt=7, cs_count=2, dyn_init=2, wt_s=0.01 * 300.0,
                        coriolis_forcing=True, ug_dim=1.0, vg_dim=0.2, ubc=1, **base)
    if name == "mid":
        return Namelist(nxt=64, nyt=64, nzt=9, cs_count=2, dyn_init=2, wt_s=-0.005 * 300.0, **base)
    if name == "mid_rect":
        return Namelist(nxt=128, nyt=64, nzt=7, cs_count=2, dyn_init=2, wt_s=0.01 * 300.0,
                        coriolis_forcing=True, ug_dim=1.0, **base)
    if name == "small":
        return Namelist(nxt=32, nyt=32, nzt=17, cs_count=3, dyn_init=3, wt_s=-0.005 * 300.0, **base)
    # the reference's own configurations (SBL/SBBL/BBL param.nml): the ocean
    # branch with a wind-stressed surface and a rough wall top, monochromatic
    # Stokes drift in its finite-depth form
    if name == "ocean_sbl":
        return Namelist(nxt=32, nyt=32, nzt=17, cs_count=3, dyn_init=3, ocean_flag=True, lbc_mom="stress free",
                        ubc_mom="wall", u_star=1.0, agl_stress=20.0, freq_coriolis=1e-4, ug_dim=0.0,
                        vg_dim=0.0, alpha_w=2.0e-4, stokes_flag=True, stokes_action="mono", amp_w=1.13,
                        lambda_w=60.0, agl_Ust=10.0, wt_s=0.01 * 300.0, **base)
    # the ocean bottom boundary layer at 512 x 512 x 256: the reference's BBL
    # physics (BBL/readin/param.nml: the ocean branch, no wind stress, a
    # geostrophic current of 0.25 m/s, a rough wall at the computational top,
    # the same 282.7 x 282.7 x 45 m domain, dt = 0.15 s, u_scale = 0.0122 m/s,
    # no surface heat flux) on a grid twice as fine horizontally and with 256
    # levels, rough wall z0 = 1e-3 z_i, the dynamic SGS from the start as in
    # the surface-layer workloads; one scalar (theta)
    if name == "bbl512":
        return Namelist(nxt=512, nyt=512, nzt=257, lx_tot=282.743334, ly_tot=282.743334, lz_tot=45.0,
                        z_i=45.0, dt=0.15, u_scale=0.0122, ocean_flag=True, lbc_mom="stress free",
                        ubc_mom="wall", u_star=0.0, freq_coriolis=1e-4, ug_dim=0.25, vg_dim=0.0,
                        inv_strength=0.1, dTdz_top=0.0, wt_s=0.0, zo=1e-3)
    # deep ocean: stress-free floor, deep-water Stokes drift, damping method 3,
    # cross flow
    if name == "ocean_deep":
        return Namelist(nxt=24, nyt=12, nzt=9, cs_count=2, dyn_init=2, ocean_flag=True, lbc_mom="stress free",
                        ubc_mom="stress free", u_star=1.0, agl_stress=-30.0, freq_coriolis=1e-2, ug_dim=0.3,
                        vg_dim=-0.2, alpha_w=2.0e-4, stokes_flag=True, stokes_action="mono", amp_w=0.5,
                        lambda_w=40.0, agl_Ust=-15.0, ubc=1, damping_method=3, inv_strength=1.0,
                        flag_crossVel=True, udim_cross=0.2, vdim_cross=0.1, wt_s=-0.005 * 300.0, **base)
    # land: prescribed surface temperature (lbc = 0), smooth walls at both ends
    if name == "land_lbc0":
        return Namelist(nxt=16, nyt=16, nzt=9, cs_count=2, dyn_init=2, lbc=0, walltype="smooth",
                        ubc_mom="wall", coriolis_forcing=True, ug_dim=1.0, **base)
    # land: damping method 2 under a stress-free top, with deep-water Stokes drift
    if name == "land_damp2":
        return Namelist(nxt=16, nyt=24, nzt=9, cs_count=2, dyn_init=2, ubc=1, damping_method=2,
                        stokes_flag=True, stokes_action="mono", amp_w=0.5, lambda_w=30.0, agl_Ust=40.0,
                        wt_s=0.01 * 300.0, **base)
    # prescribed surface temperature (lbc = 0) under the linearised step:
    # a cooled surface (stable obukhov branch) and a warmed one (unstable)
    if name == "tiny_lbc0":
        return Namelist(nxt=16, nyt=16, nzt=9, cs_count=2, dyn_init=2, lbc=0, theta_s1=299.0, **base)
    if name == "ragged_lbc0":
        return Namelist(nxt=24, nyt=12, nzt=7, cs_count=2, dyn_init=2, lbc=0, theta_s1=301.5,
                        coriolis_forcing=True, ug_dim=1.0, vg_dim=0.2, ubc=1, **base)
    # the 128^3 configuration with the surface temperature prescribed 0.15 K
    # above the air (a surface heat flux near 0.01 K m/s that depends on the
    # state), for the heat-flux objective of the adjoint
    if name == "sbl128_hf":
        return Namelist(nxt=128, nyt=128, nzt=129, lbc=0, theta_s1=300.15, passive_scalar=True,
                        coriolis_forcing=True, freq_coriolis=1e-4, ug_dim=1.0, vg_dim=0.0, **base)
    # many levels on a small plane: the TDMA's LDS and HBM levels both run.
    # dz is 1/50 of dx here, so the explicit vertical SGS diffusion near the
    # wall (Nu_t dt / dz^2) needs a tenth of the usual time step to stay
    # stable
    if name == "deep16":
        return Namelist(nxt=16, nyt=16, nzt=129, cs_count=2, dyn_init=2, wt_s=-0.005 * 300.0,
                        **dict(base, dt=2e-5 * 100.0))
    if name == "wide":        # production line lengths 256 / 384, few levels
        return Namelist(nxt=256, nyt=256, nzt=5, cs_count=2, dyn_init=2, wt_s=-0.005 * 300.0, **base)
    if name == "wide_rect":   # 512 / 768 along x, 128 / 192 along y
        return Namelist(nxt=512, nyt=128, nzt=5, cs_count=2, dyn_init=2, wt_s=0.01 * 300.0,
                        coriolis_forcing=True, ug_dim=1.0, **base)
    raise KeyError(name)


# --------------------------------------------------------------------------
# counter-based generator (splitmix64 finaliser over a keyed counter)
# --------------------------------------------------------------------------
_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _mix(x: np.ndarray) -> np.ndarray:
    x = x.astype(np.uint64)
    with np.errstate(over="ignore"):
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        x = x ^ (x >> np.uint64(31))
    return x


def uniform(seed: int, field_id: int, shape) -> np.ndarray:
    """U(-1, 1) keyed by (seed, field_id, global index); shape (nz, ny, nx)."""
    nz, ny, nx = shape
    k, j, i = np.meshgrid(np.arange(nz, dtype=np.uint64), np.arange(ny, dtype=np.uint64),
                          np.arange(nx, dtype=np.uint64), indexing="ij")
    with np.errstate(over="ignore"):
        key = (np.uint64(seed) * np.uint64(0x9E3779B97F4A7C15)
               + np.uint64(field_id) * np.uint64(0xD1B54A32D192ED03))
        ctr = key + ((k * np.uint64(1 << 20) + j) * np.uint64(1 << 20) + i)
    bits = _mix(_mix(ctr)) >> np.uint64(11)
    return (bits.astype(np.float64) * (1.0 / 9007199254740992.0)) * 2.0 - 1.0


def gaussian(seed: int, field_id: int, shape) -> np.ndarray:
    u1 = 0.5 * (uniform(seed, 2 * field_id + 1000, shape) + 1.0)
    u2 = 0.5 * (uniform(seed, 2 * field_id + 1001, shape) + 1.0)
    u1 = np.maximum(u1, 1e-300)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def initial_fields(nl: Namelist, seed: int = 1234, vw_noise: bool = False) -> dict:
    """Seeded synthetic initial state, levels 0..nzt (ghost levels BOGUS as
    data_exchange leaves them, 202_parallel.f90:722-733).

    u = (1/0.4) ln(z/z0) (1 + 0.1 U(-1,1)) for z < 0.5 and v = w = 0 (the
    benchmark configurations); vw_noise adds noise of the same amplitude to v
    and a tenth of it to w below z = 0.5.  theta = 1 + gamma max(0, z - 0.3)
    (+ N(0, 0.05) noise for a passive scalar).
    """
    p = nl.params()
    nx, ny, nz = p.nxt, p.nyt, p.nzt
    dz = p.lz_tot / (nz - 1)
    shape = (nz + 1, ny, nx)
    k = np.arange(nz + 1, dtype=np.float64)
    zuvp = (k - 0.5) * dz
    zw = (k - 1.0) * dz
    z0 = p.zo
    prof = np.where(zuvp > 0, (1.0 / 0.4) * np.log(np.maximum(zuvp, 1e-300) / z0), 0.0)
    low = (zuvp < 0.5).astype(np.float64)[:, None, None]
    loww = (zw < 0.5).astype(np.float64)[:, None, None]
    u = prof[:, None, None] * (1.0 + 0.1 * low * uniform(seed, 1, shape))
    v = 0.1 * prof[:, None, None] * low * uniform(seed, 2, shape)
    w = 0.1 * np.maximum(prof, 0.0)[:, None, None] * loww * uniform(seed, 3, shape) * 0.1
    if not vw_noise:
        v = np.zeros(shape)
        w = np.zeros(shape)
    gamma = p.dTdz_top / p.T_scale * p.z_i
    theta = 1.0 + gamma * np.maximum(0.0, zuvp - 0.3)[:, None, None] + 0.0 * u
    if nl.passive_scalar:
        theta = theta + 0.05 * gaussian(seed, 4, shape)
    w[1] = 0.0      # wall (110_project.f90:58-60)
    w[nz] = 0.0     # top (110_project.f90:55)
    for f in (u, v, w, theta):
        f[0] = BOGUS
    u[nz] = BOGUS
    v[nz] = BOGUS
    theta[nz] = BOGUS
    zeros = np.zeros(shape)
    return {
        "U": u, "V": v, "W": w, "THETA": theta,
        "RHSX": zeros.copy(), "RHSY": zeros.copy(), "RHSZ": zeros.copy(), "RHST": zeros.copy(),
        "CS_OPT2": zeros.copy(), "F_LM": zeros.copy(), "F_MM": zeros.copy(),
        "F_QN": zeros.copy(), "F_NN": zeros.copy(),
        "U_LAG": zeros.copy(), "V_LAG": zeros.copy(), "W_LAG": zeros.copy(),
        "SGS_T3": zeros.copy(),
    }


def load_state(ctx, fields: dict, filter_velocity: bool = True):
    """Upload an initial state and apply init's filter_data(u, v, w)
    (102_initial.f90:55-57)."""
    for name, arr in fields.items():
        ctx.set(name, arr)
    if filter_velocity:
        for f in ("U", "V", "W"):
            ctx.filter_data(f)


STATE_FIELDS = ("U", "V", "W", "THETA", "RHSX", "RHSY", "RHSZ", "RHST",
                "CS_OPT2", "F_LM", "F_MM", "F_QN", "F_NN", "U_LAG", "V_LAG", "W_LAG", "SGS_T3")


def physical_levels(name: str, nz: int):
    """Levels of a state field that carry physical values after a step."""
    if name == "W":
        return slice(1, nz + 1)
    if name in ("CS_OPT2", "F_LM", "F_MM", "F_QN", "F_NN"):
        return slice(1, nz + 1)
    if name == "SGS_T3":
        return slice(1, nz + 1)
    return slice(1, nz)
