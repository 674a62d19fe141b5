"""Pin the oracle's transforms against library routines, brute force and
closed forms (the reference ships no fixtures for this path).

fftn_mpi / ifftn_mpi   203_fftn_mpi.f90:4-67
ddx / ddy              004_derivatives.f90:18-81
filter_data            006_test_filter.f90:82-105
test_filter            006_test_filter.f90:19-77
zero_padding / unpad   108_01_padd.f90:73-197
"""
import math

import numpy as np
import pytest

from overlappingbls_amd import config


def mk(oracle_mod, nx=16, ny=12, nzt=6, **kw):
    nl = config.Namelist(nxt=nx, nyt=ny, nzt=nzt, lx_tot=2 * math.pi * 100, ly_tot=1.3 * 2 * math.pi * 100,
                         lz_tot=100.0, z_i=100.0, **kw)
    return nl, oracle_mod.create(nl.params())


def brute_dft2(f):
    """Long-double direct 2-D DFT, exp(-2 pi i (kx x/nx + ky y/ny))."""
    ny, nx = f.shape
    X = np.zeros((ny, nx), dtype=np.clongdouble)
    y = np.arange(ny, dtype=np.longdouble)
    x = np.arange(nx, dtype=np.longdouble)
    two_pi = np.longdouble(2) * np.longdouble(np.pi)
    Ey = np.exp(-1j * two_pi * np.outer(y, y) / ny)
    Ex = np.exp(-1j * two_pi * np.outer(x, x) / nx)
    X = Ey @ f.astype(np.longdouble) @ Ex
    return X


@pytest.mark.parametrize("nx,ny", [(16, 12), (24, 18), (12, 48), (32, 32)])
def test_fftn_matches_bruteforce_and_numpy(oracle_mod, nx, ny):
    nl, o = mk(oracle_mod, nx, ny)
    f = config.uniform(3, 1, (1, ny, nx))[0]
    got = o.fftn(f)
    ref = np.fft.fft2(f)[:, : nx // 2 + 1]
    ref[:, nx // 2] = 0          # in_hat_x(lhx,:) = 0
    ref[ny // 2, :] = 0          # out(:, nyt/2+1) = 0
    assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max()
    bf = brute_dft2(f)[:, : nx // 2 + 1].astype(np.complex128)
    bf[:, nx // 2] = 0
    bf[ny // 2, :] = 0
    assert np.abs(got - bf).max() < 1e-12 * np.abs(bf).max()


@pytest.mark.parametrize("nx,ny", [(16, 12), (36, 24)])
def test_ifftn_c2r_semantics(oracle_mod, nx, ny):
    """ifftn: y backward c2c, then x c2r ignoring Im of kx=0 and kx=nx/2."""
    nl, o = mk(oracle_mod, nx, ny)
    rng = np.random.default_rng(1)
    s = rng.standard_normal((ny, nx // 2 + 1)) + 1j * rng.standard_normal((ny, nx // 2 + 1))
    got = o.ifftn(s)
    t = np.fft.ifft(s, axis=0) * ny          # unnormalised backward along y
    t[:, 0] = t[:, 0].real
    t[:, nx // 2] = t[:, nx // 2].real
    full = np.zeros((ny, nx), dtype=complex)
    full[:, : nx // 2 + 1] = t
    full[:, nx // 2 + 1:] = np.conj(t[:, 1: nx // 2][:, ::-1])
    ref = np.fft.ifft(full, axis=1).real * nx
    assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max()


def modes(nx, ny, lx, ly, terms):
    """sum of a cos(2 pi (mx x/lx + my y/ly) + ph) on the grid, with its
    x and y derivatives."""
    x = np.arange(nx) * lx / nx
    y = np.arange(ny) * ly / ny
    X, Y = np.meshgrid(x, y)
    f = np.zeros((ny, nx)); fx = np.zeros_like(f); fy = np.zeros_like(f)
    for a, mx, my, ph in terms:
        arg = 2 * np.pi * (mx * X / lx + my * Y / ly) + ph
        f += a * np.cos(arg)
        fx += -a * np.sin(arg) * 2 * np.pi * mx / lx
        fy += -a * np.sin(arg) * 2 * np.pi * my / ly
    return f, fx, fy


def test_ddx_ddy_closed_form(oracle_mod):
    nx, ny = 24, 16
    nl, o = mk(oracle_mod, nx, ny)
    P = nl.params()
    terms = [(1.0, 1, 0, 0.3), (0.5, 3, 2, 1.1), (0.25, 11, -7, 0.0), (0.7, 0, 5, 0.2)]
    f, fx, fy = modes(nx, ny, P.lx_tot, P.ly_tot, terms)
    nyq = modes(nx, ny, P.lx_tot, P.ly_tot, [(2.0, nx // 2, 0, 0.0), (1.5, 0, ny // 2, 0.0), (0.5, nx // 2, ny // 2, 0.0)])[0]
    a = np.repeat((f + nyq)[None], o.nlev, axis=0)
    o.set("WORK0", a)
    o.ddx("WORK1", "WORK0")
    o.ddy("WORK2", "WORK0")
    # the Nyquist modes differentiate to zero (kx(lhx) = 0, ky(lhy) = 0)
    assert np.abs(o.get("WORK1")[3] - fx).max() < 1e-12 * np.abs(fx).max()
    assert np.abs(o.get("WORK2")[3] - fy).max() < 1e-12 * np.abs(fy).max()


def test_filter_data_removes_nyquist_and_is_idempotent(oracle_mod):
    nx, ny = 16, 12
    nl, o = mk(oracle_mod, nx, ny)
    P = nl.params()
    f = modes(nx, ny, P.lx_tot, P.ly_tot, [(1.0, 2, 1, 0.4), (0.3, 7, -5, 0.0)])[0]
    nyq = modes(nx, ny, P.lx_tot, P.ly_tot, [(2.0, nx // 2, 3, 0.0), (1.5, 1, ny // 2, 0.0)])[0]
    o.set("WORK0", np.repeat((f + nyq)[None], o.nlev, axis=0))
    o.filter_data("WORK0")
    g1 = o.get("WORK0")
    assert np.abs(g1[2] - f).max() < 1e-13
    o.filter_data("WORK0")
    assert np.abs(o.get("WORK0") - g1).max() < 1e-13


@pytest.mark.parametrize("which", [2, 4])
def test_test_filter_cutoff(oracle_mod, which):
    """G = 1/(nxt nyt) where k^2 < (pi/(alpha sqrt(dx dy)))^2 (006:56-61)."""
    nx, ny = 32, 32
    nl, o = mk(oracle_mod, nx, ny)
    P = nl.params()
    dx, dy = P.lx_tot / nx, P.ly_tot / ny
    kc = math.pi / (which * math.sqrt(dx * dy))
    keep, drop = [], []
    for mx in range(0, nx // 2):
        for my in range(-(ny // 2) + 1, ny // 2):
            k = math.hypot(2 * math.pi * mx / P.lx_tot, 2 * math.pi * my / P.ly_tot)
            if abs(k - kc) < 0.05 * kc or (mx, my) == (0, 0):
                continue
            (keep if k < kc else drop).append((mx, my))
    terms_k = [(1.0 / (1 + i), mx, my, 0.1 * i) for i, (mx, my) in enumerate(keep[:6])]
    terms_d = [(1.0 / (1 + i), mx, my, 0.2 * i) for i, (mx, my) in enumerate(drop[:6])]
    fk = modes(nx, ny, P.lx_tot, P.ly_tot, terms_k)[0] + 0.7
    fd = modes(nx, ny, P.lx_tot, P.ly_tot, terms_d)[0]
    o.set("WORK0", np.repeat((fk + fd)[None], o.nlev, axis=0))
    o.test_filter("WORK0", 1, which)
    g = o.get("WORK0", 1, 1)[0]
    assert np.abs(g - fk).max() < 1e-12 * np.abs(fk).max()
    o.test_filter("WORK0", 1, which)
    assert np.abs(o.get("WORK0", 1, 1)[0] - g).max() < 1e-13


@pytest.mark.parametrize("nx,ny", [(16, 12), (24, 32)])
def test_zero_padding_interpolates_band_limited_fields(oracle_mod, nx, ny):
    nl, o = mk(oracle_mod, nx, ny)
    P = nl.params()
    terms = [(1.0, 0, 0, 0.0), (0.8, 3, -2, 0.5), (0.4, nx // 2 - 1, 1, 0.3), (0.2, 1, ny // 2 - 1, 1.0)]
    f = modes(nx, ny, P.lx_tot, P.ly_tot, terms)[0]
    big = o.zero_padding(f)
    ref = modes(3 * nx // 2, 3 * ny // 2, P.lx_tot, P.ly_tot, terms)[0]
    assert np.abs(big - ref).max() < 1e-12
    # the Nyquist mode is dropped by the padding (keeps 1..nxt/2)
    nyq = modes(nx, ny, P.lx_tot, P.ly_tot, [(1.0, nx // 2, 0, 0.0)])[0]
    assert np.abs(o.zero_padding(f + nyq) - ref).max() < 1e-12
    assert np.abs(o.zero_unpadding(big) - f).max() < 1e-12


def test_dealiased_product_is_truncated_exact_product(oracle_mod):
    """unpad(pad(a) pad(b)) keeps |m| < n/2 of the exact product a*b."""
    nx, ny = 24, 24
    nl, o = mk(oracle_mod, nx, ny)
    P = nl.params()
    a = modes(nx, ny, P.lx_tot, P.ly_tot, [(1.0, 5, 3, 0.2), (0.5, 9, -4, 0.0)])[0]
    b = modes(nx, ny, P.lx_tot, P.ly_tot, [(1.0, 4, -6, 0.1), (0.3, 2, 8, 0.7)])[0]
    prod = o.zero_unpadding(o.zero_padding(a) * o.zero_padding(b))
    # analytic product: cos A cos B = (cos(A+B) + cos(A-B))/2, keep |m| < n/2
    terms = []
    for (aa, mx1, my1, p1) in [(1.0, 5, 3, 0.2), (0.5, 9, -4, 0.0)]:
        for (bb, mx2, my2, p2) in [(1.0, 4, -6, 0.1), (0.3, 2, 8, 0.7)]:
            for s in (1, -1):
                mx, my, ph = mx1 + s * mx2, my1 + s * my2, p1 + s * p2
                if abs(mx) < nx // 2 and abs(my) < ny // 2:
                    terms.append((0.5 * aa * bb, mx, my, ph))
    ref = modes(nx, ny, P.lx_tot, P.ly_tot, terms)[0]
    assert np.abs(prod - ref).max() < 1e-12
