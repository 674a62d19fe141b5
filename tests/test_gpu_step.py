"""The GPU step's composition pinned by the same whole-step closed forms as
the oracle's (tests/closed_forms.py): inertial oscillation (land, and ocean
with Stokes drift), an advected wave damped by methods 2 and 3, viscous
decay, the divergence after a step, and the dynamic-SGS gating.  FP64, with
the oracle's bounds."""
import pytest

import closed_forms as cf

pytestmark = pytest.mark.gpu


@pytest.fixture
def make(gpu_lib):
    return gpu_lib.create


@pytest.mark.parametrize("ocean", [False, True])
def test_inertial_oscillation(make, ocean):
    eu, ev, ew, moved = cf.inertial_oscillation(make, ocean)
    assert moved > 1e-3
    assert eu < 1e-14 and ev < 1e-14, (eu, ev)
    assert ew < 1e-13, ew


@pytest.mark.parametrize("method", [2, 3])
def test_advected_damped_wave(make, method):
    eu, ev, ew, decay = cf.advected_wave(make, method)
    assert decay < 0.95
    assert eu < 1e-13 and ev < 1e-12, (eu, ev)
    assert ew < 1e-12, ew


def test_viscous_decay(make):
    e_rec, e_exact, ev, ew, decayed = cf.viscous_decay(make)
    assert decayed > 1e-3
    assert e_rec < 1e-12, e_rec
    assert e_exact < 3e-6, e_exact
    assert ev < 1e-13 and ew < 1e-13, (ev, ew)


def test_divergence_free_after_a_step(make):
    assert cf.divergence(make) < 1e-11


def test_sgs_gating(make):
    assert cf.sgs_gating(make) == []
