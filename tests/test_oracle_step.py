"""The oracle's whole-step composition (orc_step, the restatement of
flosol, 103_flosol.f90:23-229) pinned by closed forms that do not come from
the restatement: see tests/closed_forms.py for each flow and what it fixes.
Run on the CPU."""
import pytest

import closed_forms as cf


@pytest.fixture
def make(oracle_mod):
    return oracle_mod.create


@pytest.mark.parametrize("ocean", [False, True])
def test_inertial_oscillation(make, ocean):
    eu, ev, ew, moved = cf.inertial_oscillation(make, ocean)
    assert moved > 1e-3                  # the state did move
    assert eu < 1e-14 and ev < 1e-14, (eu, ev)
    assert ew < 1e-13, ew


@pytest.mark.parametrize("method", [2, 3])
def test_advected_damped_wave(make, method):
    eu, ev, ew, decay = cf.advected_wave(make, method)
    assert decay < 0.95                  # the damping acted
    assert eu < 1e-13 and ev < 1e-12, (eu, ev)
    assert ew < 1e-12, ew


def test_viscous_decay(make):
    e_rec, e_exact, ev, ew, decayed = cf.viscous_decay(make)
    assert decayed > 1e-3
    assert e_rec < 1e-12, e_rec
    # against exp(-nu k^2 t): the Euler first step's error, (nu k^2 dt)^2 / 2 = 1.6e-6
    assert e_exact < 3e-6, e_exact
    assert ev < 1e-13 and ew < 1e-13, (ev, ew)


def test_divergence_free_after_a_step(make):
    assert cf.divergence(make) < 1e-11


def test_sgs_gating(make):
    assert cf.sgs_gating(make) == []
