"""The condition-aware rate tolerance (tests/rate_tol.py) on the CPU: its
scales bound the rates they belong to (a sum never exceeds the sum of its
absolute terms, so a wrong index or a dropped term fails), it equals the
plain relative bound away from cancellation, and it relaxes the bound only
on a small fraction of elements."""

import numpy as np
import pytest

from pyjac.core.states import sample_states
from rate_tol import COND, RTOL, rate_error, rate_scales


@pytest.mark.parametrize('name,conp', [('h2o2', True), ('gri30', False), ('test', True)])
def test_scales_bound_the_rates(mechs, name, conp):
    mech = mechs(name)
    param, phi = sample_states(mech, 300, seed=11, conp=conp)
    ref, scale = rate_scales(mech, conp, param, phi)
    assert np.all(scale >= 0)
    # |V wdot_k| <= V sum_r |nu_kr| |rop_r|, |dT/dt| <= sum_k |e_k wdot_k| / |sum c C|
    slack = 1e-12 * scale + 1e-300
    assert np.all(np.abs(ref[:, 2:]) <= scale[:, 2:] * (1 + 1e-12) + slack[:, 2:])
    assert np.all(np.abs(ref[:, 0]) <= scale[:, 0] * (1 + 1e-12))
    assert np.all(np.abs(ref[:, 1]) <= scale[:, 1] * (1 + 1e-12))
    # the floor only matters where the terms cancel by more than 1e-3
    ratio = COND * scale / (RTOL * np.abs(ref) + 1e-300)
    assert np.mean(ratio > 1.0) < 0.01


def test_rate_error_flags_a_planted_mistake(mechs):
    mech = mechs('gri30')
    param, phi = sample_states(mech, 50, seed=12, conp=True)
    ref, scale = rate_scales(mech, True, param, phi)
    assert rate_error(ref, ref, scale) == 0.0
    bad = ref.copy()
    k = np.argmax(np.abs(ref[:, 5]))
    bad[k, 5] *= 1 + 1e-8          # a dropped small term in one species rate
    assert rate_error(bad, ref, scale) > 1.0
