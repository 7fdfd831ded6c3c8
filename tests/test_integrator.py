"""The checker's ROS2 integrator (oracle_integrate_ros2, the scheme of
include/pyjac_integrator.h) against SciPy's Radau IIA solution at tight
tolerance on the checker's rates and Jacobian, and its second order: with
fixed steps the error against that solution falls by 4 per halving."""

import numpy as np
import pytest

import oracle
from pyjac.core.states import sample_states

scipy_integrate = pytest.importorskip('scipy.integrate')


def _radau(mech, conp, param, y0, t_end):
    fun = lambda t, x: oracle.species_rates(mech, conp, param[None], x[None, :])[0]
    jac = lambda t, x: oracle.jacobian(mech, conp, param[None], x[None, :])[0]
    r = scipy_integrate.solve_ivp(fun, (0.0, t_end), y0, method='Radau', jac=jac, rtol=1e-12,
                                  atol=1e-22)
    assert r.success
    return r.y[:, -1]


@pytest.mark.parametrize('name,conp,t_end', [('h2o2', True, 1e-4), ('h2o2', False, 1e-4),
                                             ('gri30', False, 2e-6)])
def test_adaptive_matches_radau(mechs, name, conp, t_end):
    mech = mechs(name)
    param, phi = sample_states(mech, 2, seed=8, conp=conp, T_range=(1200.0, 1500.0))
    rtol, atol = 1e-6, 1e-12
    y, stats = oracle.integrate_ros2(mech, conp, param, phi, t_end, rtol=rtol, atol=atol)
    assert stats[3] == 0 and stats[0] > 0
    for s in range(2):
        ref = _radau(mech, conp, param[s], phi[s], t_end)
        w = atol + rtol * np.abs(ref)
        assert np.max(np.abs(y[s] - ref) / w) <= 20.0


def test_second_order(mechs):
    """Fixed steps on H2/O2 at 1500 K over 1e-7 s (before the steps resolve
    ignition the error is not yet asymptotic): the species error falls by
    close to 4 per halving of the step."""
    mech = mechs('h2o2')
    param, phi = sample_states(mech, 1, seed=9, conp=True, T_range=(1500.0, 1500.0))
    t_end = 1e-7
    ref = _radau(mech, True, param[0], phi[0], t_end)
    errs = []
    for steps in (40, 80, 160, 320):
        y, _ = oracle.integrate_ros2(mech, True, param, phi, t_end, h0=t_end / steps, fixed=True)
        errs.append(np.max(np.abs(y[0, 2:] - ref[2:])) / np.max(np.abs(ref[2:])))
    ratios = [errs[i] / errs[i + 1] for i in range(3)]
    assert all(3.0 < r < 4.6 for r in ratios), (errs, ratios)
    assert ratios[-1] > 3.7  # approaching 4 = 2^2
