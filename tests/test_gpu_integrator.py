"""The batched ROS2 integrator (include/pyjac_integrator.h) on the GPU
against the checker's restatement of the same scheme (pinned by
tests/test_integrator.py against SciPy's Radau and by its order).  With fixed
steps both follow the same arithmetic and agree closely; with adaptive steps
each state's own step sequence may differ where an error estimate sits at the
acceptance threshold, so both are held to the tolerance."""

import numpy as np
import pytest

import oracle
from pyjac.core.states import sample_states
from pyjac.pywrap import PyIntegrator

pytestmark = pytest.mark.gpu
torch = pytest.importorskip('torch')


def _run(mech, conp, param, phi, **kw):
    kw.setdefault('max_steps', 100000)
    it = PyIntegrator(mech, conp)
    p = torch.from_numpy(param).cuda()
    y = torch.from_numpy(phi.copy()).cuda()
    stats = it.integrate(p, y, **kw)
    torch.cuda.synchronize()
    return y.cpu().numpy(), stats


@pytest.mark.parametrize('conp', [True, False])
def test_fixed_steps_match_checker(mechs, conp):
    """Fixed steps on realistic states (the reference's PaSR states in
    GRI-Mech 3.0; random compositions are radical-rich and ignite within
    nanoseconds, beyond any fixed step)."""
    from test_pasr_states import pasr_states
    mech = mechs('gri30')
    param, phi = pasr_states(mech, conp)
    idx = np.linspace(0, len(param) - 1, 48).astype(int)
    param, phi = param[idx], phi[idx]
    n = len(param)
    t_end, steps = 1e-5, 100
    got, stats = _run(mech, conp, param, phi, t_end=t_end, h0=t_end / steps, fixed=True,
                      max_steps=steps + 5)
    ref, rstats = oracle.integrate_ros2(mech, conp, param, phi, t_end, h0=t_end / steps,
                                        fixed=True, max_steps=steps + 5)
    assert rstats[3] == 0 and rstats[0] == n * steps
    assert stats['failed'] == 0 and stats['accepted'] == n * steps
    scale_sp = np.abs(ref[:, 2:]).max(axis=1, keepdims=True)
    assert np.max(np.abs(got[:, :2] - ref[:, :2]) / np.abs(ref[:, :2])) <= 1e-9
    assert np.max(np.abs(got[:, 2:] - ref[:, 2:]) / scale_sp) <= 1e-9


@pytest.mark.parametrize('name,conp,t_end', [('h2o2', True, 1e-5), ('gri30', False, 1e-6)])
def test_adaptive_matches_checker(mechs, name, conp, t_end):
    mech = mechs(name)
    n = 32
    param, phi = sample_states(mech, n, seed=13, conp=conp, T_range=(1200.0, 1500.0))
    rtol, atol = 1e-6, 1e-12
    got, stats = _run(mech, conp, param, phi, t_end=t_end, rtol=rtol, atol=atol)
    ref, rstats = oracle.integrate_ros2(mech, conp, param, phi, t_end, rtol=rtol, atol=atol)
    assert stats['failed'] == 0 and rstats[3] == 0
    # within the tolerance of each other (both are within it of the solution)
    w = atol + rtol * np.abs(ref)
    assert np.max(np.abs(got - ref) / w) <= 50.0
    # and the same amount of work, within a few percent
    assert abs(stats['accepted'] - rstats[0]) <= 0.05 * rstats[0] + n


@pytest.mark.parametrize('conp', [True, False])
def test_device_memory_lu_matches_checker(mechs, monkeypatch, conp):
    """The iteration matrix factorised in device memory (the path of the
    mechanisms beyond 120 species), forced on GRI-Mech 3.0's PaSR states:
    the same 1e-9 agreement with the checker after 100 fixed steps as the
    LDS path (the two paths are not compared bit for bit: the Jacobian's
    LDS atomics already vary in the last bits from call to call)."""
    from test_pasr_states import pasr_states
    monkeypatch.setenv('PYJAC_HIP_INTEGRATOR_DEVICE_LU', '1')
    mech = mechs('gri30')
    param, phi = pasr_states(mech, conp)
    idx = np.linspace(0, len(param) - 1, 32).astype(int)
    param, phi = param[idx], phi[idx]
    n = len(param)
    t_end, steps = 1e-5, 100
    got, stats = _run(mech, conp, param, phi, t_end=t_end, h0=t_end / steps, fixed=True,
                      max_steps=steps + 5)
    ref, rstats = oracle.integrate_ros2(mech, conp, param, phi, t_end, h0=t_end / steps,
                                        fixed=True, max_steps=steps + 5)
    assert stats['failed'] == 0 and stats['accepted'] == n * steps == rstats[0]
    scale_sp = np.abs(ref[:, 2:]).max(axis=1, keepdims=True)
    assert np.max(np.abs(got[:, :2] - ref[:, :2]) / np.abs(ref[:, :2])) <= 1e-9
    assert np.max(np.abs(got[:, 2:] - ref[:, 2:]) / scale_sp) <= 1e-9


def _diff(a, ref):
    scale_sp = np.abs(ref[:, 2:]).max(axis=1, keepdims=True)
    return (np.max(np.abs(a[:, :2] - ref[:, :2]) / np.abs(ref[:, :2])),
            np.max(np.abs(a[:, 2:] - ref[:, 2:]) / scale_sp))


@pytest.mark.parametrize('conp', [True, False])
def test_large_mechanism_fixed_steps(mechs, conp):
    """The 150-species synthetic mechanism (NN = 151, beyond the LDS path):
    20 fixed steps of 1e-15 s (its random kinetics move T by 10-20 K in
    2e-14 s) against the checker's integrator.  These trajectories magnify
    round-off about a million-fold: the checker itself moves by ~5e-7 when
    its inputs are perturbed by 1e-13, so the GPU is held to ten times that
    measured sensitivity."""
    mech = mechs('s150')
    param, phi = sample_states(mech, 16, seed=31, conp=conp, T_range=(1200.0, 1500.0))
    h = 1e-15
    kw = dict(h0=h, fixed=True, max_steps=25)
    got, stats = _run(mech, conp, param, phi, t_end=20 * h, **kw)
    ref, rstats = oracle.integrate_ros2(mech, conp, param, phi, 20 * h, **kw)
    rng = np.random.default_rng(1)
    phi2 = phi * (1 + 1e-13 * rng.choice([-1, 1], size=phi.shape))
    alt, _ = oracle.integrate_ros2(mech, conp, param, phi2, 20 * h, **kw)
    assert rstats[3] == 0 and stats['failed'] == 0 and stats['accepted'] == rstats[0]
    (e_T, e_sp), (s_T, s_sp) = _diff(got, ref), _diff(alt, ref)
    print('s150 fixed steps: GPU T/E {:.3g}, species {:.3g}; checker under a 1e-13 input '
          'perturbation T/E {:.3g}, species {:.3g}'.format(e_T, e_sp, s_T, s_sp))
    assert e_T <= 10 * max(s_T, 1e-12) and e_sp <= 10 * max(s_sp, 1e-12)
