"""Validation harness (reference ``pyjac/functional_tester``).

* :mod:`.pasr` -- a seeded partially stirred reactor that produces
  realistic, near-equilibrium thermochemical states without Cantera
  (reference ``functional_tester/partially_stirred_reactor.py``); its
  reaction sub-step is a constant-pressure reactor advanced by this package's
  GPU integrator, or by any callable with the same signature;
* :mod:`.norms` -- the reference's error norms for species rates and
  Jacobians (``functional_tester/test.py:685-778, 1131-1242``);
* :mod:`.validate` -- runs the HIP kernels over a set of states and scores
  them against answers the caller supplies (the reference takes them from
  Cantera and from an automatically differentiated Jacobian).
"""
