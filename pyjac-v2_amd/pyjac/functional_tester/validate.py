"""Score the HIP kernels against reference answers over a set of states
(the driver part of reference ``functional_tester/test.py``:
``species_rate_tester`` :1307, ``jacobian_tester`` :1339 and ``_run_mechanism_tests``).

The reference takes its answers from Cantera (species rates) and from an
automatically differentiated Jacobian; here the caller passes them as a
function ``answer(param, phi) -> array`` of a chunk of states, in the
kernel's own output layout (``order``, ``jac_format``), so any trusted
source can be used.
"""

import numpy as np

from .norms import JacobianErrors, RateErrors, jacobian_threshold

__all__ = ['validate_species_rates', 'validate_jacobian']


def _chunks(n, chunk):
    for s in range(0, n, chunk):
        yield s, min(n, s + chunk)


def _c_order(phi, NN):
    return np.ascontiguousarray(np.asarray(phi, dtype=np.float64).reshape(-1, NN))


def validate_species_rates(mech, param, phi, answer, conp=True, chunk=65536, rtol=1e-6,
                           atol=1e-10):
    """Error record (``RateErrors``) of the species-rate kernel (``dphi``)
    over the states ``(param, phi)`` (phi C order ``[n, NN]``)."""
    from ..pywrap import PySpeciesRatesKernel
    phi = _c_order(phi, mech.NN)
    param = np.ascontiguousarray(param, dtype=np.float64).reshape(-1)
    n = phi.shape[0]
    k = PySpeciesRatesKernel(min(n, chunk), 0, mech=mech, conp=conp, order='C')
    rec = RateErrors(rtol=rtol, atol=atol)
    for s, e in _chunks(n, chunk):
        out = np.zeros((e - s, mech.NN))
        k(param[s:e], phi[s:e], out)
        rec.update('dphi', out, np.asarray(answer(param[s:e], phi[s:e])).reshape(e - s, -1))
    return rec.errors


def validate_jacobian(mech, param, phi, answer, conp=True, order='C', jac_format=0, jac_type=0,
                      chunk=16384, rtol=1e-6, atol=1e-2):
    """Error record (``JacobianErrors``) of the Jacobian kernel over the
    states ``(param, phi)`` (phi C order ``[n, NN]``; the kernel is run in
    ``order``).  ``answer`` returns a chunk's reference Jacobians in the
    kernel's layout; the threshold is the Frobenius norm of all of them."""
    from ..pywrap import PyJacobianKernel
    phi = _c_order(phi, mech.NN)
    param = np.ascontiguousarray(param, dtype=np.float64).reshape(-1)
    n = phi.shape[0]
    answers = []
    for s, e in _chunks(n, chunk):
        answers.append(np.asarray(answer(param[s:e], phi[s:e]), dtype=np.float64))
    threshold = jacobian_threshold(answers)
    k = PyJacobianKernel(min(n, chunk), 0, mech=mech, conp=conp, order=order,
                         jac_format=jac_format, jac_type=jac_type)
    rec = JacobianErrors(rtol=rtol, atol=atol)
    for (s, e), ans in zip(_chunks(n, chunk), answers):
        ph = phi[s:e] if order == 'C' else np.ascontiguousarray(phi[s:e].T)
        out = np.zeros(ans.size)
        k(param[s:e], ph, out)
        rec.update(out, ans, threshold)
    return rec.errors
