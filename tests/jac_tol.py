"""Jacobian parity bounds of the GPU tests (against the checker's FP64
analytical Jacobian on the same states), per state:

* pyJac's relative Frobenius error (functional_tester/test.py:1131-1242,
  the LAPACK norm ``||err|| / ||ref||``) at most ``FRO = 1e-8``;
* element-wise ``|got - ref| <= RTOL |ref| + ATOL_REL max_state |ref|`` with
  ``RTOL = 1e-9`` and ``ATOL_REL = 1e-14``: an entry formed by cancelling
  contributions as large as the state's largest entry carries a few ulps of
  that entry as its rounding (the second term, ~45 ulps), and every entry is
  held to 1e-9 of itself on top (the block kernels' summation order of LDS
  atomics varies by up to 9e-10 of the entry on cancelling entries).  On the
  CPU, the checker's analytical Jacobian meets a bound ten times tighter
  against its own automatic derivative on seeded states of every bundled
  mechanism (worst 1e-2 of the 1e-10 / 1e-15 bound).
"""

import numpy as np

FRO, RTOL, ATOL_REL = 1e-8, 1e-9, 1e-14


def per_state(jac, n, order):
    """[n, entries] view of an output in the kernel's layout."""
    return jac.reshape(n, -1) if order == 'C' else jac.reshape(-1, n).T


def jacobian_errors(got, ref):
    """(largest relative Frobenius error, largest element-wise error as a
    fraction of its bound, number of entries above 1e-10 relative) over the
    states in the rows of got and ref."""
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    diff = np.abs(got - ref)
    mx = np.abs(ref).max(axis=1, keepdims=True)
    # norms of the entries scaled by the state's largest one (entries near
    # the top of the double range would overflow the sums of squares)
    sc = np.where(mx > 0, mx, 1.0)
    nrm = np.linalg.norm(ref / sc, axis=1)
    fro = np.linalg.norm(diff / sc, axis=1) / np.where(nrm > 0, nrm, 1.0)
    elem = diff / (RTOL * np.abs(ref) + ATOL_REL * mx + 1e-300)
    over = int(np.count_nonzero(diff > 1e-10 * np.abs(ref)))
    return float(fro.max()), float(elem.max()), over


def assert_jacobian_close(got, ref, what=''):
    assert np.isfinite(got).all(), 'non-finite Jacobian entries ' + what
    fro, elem, over = jacobian_errors(got, ref)
    print('jacobian {}: frobenius {:.3g}, element-wise {:.3g} of bound, {} entries '
          'above 1e-10 relative'.format(what, fro, elem, over))
    assert fro <= FRO, 'Frobenius error {:.3g} {}'.format(fro, what)
    assert elem <= 1.0, 'element-wise error {:.3g} x bound {}'.format(elem, what)
