"""Error norms of the reference's validation harness.

``JacobianErrors`` restates ``functional_tester/test.py:1131-1242``
(``jacobian_eval.eval_error``) and the threshold of :1040-1092;
``RateErrors`` restates :685-778 (``spec_rate_eval.eval_error``) for
outputs compared element by element.  Both accumulate over chunks of
states the way the reference does: norms combine as
``sqrt(a^2 + b^2)``, maxima as maxima.
"""

import numpy as np

INF_CUTOFF = 1e285
"""Values beyond this are taken as infinite (reference utils.py:280)."""


def _finite_mask(a):
    with np.errstate(invalid='ignore'):
        return np.isfinite(a) & (a < INF_CUTOFF) & (a > -INF_CUTOFF)


def jacobian_threshold(answers):
    """Frobenius norm of all reference Jacobians together, effectively
    infinite entries left out (test.py:1071-1092).  ``answers``: an iterable
    of arrays (chunks)."""
    total = 0.0
    for a in answers:
        a = np.asarray(a, dtype=np.float64)
        total += np.linalg.norm(a[_finite_mask(a)]) ** 2
    return float(np.sqrt(total))


class JacobianErrors(object):
    """Accumulated Jacobian error norms (keys as in the reference):
    ``jac`` (relative Frobenius over non-zero reference entries), ``jac_zero``
    (absolute, over zero entries), ``jac_lapack`` (||err|| / ||ref||),
    ``jac_thresholded_{20,15}`` (relative, over entries whose magnitude
    exceeds threshold / 1e20 or 1e15) with the values at their maxima,
    ``jac_weighted_{atol}`` for atol = atol0 x {1, 10, 100, 1000}
    (||err / (atol + rtol |ref|)||) with the values at their maxima,
    ``jac_max_value`` and ``jac_threshold_value``."""

    def __init__(self, rtol=1e-6, atol=1e-2):
        # the reference's Jacobian tester defaults (test.py:855)
        self.rtol = rtol
        self.atol = atol
        self.errors = {}

    def _put(self, key, value, op='norm'):
        d = self.errors
        if key in d:
            if op == 'norm':
                value = np.sqrt(d[key] * d[key] + value * value)
            else:
                value = np.maximum(d[key], value)
        d[key] = value

    def update(self, out, ans, threshold):
        """Add one chunk: ``out`` the kernel's Jacobians, ``ans`` the
        reference's, in the same layout."""
        out = np.asarray(out, dtype=np.float64).reshape(-1)
        ans = np.asarray(ans, dtype=np.float64).reshape(-1)
        if not np.all(np.isfinite(out)):
            raise AssertionError('NaN or Inf in the Jacobian')
        good = _finite_mask(out)
        if not np.all(good):
            # effectively infinite outputs must sit where the reference is
            # infinite or huge too, and nowhere else may the reference be
            bad_ans = ~_finite_mask(ans)
            if not np.all(bad_ans[~good]):
                raise AssertionError('effectively infinite Jacobian entries where the reference '
                                     'is finite')
            if np.any(bad_ans[good]):
                raise AssertionError('infinite reference entries where the Jacobian is finite')
            out, ans = out[good], ans[good]
        err = np.abs(out - ans)
        mag = np.abs(out)
        denom = np.abs(ans)
        nz = denom > 0
        self._put('jac', np.linalg.norm(err[nz] / denom[nz]))
        self._put('jac_zero', np.linalg.norm(err[~nz]))
        self._put('jac_lapack', np.linalg.norm(err) / np.linalg.norm(denom))
        for exp in (20, 15):
            sel = mag > threshold / 10.0 ** exp
            with np.errstate(divide='ignore', invalid='ignore'):
                rel = err[sel] / denom[sel]
            key = 'jac_thresholded_{}'.format(exp)
            self._put(key, np.linalg.norm(rel))
            if rel.size:
                k = int(np.argmax(rel))
                self._put(key + '_PJ_amax', mag[sel][k], op='max')
                self._put(key + '_AD_amax', denom[sel][k], op='max')
        scaled = self.rtol * denom
        for mul in (1, 10, 100, 1000):
            atol = self.atol * mul
            w = err / (atol + scaled)
            k = int(np.argmax(w))
            key = 'jac_weighted_{}'.format(atol)
            self._put(key, np.linalg.norm(w))
            self._put(key + '_PJ_amax', mag[k], op='max')
            self._put(key + '_AD_amax', scaled[k] / self.rtol, op='max')
        self._put('jac_max_value', np.amax(mag), op='max')
        self._put('jac_threshold_value', threshold, op='max')
        return self.errors


class RateErrors(object):
    """Per-output-entry maximum weighted errors over states (test.py:685-778):
    for each named output ``[n_states, n_values]``, the state with the
    largest ``|out - ans| / (atol + rtol |ans|)`` for each value, and there
    the absolute error (``name``), the weighted error (``name_store``) and
    the reference value (``name_value``); later chunks replace a value's
    record where their weighted error is at least as large."""

    def __init__(self, rtol=1e-6, atol=1e-10):
        # the reference's species-rate tester defaults (test.py:516)
        self.rtol = rtol
        self.atol = atol
        self.errors = {}

    def update(self, name, out, ans):
        out = np.asarray(out, dtype=np.float64)
        ans = np.asarray(ans, dtype=np.float64)
        out = out.reshape(out.shape[0], -1)
        ans = ans.reshape(ans.shape[0], -1)
        err = np.abs(out - ans)
        cmp = err / (self.atol + self.rtol * np.abs(ans))
        loc = np.argmax(cmp, axis=0)
        cols = np.arange(out.shape[1])
        store, abs_err, value = cmp[loc, cols], err[loc, cols], ans[loc, cols]
        d = self.errors
        if name not in d:
            d[name] = np.zeros_like(abs_err)
            d[name + '_store'] = np.zeros_like(abs_err)
            d[name + '_value'] = np.zeros_like(abs_err)
        upd = store >= d[name + '_store']
        d[name][upd] = abs_err[upd]
        d[name + '_store'][upd] = store[upd]
        d[name + '_value'][upd] = value[upd]
        return d
