"""CPU checker: ctypes binding of ``liboracle.so`` (built from
``pyjac_oracle.cpp`` by ``oracle/Makefile``).

Test infrastructure.  Only the tests, the smoke check and the CPU-baseline
leg of ``bench.py`` use this package; the product library never does.

All arrays are per-state C order: ``phi``/``dphi`` ``[n, NN]``, ``param``
``[n]``, dense Jacobians ``[n, NN, NN]`` indexed ``[state, row, col]``.
"""

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'liboracle.so')

_lib = None


def build(force=False):
    """Compile liboracle.so with make (g++, OpenMP, strict IEEE)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.check_call(['make', '-s', '-B' if force else '-s', '-C', HERE])
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        L.oracle_species_rates.argtypes = [P, ctypes.c_int, ctypes.c_long, P, P, P, ctypes.c_int]
        L.oracle_species_rates_full.argtypes = [P, ctypes.c_int, ctypes.c_long, P, P, P, P, P, P, P]
        L.oracle_species_rates_ld.argtypes = [P, ctypes.c_int, ctypes.c_long, P, P, P, P,
                                              ctypes.c_int]
        L.oracle_chem_utils.argtypes = [P, ctypes.c_int, ctypes.c_long, P, P, P, P]
        L.oracle_jacobian.argtypes = [P, ctypes.c_int, ctypes.c_long, P, P, P, ctypes.c_int]
        L.oracle_jacobian_ad.argtypes = [P, ctypes.c_int, ctypes.c_long, P, P, P]
        L.oracle_jacobian_masked.argtypes = [P, ctypes.c_int, ctypes.c_long, P, P, P, P]
        L.oracle_integrate_ros2.argtypes = [P, ctypes.c_int, ctypes.c_long, P, P, ctypes.c_double,
                                            ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                            ctypes.c_int, ctypes.c_long, P]
        L.oracle_fd_jacobian.argtypes = [P, ctypes.c_int, ctypes.c_long, P, P, P, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_double, ctypes.c_double]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _prep(mech, param, phi):
    phi = np.ascontiguousarray(phi, dtype=np.float64).reshape(-1, mech.NN)
    param = np.ascontiguousarray(param, dtype=np.float64).reshape(-1)
    assert param.shape[0] == phi.shape[0]
    return param, phi


def species_rates(mech, conp, param, phi, nthreads=0):
    param, phi = _prep(mech, param, phi)
    out = np.empty_like(phi)
    st = lib().oracle_species_rates(ctypes.cast(mech.pointer(), ctypes.c_void_p), int(conp),
                                    phi.shape[0], _p(param), _p(phi), _p(out), int(nthreads))
    assert st == 0, st
    return out


def species_rates_ld(mech, conp, param, phi, nthreads=0):
    """The species rates evaluated in x87 extended precision from the same
    binary64 inputs: (hi, lo) with hi + lo the extended result, ``[n, NN]``
    each."""
    param, phi = _prep(mech, param, phi)
    hi = np.empty_like(phi)
    lo = np.empty_like(phi)
    st = lib().oracle_species_rates_ld(ctypes.cast(mech.pointer(), ctypes.c_void_p), int(conp),
                                       phi.shape[0], _p(param), _p(phi), _p(hi), _p(lo),
                                       int(nthreads))
    assert st == 0, st
    return hi, lo


def species_rates_full(mech, conp, param, phi):
    """Rates plus rop_fwd, rop_rev, pres_mod, rop_net, each ``[n, Nr]``."""
    param, phi = _prep(mech, param, phi)
    n = phi.shape[0]
    out = np.empty_like(phi)
    extra = [np.empty((n, mech.Nr)) for _ in range(4)]
    st = lib().oracle_species_rates_full(ctypes.cast(mech.pointer(), ctypes.c_void_p), int(conp), n,
                                         _p(param), _p(phi), _p(out), *[_p(e) for e in extra])
    assert st == 0, st
    return (out,) + tuple(extra)


def chem_utils(mech, conp, phi):
    phi = np.ascontiguousarray(phi, dtype=np.float64).reshape(-1, mech.NN)
    n = phi.shape[0]
    outs = [np.empty((n, mech.Ns)) for _ in range(3)]
    st = lib().oracle_chem_utils(ctypes.cast(mech.pointer(), ctypes.c_void_p), int(conp), n,
                                 _p(phi), *[_p(o) for o in outs])
    assert st == 0, st
    return tuple(outs)


def jacobian(mech, conp, param, phi, nthreads=0):
    """Dense analytical Jacobian ``[n, NN, NN]``."""
    param, phi = _prep(mech, param, phi)
    n = phi.shape[0]
    out = np.empty((n, mech.NN, mech.NN))
    st = lib().oracle_jacobian(ctypes.cast(mech.pointer(), ctypes.c_void_p), int(conp), n,
                               _p(param), _p(phi), _p(out), int(nthreads))
    assert st == 0, st
    return out


def jacobian_ad(mech, conp, param, phi):
    """Dense forward-mode AD Jacobian of the species rates ``[n, NN, NN]``."""
    param, phi = _prep(mech, param, phi)
    n = phi.shape[0]
    out = np.empty((n, mech.NN, mech.NN))
    st = lib().oracle_jacobian_ad(ctypes.cast(mech.pointer(), ctypes.c_void_p), int(conp), n,
                                  _p(param), _p(phi), _p(out))
    assert st == 0, st
    return out


def jacobian_on_pattern(mech, conp, param, phi):
    """Dense analytical Jacobian ``[n, NN, NN]`` on the mechanism's sparse
    pattern (``mech.sparse_rows_cols``): the T and extra-variable rows sum
    only the stored species entries.  With an approximate pattern, this is
    the approximate Jacobian."""
    param, phi = _prep(mech, param, phi)
    n = phi.shape[0]
    rows, cols = mech.sparse_rows_cols('C')
    mask = np.zeros((mech.NN, mech.NN), dtype=np.uint8)
    mask[rows, cols] = 1
    out = np.empty((n, mech.NN, mech.NN))
    st = lib().oracle_jacobian_masked(ctypes.cast(mech.pointer(), ctypes.c_void_p), int(conp), n,
                                      _p(param), _p(phi), _p(out), _p(mask))
    assert st == 0, st
    return out


def fd_jacobian(mech, conp, param, phi, mode=0, order=1, rtol=1e-8, atol=1e-15):
    """Dense finite-difference Jacobian ``[n, NN, NN]`` (mode 0 forward, 1
    central, 2 backward)."""
    param, phi = _prep(mech, param, phi)
    n = phi.shape[0]
    out = np.empty((n, mech.NN, mech.NN))
    st = lib().oracle_fd_jacobian(ctypes.cast(mech.pointer(), ctypes.c_void_p), int(conp), n,
                                  _p(param), _p(phi), _p(out), int(mode), int(order),
                                  float(rtol), float(atol))
    assert st == 0, st
    return out


def integrate_ros2(mech, conp, param, phi, t_end, rtol=1e-6, atol=1e-12, h0=0.0, fixed=False,
                   max_steps=100000):
    """ROS2 integration (the scheme of include/pyjac_integrator.h) of each
    state over [0, t_end]; returns (phi at t_end, stats [accepted, rejected,
    largest attempts, failed])."""
    param, phi = _prep(mech, param, phi)
    y = phi.copy()
    stats = np.zeros(4, dtype=np.int64)
    st = lib().oracle_integrate_ros2(ctypes.cast(mech.pointer(), ctypes.c_void_p), int(conp),
                                     y.shape[0], _p(param), _p(y), float(t_end), float(rtol),
                                     float(atol), float(h0), int(bool(fixed)), int(max_steps),
                                     _p(stats))
    assert st == 0, st
    return y, stats


def to_layout(mech, jac_dense, order, jac_format):
    """Re-lay a dense ``[n, NN, NN]`` Jacobian into the kernel output layout:
    full C ``[n, NN, NN]``, full F ``[NN(col), NN(row), n]``, sparse C (CRS)
    ``[n, nnz]``, sparse F (CCS) ``[nnz, n]``."""
    if jac_format == 0:
        if order == 'C':
            return np.ascontiguousarray(jac_dense)
        return np.ascontiguousarray(np.transpose(jac_dense, (2, 1, 0)))
    rows, cols = mech.sparse_rows_cols(order)
    vals = jac_dense[:, rows, cols]
    return np.ascontiguousarray(vals if order == 'C' else vals.T)
