"""The mechanism-specialised analytical Jacobian generator
(pyjac.libgen.lanejac) on the CPU: its source compiled for the host with g++
(the same per-state code the gfx950 kernels run, in a loop) against the
oracle's analytical Jacobian, in all four output layouts, on seeded and guard
states; a planted generator slip fails; the module's entry list is the
tables' pattern."""

import ctypes
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

import oracle
from jac_tol import assert_jacobian_close, jacobian_errors
from test_gpu_guards import guard_states
from pyjac.core.states import sample_states
from pyjac.libgen import lanejac

D = ctypes.POINTER(ctypes.c_double)

_BUILT = {}


def _compile(src, d, name):
    cpp = os.path.join(str(d), name + '.cpp')
    so = os.path.join(str(d), name + '.so')
    with open(cpp, 'w') as f:
        f.write(src)
    gxx = shutil.which('g++') or 'g++'
    res = subprocess.run([gxx, '-O1', '-ffp-contract=off', '-shared', '-fPIC', '-o', so, cpp],
                         stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert res.returncode == 0, res.stdout[-3000:]
    lib = ctypes.CDLL(so)
    fn = lib.pyjac_ljac_host
    fn.argtypes = [ctypes.c_long, D, D, D, D, ctypes.c_int, ctypes.c_long, ctypes.c_long,
                   ctypes.c_long, ctypes.c_int]
    _LIBS[id(fn)] = (fn, lib)
    return fn


_LIBS = {}


def _host(mech, name, tmp_path_factory):
    if name not in _BUILT:
        d = tmp_path_factory.mktemp('ljac_' + name)
        _BUILT[name] = _compile(lanejac.generate(mech, 'host'), d, name)
    return _BUILT[name]


def _strided(fn):
    """The host build's store kernel with explicit output strides."""
    g = _LIBS[id(fn)][1].pyjac_ljac_host_strided
    g.argtypes = [ctypes.c_long, D, D, D, D, ctypes.c_int] + [ctypes.c_long] * 5 + [ctypes.c_int]
    return g


def _run(fn, mech, conp, param, phi, order, fmt):
    """The host build's Jacobian in the layout of (order, fmt)."""
    n, NN = len(param), mech.NN
    layout = (2 if fmt == 0 else 0) + (1 if order == 'F' else 0)
    x = np.ascontiguousarray(phi if order == 'C' else phi.T)
    width = NN * NN if fmt == 0 else mech.nnz
    out = np.full(n * width, np.nan)
    scr = np.zeros(lanejac.LaneJacSource(mech).nscr * n)
    sN, sI = (NN, 1) if order == 'C' else (1, n)
    jN = width if order == 'C' else 1
    fn(n, param.ctypes.data_as(D), x.ctypes.data_as(D), scr.ctypes.data_as(D),
       out.ctypes.data_as(D), layout, sN, sI, jN, int(conp))
    return out


@pytest.mark.parametrize('name', ['h2o2', 'gri30', 'test'])
def test_host_build_matches_oracle(mechs, tmp_path_factory, name):
    """Seeded states (T 900-2800 K), CONP and CONV: within the GPU parity
    tests' Jacobian bounds of the oracle in every layout (sparse C and F,
    full C and F; the full layouts zero outside the pattern)."""
    mech = mechs(name)
    fn = _host(mech, name, tmp_path_factory)
    for conp in (True, False):
        param, phi = sample_states(mech, 240, seed=5, T_range=(900.0, 2800.0), alpha=0.3,
                                   conp=conp)
        dense = oracle.jacobian(mech, conp, param, phi)
        for order in ('C', 'F'):
            for fmt in (0, 1):
                got = _run(fn, mech, conp, param, phi, order, fmt)
                ref = oracle.to_layout(mech, dense, order, fmt).reshape(-1)
                n = len(param)
                per = lambda a: a.reshape(n, -1) if order == 'C' else a.reshape(-1, n).T
                assert_jacobian_close(per(got), per(ref), '%s conp=%s %s fmt=%d'
                                      % (name, conp, order, fmt))


@pytest.mark.parametrize('name', ['h2o2', 'gri30', 'test'])
def test_host_build_guards(mechs, tmp_path_factory, name):
    """States outside the temperature clamp, below the volume floor, with
    zero and negative moles: the oracle's values (finite where the oracle's
    are)."""
    mech = mechs(name)
    fn = _host(mech, name, tmp_path_factory)
    for conp in (True, False):
        param, phi = guard_states(mech, conp)
        n = len(param)
        ref = oracle.jacobian(mech, conp, param, phi).reshape(n, -1)
        got = _run(fn, mech, conp, param, phi, 'C', 0).reshape(n, -1)
        fin = np.isfinite(ref).all(axis=1)
        assert (np.isfinite(got).all(axis=1) == fin).all()
        assert_jacobian_close(got[fin], ref[fin], '%s guards conp=%s' % (name, conp))


@pytest.mark.parametrize('name', ['h2o2', 'gri30', 'test'])
def test_c_order_through_f_lines(mechs, tmp_path_factory, name):
    """The runtime's C-order path through lines: the F-layout store kernel
    writing lines of the states (entry stride = the number of states, state
    offset 0), each C-order entry then read from the line of its F-layout
    position (the CCS index of the CRS entry; column-major for full), is the
    C-layout kernel's C-order output bit for bit, sparse and full, CONP and
    CONV."""
    mech = mechs(name)
    g = _strided(_host(mech, name, tmp_path_factory))
    NN = mech.NN
    rc, cc = mech.sparse_rows_cols('C')
    rf, cf = mech.sparse_rows_cols('F')
    ccs = {(int(a), int(b)): q for q, (a, b) in enumerate(zip(rf, cf))}
    fpos_sparse = np.array([ccs[(int(a), int(b))] for a, b in zip(rc, cc)])
    fpos_full = np.array([col * NN + row for row in range(NN) for col in range(NN)])
    for conp in (True, False):
        param, phi = sample_states(mech, 150, seed=9, T_range=(900.0, 2800.0), alpha=0.3,
                                   conp=conp)
        n = len(param)
        phi = np.ascontiguousarray(phi)
        for layout, ne, fpos in ((0, mech.nnz, fpos_sparse), (2, NN * NN, fpos_full)):
            outs = []
            for L, jN, jI in ((layout, ne, 1), (layout | 1, 1, n)):
                scr = np.zeros(lanejac.LaneJacSource(mech).nscr * n)
                out = np.full(n * ne, np.nan)
                g(n, param.ctypes.data_as(D), phi.ctypes.data_as(D), scr.ctypes.data_as(D),
                  out.ctypes.data_as(D), L, NN, 1, jN, 0, jI, int(conp))
                outs.append(out)
            direct, lines = outs
            assert not np.isnan(lines).any()
            back = lines.reshape(ne, n)[fpos].T.reshape(-1)
            assert np.array_equal(direct, back, equal_nan=True)


def test_planted_mistake_fails(mechs, tmp_path_factory):
    """A generator slip (one temperature-derivative Arrhenius exponent off by
    1e-6) is caught by the bounds above."""
    mech = mechs('h2o2')
    src = lanejac.generate(mech, 'host')
    m = re.search(r'const double dkf = \((\(-?[0-9.e+-]+\)|[0-9.e+-]+) \+', src)
    b = float(m.group(1).strip('()'))
    assert b != 0.0
    bad = src[:m.start(1)] + '(%r)' % (b * (1 + 1e-6)) + src[m.end(1):]
    fn = _compile(bad, tmp_path_factory.mktemp('ljac_bad'), 'bad')
    param, phi = sample_states(mech, 200, seed=5, conp=True)
    n = len(param)
    ref = oracle.jacobian(mech, True, param, phi).reshape(n, -1)
    got = _run(fn, mech, True, param, phi, 'C', 0).reshape(n, -1)
    fro, elem, _ = jacobian_errors(got, ref)
    assert elem > 1.0


@pytest.mark.parametrize('name', ['h2o2', 'gri30', 'test'])
def test_entries_are_the_pattern(mechs, name):
    """The module stores each entry of the tables' sparsity pattern once,
    column by column."""
    mech = mechs(name)
    ent = lanejac.entry_list(mech)
    rows, cols = mech.sparse_rows_cols('C')
    assert sorted(ent) == sorted(zip(rows.tolist(), cols.tolist()))
    assert len(set(ent)) == len(ent) == mech.nnz


def test_source_is_deterministic_and_signed(mechs):
    mech = mechs('gri30')
    a, b = lanejac.generate(mech), lanejac.generate(mech)
    assert a == b
    assert '0x%016xULL' % lanejac.tables_hash(mech) in a
    for k in ('pyjac_ljac_a', 'pyjac_ljac_b0', 'pyjac_ljac_b3', 'pyjac_ljac_sig',
              'pyjac_ljac_entries'):
        assert k in a


def test_refuses_large_and_approximate(mechs, tmp_path):
    from pyjac.core.enum_types import JacobianType
    with pytest.raises(ValueError, match='species'):
        lanejac.build(mechs('s150'), str(tmp_path / 'x.hsaco'))
    with pytest.raises(ValueError, match='exact'):
        lanejac.build(mechs('gri30', jac_type=JacobianType.approximate),
                      str(tmp_path / 'y.hsaco'))
    assert lanejac.ensure_module(mechs('s150'), build_missing=False) is None
