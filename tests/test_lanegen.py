"""The mechanism-specialised species-rate generator (pyjac.libgen.lanegen) on
the CPU: its source compiled for the host with g++ (the same per-state code
the gfx950 module runs, in a loop) against the oracle, the tables'
fingerprint in Python against the library's, and the refusals."""

import ctypes
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from rate_tol import rate_error, rate_scales
from test_gpu_guards import guard_states
from pyjac.core.states import sample_states
from pyjac.libgen import lanegen

D = ctypes.POINTER(ctypes.c_double)

_BUILT = {}


def _host_lib(mech, name, tmp_path_factory):
    if name not in _BUILT:
        d = tmp_path_factory.mktemp('lane_' + name)
        src = os.path.join(str(d), name + '.cpp')
        so = os.path.join(str(d), name + '.so')
        with open(src, 'w') as f:
            f.write(lanegen.generate(mech, 'host'))
        gxx = shutil.which('g++') or 'g++'
        res = subprocess.run([gxx, '-O1', '-ffp-contract=off', '-shared', '-fPIC', '-o', so, src],
                             stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        assert res.returncode == 0, res.stdout[-3000:]
        lib = ctypes.CDLL(so)
        fn = lib.pyjac_lane_rates_host
        fn.argtypes = [ctypes.c_long, D, D, D, ctypes.c_long, ctypes.c_long, ctypes.c_int]
        _BUILT[name] = (lib, fn)
    return _BUILT[name][1]


def _run(fn, mech, conp, param, phi, order):
    n = len(param)
    NN = mech.NN
    x = phi if order == 'C' else np.ascontiguousarray(phi.T)
    out = np.full_like(x, np.nan)
    sN, sI = (NN, 1) if order == 'C' else (1, n)
    fn(n, param.ctypes.data_as(D), x.ctypes.data_as(D), out.ctypes.data_as(D), sN, sI, int(conp))
    return out if order == 'C' else np.ascontiguousarray(out.T)


@pytest.mark.parametrize('name', ['h2o2', 'gri30', 'test'])
def test_host_build_matches_oracle(mechs, tmp_path_factory, name):
    """Seeded states (T 900-2800 K), CONP and CONV, C and F strides: within
    the GPU parity tests' bound of the oracle (the host build evaluates the
    same expressions in the same order, so in practice bit for bit)."""
    mech = mechs(name)
    fn = _host_lib(mech, name, tmp_path_factory)
    for conp in (True, False):
        param, phi = sample_states(mech, 400, seed=5, T_range=(900.0, 2800.0), alpha=0.3,
                                   conp=conp)
        ref, scale = rate_scales(mech, conp, param, phi)
        for order in ('C', 'F'):
            got = _run(fn, mech, conp, param, phi, order)
            assert np.isfinite(got).all()
            err = rate_error(got, ref, scale)
            print(name, conp, order, 'max error / bound', err,
                  'elements differing', int((got != ref).sum()))
            assert err <= 1.0


@pytest.mark.parametrize('name', ['h2o2', 'gri30', 'test'])
def test_host_build_guards(mechs, tmp_path_factory, name):
    """States outside the temperature clamp, below the volume floor, with
    zero and negative moles: the same values as the oracle."""
    mech = mechs(name)
    fn = _host_lib(mech, name, tmp_path_factory)
    for conp in (True, False):
        param, phi = guard_states(mech, conp)
        ref, scale = rate_scales(mech, conp, param, phi)
        got = _run(fn, mech, conp, param, phi, 'C')
        fin = np.isfinite(ref)
        assert (np.isfinite(got) == fin).all()
        assert rate_error(got[fin.all(axis=1)], ref[fin.all(axis=1)],
                          scale[fin.all(axis=1)]) <= 1.0


def test_planted_mistake_fails(mechs, tmp_path_factory):
    """A generator slip (one Arrhenius exponent off by 1e-6) is caught by
    the comparison above."""
    mech = mechs('h2o2')
    src = lanegen.generate(mech, 'host')
    m = re.search(r'logT \* (\(-?[0-9.e+-]+\)|[0-9.e+-]+)', src)
    b = float(m.group(1).strip('()'))
    bad = src[:m.start(1)] + '(%r)' % (b * (1 + 1e-6)) + src[m.end(1):]
    d = tmp_path_factory.mktemp('lane_bad')
    cpp, so = os.path.join(str(d), 'bad.cpp'), os.path.join(str(d), 'bad.so')
    with open(cpp, 'w') as f:
        f.write(bad)
    subprocess.check_call([shutil.which('g++') or 'g++', '-O1', '-ffp-contract=off', '-shared',
                           '-fPIC', '-o', so, cpp])
    fn = ctypes.CDLL(so).pyjac_lane_rates_host
    fn.argtypes = [ctypes.c_long, D, D, D, ctypes.c_long, ctypes.c_long, ctypes.c_int]
    param, phi = sample_states(mech, 200, seed=5, conp=True)
    ref, scale = rate_scales(mech, True, param, phi)
    assert rate_error(_run(fn, mech, True, param, phi, 'C'), ref, scale) > 1.0


def test_fingerprint_matches_library(mechs):
    """lanegen.tables_hash and pyjac_hip_tables_hash agree on every bundled
    and synthetic mechanism and tell them apart."""
    from pyjac.pywrap import load_library
    L = load_library()
    seen = set()
    for name in ('h2o2', 'gri30', 'test', 's150'):
        mech = mechs(name)
        out = ctypes.c_ulonglong()
        assert L.pyjac_hip_tables_hash(ctypes.cast(mech.pointer(), ctypes.c_void_p),
                                       ctypes.byref(out)) == 0
        assert out.value == lanegen.tables_hash(mech)
        seen.add(out.value)
    assert len(seen) == 4


def test_source_is_deterministic_and_signed(mechs):
    mech = mechs('gri30')
    a, b = lanegen.generate(mech), lanegen.generate(mech)
    assert a == b
    assert '0x%016xULL' % lanegen.tables_hash(mech) in a
    assert 'pyjac_lane_rates' in a and 'pyjac_lane_sig' in a


def test_refuses_large_mechanisms(mechs, tmp_path):
    with pytest.raises(ValueError, match='species'):
        lanegen.build(mechs('s150'), str(tmp_path / 'x.hsaco'))
