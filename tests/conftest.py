import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, 'pyjac-v2_amd'))
sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (run with -m gpu)')


_MECHS = {}


def get_mech(name, **kw):
    from pyjac.core.mech_tables import MechTables
    from pyjac._paths import mechanism_files
    key = (name, tuple(sorted(kw.items())))
    if key not in _MECHS:
        mf, tf = mechanism_files(name)
        _MECHS[key] = MechTables.from_file(mf, tf, **kw)
    return _MECHS[key]


@pytest.fixture(scope='session')
def mechs():
    return get_mech
