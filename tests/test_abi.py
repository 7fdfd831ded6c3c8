"""The C ABI library and the checker load and export what their headers
declare (no device work: these run without a GPU)."""

import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, 'include', header)).read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    return sorted(set(re.findall(r'\b(pyjac_\w+)\s*\(', text)))


@pytest.mark.parametrize('header,least', [('pyjac_hip.h', 15), ('pyjac_surrogate.h', 11)])
def test_library_exports_every_declared_symbol(header, least):
    from pyjac.pywrap import load_library
    lib = load_library()
    names = _declared(header)
    assert len(names) >= least
    for name in names:
        assert hasattr(lib, name), name


def test_mech_struct_matches_header():
    """Field order of the ctypes mirror equals the C struct's."""
    from pyjac.core.mech_tables import CMechTables
    text = open(os.path.join(ROOT, 'include', 'pyjac_hip.h')).read()
    body = text[text.index('typedef struct pyjac_mech_tables {'):text.index('} pyjac_mech_tables;')]
    body = re.sub(r'/\*.*?\*/', '', body, flags=re.S)
    fields = []
    for decl in body.split(';'):
        decl = decl.strip()
        if not decl or decl.startswith('typedef'):
            decl = decl.split('{', 1)[-1].strip()
        if not decl:
            continue
        decl = re.sub(r'^(const\s+)?(int32_t|double|char\s*\*\s*const)\s*', '', decl)
        fields += [f.strip().lstrip('*') for f in decl.split(',')]
    assert fields == [f[0] for f in CMechTables._fields_]


def test_last_error_without_gpu():
    from pyjac.pywrap import load_library
    lib = load_library()
    assert isinstance(lib.pyjac_hip_last_error(), bytes)
    # argument errors are reported as status codes, never by exiting
    h = ctypes.c_void_p()
    assert lib.pyjac_hip_create(None, 1, b'C', 0, 0, ctypes.byref(h)) == 1
    assert b'null' in lib.pyjac_hip_last_error()
