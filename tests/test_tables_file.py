"""The tables file (MechTables.save_binary -> pyjac_hip_tables_load, no GPU
needed): every scalar and every array comes back bit for bit, in the
struct's field order, with the species names and reaction strings of the
trailer, for mechanisms with and without PLOG / Chebyshev / falloff tables;
damaged files are errors, and a file without the trailer (the first layout)
still loads, without names."""

import ctypes

import numpy as np
import pytest

from pyjac.core.exceptions import HipRuntimeError
from pyjac.core.mech_tables import CMechTables, _FIELDS
from pyjac.pywrap import check, load_library


def _load(path):
    L = load_library()
    L.pyjac_hip_tables_load.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.POINTER(CMechTables))]
    L.pyjac_hip_tables_free.argtypes = [ctypes.POINTER(CMechTables)]
    out = ctypes.POINTER(CMechTables)()
    check(L.pyjac_hip_tables_load(str(path).encode(), ctypes.byref(out)))
    return L, out


@pytest.mark.parametrize('name', ['h2o2', 'gri30', 'test', 's150'])
def test_round_trip(mechs, tmp_path, name):
    mech = mechs(name)
    path = mech.save_binary(str(tmp_path / 'tables.pjt'))
    L, out = _load(path)
    try:
        t = out.contents
        for field, ctype in _FIELDS:
            if field in ('species_names', 'reaction_strings'):
                names = mech.species_names if field == 'species_names' else mech.reaction_strings
                got = [getattr(t, field)[i].decode() for i in range(len(names))]
                assert got == names, field
            elif ctype is ctypes.c_int32:
                assert getattr(t, field) == getattr(mech.cstruct, field), field
            else:
                ref = mech.arrays[field]
                got = np.ctypeslib.as_array(getattr(t, field), shape=(ref.size,))
                assert np.array_equal(got.view(np.uint8), np.ascontiguousarray(ref).view(np.uint8)), field
    finally:
        L.pyjac_hip_tables_free(out)


def test_damaged_files(mechs, tmp_path):
    mech = mechs('h2o2')
    good = open(mech.save_binary(str(tmp_path / 'good.pjt')), 'rb').read()
    (tmp_path / 'short.pjt').write_bytes(good[:len(good) // 2])
    (tmp_path / 'magic.pjt').write_bytes(b'XXXXXXXX' + good[8:])
    with pytest.raises(HipRuntimeError, match='truncated|missing'):
        _load(tmp_path / 'short.pjt')
    with pytest.raises(HipRuntimeError, match='not a pyJac'):
        _load(tmp_path / 'magic.pjt')
    with pytest.raises(HipRuntimeError, match='could not open'):
        _load(tmp_path / 'none.pjt')


def test_file_without_names_trailer(mechs, tmp_path):
    """A tables file of the first layout (no names trailer) loads, and its
    struct carries NULL name fields."""
    mech = mechs('h2o2')
    good = open(mech.save_binary(str(tmp_path / 'good.pjt')), 'rb').read()
    cut = good.index(mech.NAMES_MAGIC)
    (tmp_path / 'v1.pjt').write_bytes(good[:cut])
    L, out = _load(tmp_path / 'v1.pjt')
    try:
        assert not out.contents.species_names
        assert not out.contents.reaction_strings
        assert out.contents.n_species == mech.Ns
    finally:
        L.pyjac_hip_tables_free(out)
