"""Pack the rate tables into the ``pyjac_mech_tables`` C struct.

The struct (declared in ``include/pyjac_hip.h``) is the one mechanism format
both the HIP library and the CPU checker read.  :class:`MechTables` owns the
numpy arrays and hands out a ``ctypes`` pointer to a filled struct.
"""

import ctypes

import numpy as np

from .enum_types import RateSpecialization, JacobianType
from .mech_interpret import load_mechanism
from .rate_tables import determine_jac_inds, named_tables

ABI_VERSION = 2

ARR_A, ARR_AT_B, ARR_FULL, ARR_NO_TA, ARR_NO_B = range(5)

_I = ctypes.POINTER(ctypes.c_int32)
_D = ctypes.POINTER(ctypes.c_double)

_FIELDS = [
    ('abi_version', ctypes.c_int32), ('n_species', ctypes.c_int32),
    ('n_reactions', ctypes.c_int32), ('n_rxn_spec', ctypes.c_int32),
    ('n_simple', ctypes.c_int32), ('n_rev', ctypes.c_int32),
    ('n_thd', ctypes.c_int32), ('n_thd_spec', ctypes.c_int32),
    ('n_fall', ctypes.c_int32), ('n_troe', ctypes.c_int32),
    ('n_sri', ctypes.c_int32), ('n_plog', ctypes.c_int32),
    ('plog_max_p', ctypes.c_int32), ('n_cheb', ctypes.c_int32),
    ('cheb_max_t', ctypes.c_int32), ('cheb_max_p', ctypes.c_int32),
    ('n_rxn_has_ns', ctypes.c_int32), ('jac_nnz', ctypes.c_int32),
    ('rxn_off', _I), ('rxn_spec', _I), ('nu_prod', _I), ('nu_reac', _I),
    ('nu_sum', _I),
    ('rev_mask', _I), ('thd_mask', _I), ('simple_mask', _I),
    ('plog_mask', _I), ('cheb_mask', _I), ('fall_mask', _I),
    ('simple_form', _I), ('simple_A', _D), ('simple_b', _D),
    ('simple_Ta', _D),
    ('fall_map', _I), ('fall_type', _I), ('fall_form', _I),
    ('fall_blend', _I), ('fall_param', _I), ('fall_to_thd', _I),
    ('fall_A', _D), ('fall_b', _D), ('fall_Ta', _D),
    ('troe_a', _D), ('troe_T1inv', _D), ('troe_T2', _D), ('troe_T3inv', _D),
    ('sri_a', _D), ('sri_b', _D), ('sri_c', _D), ('sri_d', _D), ('sri_e', _D),
    ('thd_map', _I), ('thd_type', _I), ('thd_off', _I), ('thd_spec', _I),
    ('thd_eff', _D),
    ('plog_map', _I), ('plog_np', _I), ('plog_params', _D),
    ('cheb_map', _I), ('cheb_nt', _I), ('cheb_np', _I),
    ('cheb_params', _D), ('cheb_plim', _D), ('cheb_tlim', _D),
    ('a_lo', _D), ('a_hi', _D), ('T_mid', _D), ('mw_post', _D),
    ('rxn_has_ns', _I),
    ('ccs_col_ptr', _I), ('ccs_row_ind', _I), ('crs_row_ptr', _I),
    ('crs_col_ind', _I),
    ('species_names', ctypes.POINTER(ctypes.c_char_p)),
    ('reaction_strings', ctypes.POINTER(ctypes.c_char_p)),
]
_NAME_FIELDS = ('species_names', 'reaction_strings')


class CMechTables(ctypes.Structure):
    """ctypes mirror of ``pyjac_mech_tables``."""
    _fields_ = _FIELDS


def _form(rate_spec, rtype):
    rtype = np.asarray(rtype, dtype=np.int32)
    if rate_spec == RateSpecialization.fixed:
        return np.full(rtype.shape, ARR_FULL, dtype=np.int32)
    if rate_spec == RateSpecialization.hybrid:
        lut = np.array([ARR_A, ARR_AT_B, ARR_FULL], dtype=np.int32)
    else:
        lut = np.array([ARR_A, ARR_AT_B, ARR_NO_TA, ARR_NO_B, ARR_FULL],
                       dtype=np.int32)
    return lut[rtype] if rtype.size else rtype


class MechTables(object):
    """Mechanism tables for one (mechanism, rate specialisation, Jacobian
    type) choice.

    Attributes of interest: ``Ns``, ``Nr``, ``NN``, ``nnz``,
    ``species_names``, ``reaction_strings``, ``info`` (the
    ``assign_rates``/``determine_jac_inds`` dict) and ``arrays`` (every
    packed array by struct field name).
    """

    def __init__(self, specs, reacs, rate_spec=RateSpecialization.hybrid,
                 jac_type=JacobianType.exact, name='mech'):
        self.name = name
        self.rate_spec = RateSpecialization(rate_spec)
        self.jac_type = JacobianType(jac_type)
        self.specs = specs
        self.reacs = reacs
        info = determine_jac_inds(reacs, specs, self.rate_spec, self.jac_type)
        if not info['net']['allint']:
            raise NotImplementedError(
                'non-integer stoichiometric coefficients are not supported '
                'by the HIP kernels')
        self.info = info
        self.named = named_tables(info)
        self.Ns = info['Ns']
        self.Nr = info['Nr']
        self.NN = self.Ns + 1
        self.nnz = int(len(info['jac_inds']['flat_C']))
        self.species_names = [s.name for s in specs]
        self.reaction_strings = [str(r) for r in reacs]
        self.mw = np.asarray(info['mws'], dtype=np.float64)
        self._pack()

    @classmethod
    def from_file(cls, mech, therm=None, last_spec=None, **kw):
        _, specs, reacs = load_mechanism(mech, therm, last_spec=last_spec)
        kw.setdefault('name', str(mech).rsplit('/', 1)[-1].rsplit('.', 1)[0])
        return cls(specs, reacs, **kw)

    def _pack(self):
        info, nt = self.info, self.named
        Nr, Ns = self.Nr, self.Ns
        i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32).reshape(-1)
        f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        A = {}
        A['rxn_off'] = i32(nt['net_reac_to_spec_offsets'])
        A['rxn_spec'] = i32(nt['rxn_to_spec'])
        nu = np.asarray(info['net']['nu']).reshape(-1, 2)
        A['nu_prod'] = i32(nu[:, 0])
        A['nu_reac'] = i32(nu[:, 1])
        A['nu_sum'] = i32(info['net']['nu_sum'])

        def mask(mapping, n):
            m = -np.ones(Nr, dtype=np.int32)
            m[np.asarray(mapping, dtype=np.int64)] = np.arange(n, dtype=np.int32)
            return m

        s = info['simple']
        A['rev_mask'] = mask(info['rev']['map'], info['rev']['num'])
        A['thd_mask'] = mask(info['thd']['map'], info['thd']['num'])
        A['simple_mask'] = mask(s['map'], s['num'])
        A['plog_mask'] = mask(info['plog']['map'], info['plog']['num'])
        A['cheb_mask'] = mask(info['cheb']['map'], info['cheb']['num'])
        f = info['fall']
        A['fall_mask'] = mask(f['map'], f['num'])
        A['simple_form'] = i32(_form(self.rate_spec, s['type']))
        A['simple_A'] = f64(s['A'])
        A['simple_b'] = f64(s['b'])
        A['simple_Ta'] = f64(s['Ta'])

        nf = f['num']
        A['fall_map'] = i32(f['map'])
        A['fall_type'] = i32(f['ftype'] if nf else [])
        A['fall_form'] = i32(_form(self.rate_spec, f['type']) if nf else [])
        A['fall_blend'] = i32(f['blend'])
        param = np.zeros(nf, dtype=np.int32)
        param[f['troe']['map']] = np.arange(f['troe']['num'])
        param[f['sri']['map']] = np.arange(f['sri']['num'])
        A['fall_param'] = param
        A['fall_to_thd'] = i32(nt.get('fall_to_thd_map', []))
        A['fall_A'] = f64(f['A'])
        A['fall_b'] = f64(f['b'])
        A['fall_Ta'] = f64(f['Ta'])
        tr = f['troe']
        A['troe_a'] = f64(tr['a'])
        A['troe_T1inv'] = f64(tr['T1'])
        A['troe_T2'] = f64(tr['T2'])
        A['troe_T3inv'] = f64(tr['T3'])
        for k in 'abcde':
            A['sri_' + k] = f64(f['sri'][k])

        t = info['thd']
        A['thd_map'] = i32(t['map'])
        A['thd_type'] = i32(t['type'])
        A['thd_off'] = i32(np.concatenate(([0], np.cumsum(t['spec_num']))))
        A['thd_spec'] = i32(t['spec'])
        A['thd_eff'] = f64(t['eff'])

        p = info['plog']
        A['plog_map'] = i32(p['map'])
        A['plog_np'] = i32(p['num_P'])
        A['plog_params'] = f64(p['post_process']['params'])
        c = info['cheb']
        A['cheb_map'] = i32(c['map'])
        A['cheb_nt'] = i32(c['num_T'])
        A['cheb_np'] = i32(c['num_P'])
        A['cheb_params'] = f64(c['post_process']['params'])
        A['cheb_plim'] = f64(c['post_process']['Plim'])
        A['cheb_tlim'] = f64(c['post_process']['Tlim'])

        A['a_lo'] = f64(info['thermo']['a_lo'])
        A['a_hi'] = f64(info['thermo']['a_hi'])
        A['T_mid'] = f64(info['thermo']['T_mid'])
        A['mw_post'] = f64(info['mw_post'])
        A['rxn_has_ns'] = i32(info['reac_has_ns'])
        ji = info['jac_inds']
        A['ccs_col_ptr'] = i32(ji['ccs']['col_ptr'])
        A['ccs_row_ind'] = i32(ji['ccs']['row_ind'])
        A['crs_row_ptr'] = i32(ji['crs']['row_ptr'])
        A['crs_col_ind'] = i32(ji['crs']['col_ind'])
        # never hand out a NULL pointer, even for empty tables
        for k, v in list(A.items()):
            if v.size == 0:
                A[k] = np.zeros(1, dtype=v.dtype)
        self.arrays = A

        st = CMechTables()
        st.abi_version = ABI_VERSION
        st.n_species = Ns
        st.n_reactions = Nr
        st.n_rxn_spec = int(len(nt['rxn_to_spec']))
        st.n_simple = s['num']
        st.n_rev = info['rev']['num']
        st.n_thd = t['num']
        st.n_thd_spec = int(len(t['spec']))
        st.n_fall = nf
        st.n_troe = tr['num']
        st.n_sri = f['sri']['num']
        st.n_plog = p['num']
        st.plog_max_p = int(p['max_P'] or 0)
        st.n_cheb = c['num']
        st.cheb_max_t = int(np.max(c['num_T'])) if c['num'] else 0
        st.cheb_max_p = int(np.max(c['num_P'])) if c['num'] else 0
        st.n_rxn_has_ns = int(len(info['reac_has_ns']))
        st.jac_nnz = self.nnz
        for name, ctype in _FIELDS:
            if name in A:
                arr = A[name]
                setattr(st, name, arr.ctypes.data_as(ctype))
        # the names the introspection entry points return (kept alive here)
        self._names_c = []
        for field, names in zip(_NAME_FIELDS, (self.species_names, self.reaction_strings)):
            arr = (ctypes.c_char_p * max(1, len(names)))(*[n.encode() for n in names])
            self._names_c.append(arr)
            setattr(st, field, ctypes.cast(arr, ctypes.POINTER(ctypes.c_char_p)))
        self.cstruct = st

    def pointer(self):
        """``ctypes`` pointer to the filled ``pyjac_mech_tables``."""
        return ctypes.pointer(self.cstruct)

    TABLES_MAGIC = b'PJHIPTB1'
    NAMES_MAGIC = b'PJNAMES1'

    def save_binary(self, path):
        """Write the tables as a binary file that C and C++ callers load with
        ``pyjac_hip_tables_load`` (include/pyjac_hip.h): the magic
        ``PJHIPTB1``, the number of scalar fields and their int32 values, the
        number of arrays, then per array (in the struct's field order) its
        element type (int32 0 / float64 1), int64 length and data; then the
        trailer ``PJNAMES1`` with the species names and the reaction strings
        (per list an int32 count, per name an int32 length and its bytes)."""
        scalars = [(n, t) for n, t in _FIELDS if t is ctypes.c_int32]
        arrays = [(n, t) for n, t in _FIELDS
                  if t is not ctypes.c_int32 and n not in _NAME_FIELDS]
        with open(path, 'wb') as f:
            f.write(self.TABLES_MAGIC)
            f.write(np.int32(len(scalars)).tobytes())
            f.write(np.array([getattr(self.cstruct, n) for n, _ in scalars], dtype=np.int32)
                    .tobytes())
            f.write(np.int32(len(arrays)).tobytes())
            for n, t in arrays:
                is_d = t is _D
                a = np.ascontiguousarray(self.arrays[n], dtype=np.float64 if is_d else np.int32)
                f.write(np.uint8(1 if is_d else 0).tobytes())
                f.write(np.int64(a.size).tobytes())
                f.write(a.tobytes())
            # trailer: species names and reaction strings
            f.write(self.NAMES_MAGIC)
            for names in (self.species_names, self.reaction_strings):
                f.write(np.int32(len(names)).tobytes())
                for nm in names:
                    b = nm.encode()
                    f.write(np.int32(len(b)).tobytes())
                    f.write(b)
        return path

    # sparse index helpers --------------------------------------------------
    def sparse_rows_cols(self, order):
        """(rows, cols) of each stored sparse value: CCS for 'F', CRS for
        'C' (reference utils.py:1014-1022)."""
        ji = self.info['jac_inds']
        if order == 'F':
            return ji['flat_F'][:, 0], ji['flat_F'][:, 1]
        return ji['flat_C'][:, 0], ji['flat_C'][:, 1]
