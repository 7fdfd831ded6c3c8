"""Python bindings of the HIP library (ctypes over the C ABI).

Mirrors the reference's Cython wrapper (``pywrap/pyjacob_wrapper.pyx.in``):
``Py<Name>Kernel(problem_size, work_size, do_not_compile=False)`` with
``species_names()``, ``reaction_strings()``, ``num_species()``,
``num_reactions()``, ``required_working_memory()``, ``resize()`` and
``__call__`` taking the kernel's arrays in the reference's argument order
(``utils.kernel_argument_ordering``, utils.py:203-277):

* Jacobian:       ``kernel(param, phi, jac)``
* species rates:  ``kernel(param, phi, dphi)``
* chem utils:     ``kernel(phi, b, specific_energy, specific_heat)``

Arrays are numpy float64 in the kernel's data order ('C' or 'F').  The
calls also accept device arrays (``torch`` CUDA tensors or any object with
``data_ptr()``), in which case no host copy is made.  There is no CPU
fallback: a missing library raises.
"""

import ctypes
import os

import numpy as np

from ..core.exceptions import HipRuntimeError, InvalidInputSpecificationException
from ..core.enum_types import JacobianFormat, JacobianType

LIB_PATH = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                        'libpyjac_hip.so')

_lib = None


def _share_hip_runtime():
    """Load torch's HIP runtime first when torch is installed.  torch links
    its bundled ``libamdhip64`` under a name that does not match a copy from
    /opt/rocm already in the process, so loading this library first would put
    two HIP runtimes in one process, and the second finds no device.  Loaded
    after torch, the library binds to torch's copy."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def load_library(path=None):
    """Load ``libpyjac_hip.so`` (raises if it has not been built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    path = path or os.environ.get('PYJAC_HIP_LIB', LIB_PATH)
    if not os.path.exists(path):
        raise HipRuntimeError(-1, 'HIP library not built: {} (run '
                                  '__graft_entry__.build() or '
                                  'pyjac.libgen.build_library())'.format(path))
    _share_hip_runtime()
    L = ctypes.CDLL(path)
    P, I, S = ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t
    L.pyjac_hip_last_error.restype = ctypes.c_char_p
    L.pyjac_hip_create.argtypes = [P, I, ctypes.c_char, I, I, ctypes.POINTER(P)]
    L.pyjac_hip_destroy.argtypes = [P]
    L.pyjac_hip_resize.argtypes = [P, S, S]
    for name in ('pyjac_hip_species_rates', 'pyjac_hip_jacobian'):
        getattr(L, name).argtypes = [P, P, P, P]
    L.pyjac_hip_chem_utils.argtypes = [P, P, P, P, P]
    for name in ('pyjac_hip_species_rates_sized', 'pyjac_hip_jacobian_sized'):
        getattr(L, name).argtypes = [P, S, P, S, P, S, P, S]
    L.pyjac_hip_chem_utils_sized.argtypes = [P, S, P, S, P, P, P, S]
    L.pyjac_hip_set_input_check.argtypes = [P, I]
    L.pyjac_hip_nonfinite_dev.argtypes = [P, S, P, I, P, ctypes.POINTER(S), ctypes.POINTER(S),
                                          P]
    for name in ('pyjac_hip_species_name', 'pyjac_hip_reaction_string'):
        getattr(L, name).argtypes = [P, I]
        getattr(L, name).restype = ctypes.c_char_p
    L.pyjac_hip_order.argtypes = [P]
    L.pyjac_hip_order.restype = ctypes.c_char
    L.pyjac_hip_vector_width.argtypes = [P]
    for name in ('pyjac_hip_species_rates_dev', 'pyjac_hip_jacobian_dev'):
        getattr(L, name).argtypes = [P, S, P, P, P, P]
    for name in ('pyjac_hip_num_species', 'pyjac_hip_num_reactions',
                 'pyjac_hip_jacobian_size'):
        getattr(L, name).argtypes = [P]
    L.pyjac_hip_required_memory.argtypes = [P]
    L.pyjac_hip_required_memory.restype = S
    L.pyjac_hip_set_variant.argtypes = [P, I, I]
    L.pyjac_hip_get_variant.argtypes = [P, I]
    L.pyjac_hip_batch_create.argtypes = [ctypes.POINTER(P), I, ctypes.POINTER(P)]
    L.pyjac_hip_batch_set.argtypes = [P, I, S, P, P, P]
    L.pyjac_hip_batch_run.argtypes = [P, I, I, P]
    L.pyjac_hip_batch_destroy.argtypes = [P]
    L.pyjac_hip_batch_set_queue.argtypes = [P, I]
    L.pyjac_hip_batch_slot_failures.argtypes = [P, ctypes.POINTER(I)]
    L.pyjac_hip_autotune.argtypes = [P, I, S, P, P, P, P, ctypes.POINTER(I),
                                     ctypes.POINTER(ctypes.c_float)]
    L.pyjac_hip_set_fd_options.argtypes = [P, I, I, ctypes.c_double, ctypes.c_double]
    L.pyjac_hip_set_tuning.argtypes = [P, I, I, ctypes.c_long]
    L.pyjac_hip_get_tuning.argtypes = [P, I, ctypes.POINTER(I), ctypes.POINTER(ctypes.c_long),
                                       ctypes.POINTER(I)]
    L.pyjac_hip_tables_hash.argtypes = [P, ctypes.POINTER(ctypes.c_ulonglong)]
    L.pyjac_hip_load_module.argtypes = [P, ctypes.c_char_p]
    L.pyjac_hip_read_initial_conditions.argtypes = [ctypes.c_char_p, S, I, P, P, ctypes.c_char]
    L.pyjac_hip_write_data.argtypes = [ctypes.c_char_p, P, S]
    if path == LIB_PATH or _lib is None:
        _lib = L
    return L


def check(status):
    if status != 0:
        msg = _lib.pyjac_hip_last_error().decode() if _lib else ''
        raise HipRuntimeError(status, msg)


def _ptr(a):
    """Pointer of a numpy array or a device tensor; (pointer, on_device)."""
    if isinstance(a, np.ndarray):
        if a.dtype != np.float64:
            raise TypeError('arrays must be float64')
        if not (a.flags['C_CONTIGUOUS'] or a.flags['F_CONTIGUOUS']):
            raise ValueError('arrays must be contiguous')
        return ctypes.c_void_p(a.ctypes.data), False
    if hasattr(a, 'data_ptr'):
        if hasattr(a, 'dtype') and str(a.dtype) not in ('torch.float64',):
            raise TypeError('device arrays must be float64')
        if hasattr(a, 'is_contiguous') and not a.is_contiguous():
            raise ValueError('device arrays must be contiguous')
        return ctypes.c_void_p(a.data_ptr()), True
    raise TypeError('unsupported array type {}'.format(type(a)))


def _numel(a):
    return a.size if isinstance(a, np.ndarray) else a.numel()


class _KernelBase(object):
    """Holds one handle (mechanism + options) of the C ABI."""

    kernel_type = None

    def __init__(self, problem_size, work_size=0, do_not_compile=False,
                 mech=None, conp=True, order='C', jac_format=JacobianFormat.sparse,
                 jac_type=JacobianType.exact, **kwargs):
        if mech is None:
            raise InvalidInputSpecificationException('mech')
        if order not in ('C', 'F'):
            raise InvalidInputSpecificationException('order')
        self.lib = load_library()
        self.mech = mech
        self.conp = bool(conp)
        self.order = order
        self.jac_format = int(JacobianFormat(jac_format))
        self.jac_type = int(JacobianType(jac_type))
        if (self.jac_type == JacobianType.approximate) != \
                (getattr(mech, 'jac_type', JacobianType.exact) == JacobianType.approximate):
            # the pattern belongs to the tables (determine_jac_inds)
            raise InvalidInputSpecificationException(
                'jac_type: an approximate Jacobian needs tables built with '
                'MechTables(..., jac_type=JacobianType.approximate), and only then')
        h = ctypes.c_void_p()
        check(self.lib.pyjac_hip_create(ctypes.cast(mech.pointer(), ctypes.c_void_p),
                                        int(self.conp), order.encode(),
                                        self.jac_format, self.jac_type,
                                        ctypes.byref(h)))
        self._h = h
        self.problem_size = 0
        self.module = None
        self.jac_module = None
        module = kwargs.pop('module', None)
        module_dir = kwargs.pop('module_dir', None)
        if module is not None:
            self.load_module(module, module_dir)
        if self._variant_slot <= 1:
            self._apply_tuned()
        if not kwargs.pop('testing_only_skip_compilation', False):
            self.resize(problem_size, work_size, do_not_compile)

    def __del__(self):
        h = getattr(self, '_h', None)
        if h is not None and h.value and self.lib is not None:
            self.lib.pyjac_hip_destroy(h)
            self._h = None

    def load_module(self, module='auto', directory=None):
        """Attach mechanism-specialised code objects: a species-rate module
        (``pyjac.libgen.lanegen``) or an analytical Jacobian module
        (``pyjac.libgen.lanejac``), by path, or 'auto' for the cached
        modules of this mechanism's tables (built by
        ``__graft_entry__.build()`` or ``pyjac --lang hip``; nothing is
        compiled here): the species-rate module always, the Jacobian module
        for a Jacobian kernel of the exact Jacobian.  The species rates (and
        the finite-difference Jacobian's perturbed states) then run on the
        first, the analytical Jacobian on the second.  'auto' looks in
        ``directory`` (a ``pyjac --lang hip`` build folder) when given, else in
        the package's cache.  Returns the species-rate module's path (the
        Jacobian module's is ``jac_module``), or None when 'auto' finds
        neither."""
        if module == 'auto':
            from ..libgen import lanegen, lanejac
            rates = lanegen.ensure_module(self.mech, build_missing=False, directory=directory)
            jac = None
            if self._variant_slot == 1 and self.jac_type == JacobianType.exact:
                jac = lanejac.ensure_module(self.mech, build_missing=False, directory=directory)
            if rates:
                check(self.lib.pyjac_hip_load_module(self._h, os.fsencode(rates)))
            if jac:
                self._load_jac_parts(jac)
            self.module = rates
            self.jac_module = jac
            return rates if rates or not jac else jac
        if os.path.basename(module).startswith('ljac_'):
            self._load_jac_parts(module)
            self.jac_module = module
        else:
            check(self.lib.pyjac_hip_load_module(self._h, os.fsencode(module)))
            self.module = module
        return module

    def _load_jac_parts(self, path):
        """A Jacobian module is five code objects (lanejac.part_paths):
        kernel A at ``path`` and the store kernel of each output layout
        beside it; every one that exists is loaded."""
        from ..libgen import lanejac
        for part in lanejac.part_paths(path):
            if part == path or os.path.exists(part):
                check(self.lib.pyjac_hip_load_module(self._h, os.fsencode(part)))

    def species_names(self):
        """Species names in the kernel's order (bath gas last), from the
        library's copy of the tables (``Kernel::speciesNames``)."""
        return [self.lib.pyjac_hip_species_name(self._h, i).decode()
                for i in range(self.num_species())]

    def reaction_strings(self):
        return [self.lib.pyjac_hip_reaction_string(self._h, i).decode()
                for i in range(self.num_reactions())]

    def vector_width(self):
        return self.lib.pyjac_hip_vector_width(self._h)

    def set_input_check(self, on=True):
        """Check device inputs for NaN / infinity before each device call
        (host arrays are always checked)."""
        check(self.lib.pyjac_hip_set_input_check(self._h, int(bool(on))))

    def nonfinite_states(self, arr, width=None):
        """Per-state finiteness report of a device array in the kernel's
        data order: (flags (int32 device tensor, 1 = the state holds a NaN
        or an infinity), count, first such state or n)."""
        import torch
        width = int(width or self._out_width())
        n = _numel(arr) // width
        if n * width != _numel(arr):
            raise ValueError('array size is not a multiple of {}'.format(width))
        pa, on_dev = _ptr(arr)
        if not on_dev:
            raise ValueError('nonfinite_states takes a device array')
        flags = torch.empty(n, dtype=torch.int32, device=arr.device)
        bad, first = ctypes.c_size_t(), ctypes.c_size_t()
        check(self.lib.pyjac_hip_nonfinite_dev(self._h, n, pa, width,
                                               ctypes.c_void_p(flags.data_ptr()),
                                               ctypes.byref(bad), ctypes.byref(first),
                                               _current_stream_ptr()))
        return flags, int(bad.value), int(first.value)

    def _out_width(self):
        return self.mech.NN

    def num_species(self):
        return self.lib.pyjac_hip_num_species(self._h)

    def num_reactions(self):
        return self.lib.pyjac_hip_num_reactions(self._h)

    def required_working_memory(self):
        return self.lib.pyjac_hip_required_memory(self._h)

    def order_str(self):
        return self.order

    def resize(self, problem_size, work_size=0, do_not_compile=False):
        """Resize the staging buffers for ``problem_size`` states."""
        check(self.lib.pyjac_hip_resize(self._h, int(problem_size), int(work_size)))
        self.problem_size = int(problem_size)

    def set_fd_options(self, mode=0, order=1, rtol=1e-8, atol=1e-15):
        """Finite-difference Jacobian options (a kernel made with
        ``jac_type=JacobianType.finite_difference``): mode 0 forward, 1
        central, 2 backward (``FiniteDifferenceMode``), the order, and the
        tolerances of the steps (finite_difference_jacobian,
        core/create_jacobian.py:4523-4526)."""
        check(self.lib.pyjac_hip_set_fd_options(self._h, int(mode), int(order), float(rtol),
                                                float(atol)))

    def set_variant(self, variant):
        check(self.lib.pyjac_hip_set_variant(self._h, self._variant_slot, int(variant)))

    def autotune(self, param, phi, out, cache=True):
        """Time the kernel families on these resident states (device
        arrays; out receives a result) and keep the fastest.  The choice is
        remembered in a JSON cache keyed by mechanism, options and device
        (``PYJAC_HIP_TUNE_CACHE``, default ~/.cache/pyjac_hip/autotune.json)
        and reapplied by later kernels of the same key.  Returns
        (variant, milliseconds per call)."""
        if self._variant_slot > 1:
            raise InvalidInputSpecificationException('autotune: rates or Jacobian only')
        pp, dp = _ptr(param)
        ph, dph = _ptr(phi)
        po, do = _ptr(out)
        if not (dp and dph and do):
            raise ValueError('autotune needs device arrays')
        v = ctypes.c_int()
        ms = ctypes.c_float()
        check(self.lib.pyjac_hip_autotune(self._h, self._variant_slot, _numel(param), pp, ph, po,
                                          _current_stream_ptr(), ctypes.byref(v), ctypes.byref(ms)))
        occ, chunk, _ = self.tuning()
        if cache:
            _tune_store(self._tune_key(), int(v.value), float(ms.value), occ, chunk)
        return int(v.value), float(ms.value)

    def tuning(self):
        """(occupancy target, F-order chunk states, waves per state) in use
        (0: the default build / size)."""
        occ, ws = ctypes.c_int(), ctypes.c_int()
        chunk = ctypes.c_long()
        check(self.lib.pyjac_hip_get_tuning(self._h, self._variant_slot, ctypes.byref(occ),
                                            ctypes.byref(chunk), ctypes.byref(ws)))
        return int(occ.value), int(chunk.value), int(ws.value)

    def _tune_key(self):
        import hashlib
        try:
            import torch
            dev = torch.cuda.get_device_name(0)
        except Exception:
            dev = 'unknown'
        h = hashlib.sha1('\n'.join(self.mech.species_names + self.mech.reaction_strings)
                         .encode()).hexdigest()[:16]
        return '{}|{}|conp={}|{}|fmt={}|{}'.format(h, self.kernel_type, int(self.conp),
                                                   self.order, self.jac_format, dev)

    def _apply_tuned(self):
        choice = _tune_load().get(self._tune_key())
        if choice is not None:
            # a cached choice this build cannot run keeps the automatic one
            self.lib.pyjac_hip_set_variant(self._h, self._variant_slot, int(choice['variant']))
            self.lib.pyjac_hip_set_tuning(self._h, self._variant_slot,
                                          int(choice.get('occupancy', 0)),
                                          int(choice.get('chunk', 0)))

    def _check_sizes(self, param, phi, out):
        """n from param; phi must hold n * NN values and out n * (width of
        the kernel's output), or ValueError before anything is called."""
        n = _numel(param)
        NN, w = self.mech.NN, self._out_width()
        for a, need, name in ((phi, n * NN, 'phi'), (out, n * w, 'output')):
            if _numel(a) != need:
                raise ValueError('{} holds {} values; {} states need {} ({} per state)'
                                 .format(name, _numel(a), n, need, need // max(n, 1)))
        if isinstance(out, np.ndarray) and not out.flags['WRITEABLE']:
            raise ValueError('output array is not writeable')
        return n

    def _run(self, sized_fn, dev_fn, param, phi, out):
        pp, dp = _ptr(param)
        ph, dph = _ptr(phi)
        po, do = _ptr(out)
        n = self._check_sizes(param, phi, out)
        if dp or dph or do:
            if not (dp and dph and do):
                raise ValueError('mix of host and device arrays')
            stream = _current_stream_ptr()
            check(dev_fn(self._h, n, pp, ph, po, stream))
            return out
        check(sized_fn(self._h, n, pp, _numel(param), ph, _numel(phi), po, _numel(out)))
        self.problem_size = n
        return out


def _tune_path():
    return os.environ.get('PYJAC_HIP_TUNE_CACHE',
                          os.path.join(os.path.expanduser('~'), '.cache', 'pyjac_hip',
                                       'autotune.json'))


def _tune_load():
    import json
    try:
        with open(_tune_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _tune_store(key, variant, ms, occupancy=0, chunk=0):
    import json
    d = _tune_load()
    d[key] = {'variant': variant, 'ms': ms, 'occupancy': occupancy, 'chunk': chunk}
    path = _tune_path()
    os.makedirs(os.path.dirname(path), exist_ok=True)
    tmp = path + '.%d.tmp' % os.getpid()
    with open(tmp, 'w') as f:
        json.dump(d, f, indent=1)
    os.replace(tmp, path)


class Batch(object):
    """Ragged batch over several kernels (mechanisms / option sets) of the
    same kind: ``set(i, param, phi, out)`` with device arrays per member, then
    ``run()`` evaluates all of them -- one launch for the members the
    block-per-state kernels take, captured with the others into one HIP graph
    that later calls replay."""

    def __init__(self, kernels):
        if not kernels:
            raise InvalidInputSpecificationException('kernels')
        kinds = {k._variant_slot for k in kernels}
        if len(kinds) != 1 or kinds.pop() > 1:
            raise InvalidInputSpecificationException('kernels must all be rates or all Jacobian')
        self.kernels = list(kernels)
        self.lib = kernels[0].lib
        arr = (ctypes.c_void_p * len(kernels))(*[k._h.value for k in kernels])
        b = ctypes.c_void_p()
        check(self.lib.pyjac_hip_batch_create(arr, len(kernels), ctypes.byref(b)))
        self._b = b
        self._keep = [None] * len(kernels)

    def set(self, i, param, phi, out):
        pp, _ = _ptr(param)
        ph, _ = _ptr(phi)
        po, _ = _ptr(out)
        self._keep[i] = (param, phi, out)
        check(self.lib.pyjac_hip_batch_set(self._b, i, _numel(param), pp, ph, po))

    def run(self, graph=True):
        check(self.lib.pyjac_hip_batch_run(self._b, self.kernels[0]._variant_slot, int(graph),
                                           _current_stream_ptr()))

    def set_queue(self, on=True):
        """Jacobian batches: evaluate every analytical-Jacobian member, of
        either Jacobian variant, in one launch (``pyjac_hip_batch_set_queue``)."""
        check(self.lib.pyjac_hip_batch_set_queue(self._b, int(bool(on))))

    def slot_failures(self):
        """Contribution-slot claims of the one-launch kernel that found no
        free slot (0 unless the pool is too small)."""
        v = ctypes.c_int()
        check(self.lib.pyjac_hip_batch_slot_failures(self._b, ctypes.byref(v)))
        return int(v.value)

    def __del__(self):
        b = getattr(self, '_b', None)
        if b is not None and b.value:
            self.lib.pyjac_hip_batch_destroy(b)
            self._b = None


def _current_stream_ptr():
    try:
        import torch
        if torch.cuda.is_available():
            return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    except ImportError:
        pass
    return ctypes.c_void_p(0)


class PyJacobianKernel(_KernelBase):
    """Analytical Jacobian: ``kernel(param, phi, jac)``."""
    kernel_type = 'jacobian'
    _variant_slot = 1

    def __call__(self, param, phi, jac):
        return self._run(self.lib.pyjac_hip_jacobian_sized, self.lib.pyjac_hip_jacobian_dev,
                         param, phi, jac)

    def jacobian_size(self):
        return self.lib.pyjac_hip_jacobian_size(self._h)

    def _out_width(self):
        return self.jacobian_size()


class PySpeciesRatesKernel(_KernelBase):
    """Species rates: ``kernel(param, phi, dphi)``."""
    kernel_type = 'species_rates'
    _variant_slot = 0

    def __call__(self, param, phi, dphi):
        return self._run(self.lib.pyjac_hip_species_rates_sized,
                         self.lib.pyjac_hip_species_rates_dev, param, phi, dphi)


class PyChemUtilsKernel(_KernelBase):
    """Thermodynamic properties: ``kernel(phi, b, specific_energy,
    specific_heat)`` (h, cp for CONP; u, cv for CONV)."""
    kernel_type = 'chem_utils'
    _variant_slot = 2

    def __call__(self, phi, b, specific_energy, specific_heat):
        NN, Ns = self.mech.NN, self.mech.Ns
        n = _numel(phi) // NN
        if n * NN != _numel(phi):
            raise ValueError('phi holds {} values, not a multiple of NN = {}'
                             .format(_numel(phi), NN))
        for a, name in ((b, 'b'), (specific_energy, 'specific_energy'),
                        (specific_heat, 'specific_heat')):
            if _numel(a) != n * Ns:
                raise ValueError('{} holds {} values; {} states need {}'
                                 .format(name, _numel(a), n, n * Ns))
            if _ptr(a)[1]:
                raise ValueError('chem utils take host arrays')
        args = [_ptr(a)[0] for a in (phi, b, specific_energy, specific_heat)]
        check(self.lib.pyjac_hip_chem_utils_sized(self._h, n, args[0], _numel(phi), args[1],
                                                  args[2], args[3], n * Ns))
        self.problem_size = n
        return b, specific_energy, specific_heat


def _host_f64(a, name):
    if not isinstance(a, np.ndarray) or a.dtype != np.float64 or not (
            a.flags['C_CONTIGUOUS'] or a.flags['F_CONTIGUOUS']) or not a.flags['WRITEABLE']:
        raise TypeError('{} must be a writeable contiguous float64 numpy array'.format(name))


def read_initial_conditions(filename, num, param, phi, order='C'):
    """Fill ``param[num]`` and ``phi`` (``num * NN`` doubles in ``order``)
    from a ``data.bin`` file of ``num`` records ``[T, param, V|P, n_0 ..
    n_{Ns-2}]`` (read_initial_conditions.cpp.in:41-116; the Cython
    ``read_ics`` of tests/test_utils/read_ic_wrapper.pyx, with the
    reference's ``(param, phi)`` argument order).  Raises
    ``HipRuntimeError`` where the reference exits."""
    _host_f64(param, 'param')
    _host_f64(phi, 'phi')
    num = int(num)
    if param.size < num or (num and phi.size % num) or phi.size < num * 2:
        raise ValueError('param needs {0} entries and phi a multiple of {0}'.format(num))
    nn = phi.size // num if num else 2
    L = load_library()
    check(L.pyjac_hip_read_initial_conditions(os.fsencode(filename), num, nn,
                                              ctypes.c_void_p(param.ctypes.data),
                                              ctypes.c_void_p(phi.ctypes.data),
                                              order.encode()))
    return param, phi


def write_data(filename, arr):
    """Raw dump of a float64 array in its memory order (write_data.hpp:15-32)."""
    a = np.asarray(arr)
    if a.dtype != np.float64 or not (a.flags['C_CONTIGUOUS'] or a.flags['F_CONTIGUOUS']):
        raise TypeError('write_data takes a contiguous float64 array')
    L = load_library()
    check(L.pyjac_hip_write_data(os.fsencode(filename), ctypes.c_void_p(a.ctypes.data), a.size))


# names under which the reference's generated module exposes the classes
PySpecies_RatesKernel = PySpeciesRatesKernel
PyChem_UtilsKernel = PyChemUtilsKernel


def bind(mech, conp=True, order='C', jac_format=JacobianFormat.sparse,
         jac_type=JacobianType.exact, fd_mode=0, fd_order=1, module_dir=None):
    """Kernel classes with the reference's constructor signature
    ``(problem_size, work_size, do_not_compile=False)`` bound to one
    mechanism and option set -- what the reference's generated
    ``pyjac_<lang>`` module provides.  With ``module_dir`` (a build folder
    holding the mechanism-specialised modules) the kernels attach them."""
    opts = dict(mech=mech, conp=conp, order=order, jac_format=jac_format,
                jac_type=jac_type)
    if module_dir:
        opts.update(module='auto', module_dir=module_dir)

    def make(base):
        class Bound(base):
            def __init__(self, problem_size, work_size=0, do_not_compile=False, **kw):
                kw2 = dict(opts)
                kw2.update(kw)
                base.__init__(self, problem_size, work_size, do_not_compile, **kw2)
                if base is PyJacobianKernel and \
                        JacobianType(kw2['jac_type']) == JacobianType.finite_difference:
                    self.set_fd_options(fd_mode, fd_order)
        Bound.__name__ = base.__name__
        return Bound

    return {'PyJacobianKernel': make(PyJacobianKernel),
            'PySpeciesRatesKernel': make(PySpeciesRatesKernel),
            'PySpecies_RatesKernel': make(PySpeciesRatesKernel),
            'PyChemUtilsKernel': make(PyChemUtilsKernel),
            'PyChem_UtilsKernel': make(PyChemUtilsKernel)}


class IntegrateStats(ctypes.Structure):
    _fields_ = [('accepted', ctypes.c_long), ('rejected', ctypes.c_long),
                ('max_steps', ctypes.c_long), ('failed', ctypes.c_long),
                ('launches', ctypes.c_long)]


class PyIntegrator(object):
    """Batched implicit (ROS2) integration of the species-rate ODE on the GPU
    (include/pyjac_integrator.h): every state of ``phi`` (C order ``[n, NN]``,
    a device tensor, updated in place) over ``[0, t_end]`` with its own
    adaptive step.  No counterpart in the reference, whose Jacobian exists for
    such integrators."""

    def __init__(self, mech, conp=True):
        self.lib = load_library()
        L = self.lib
        P = ctypes.c_void_p
        L.pyjac_hip_integrator_create.argtypes = [P, ctypes.c_int, ctypes.POINTER(P)]
        L.pyjac_hip_integrator_destroy.argtypes = [P]
        L.pyjac_hip_integrate_dev.argtypes = [P, ctypes.c_size_t, P, P, ctypes.c_double,
                                              ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                              ctypes.c_int, ctypes.c_long, P,
                                              ctypes.POINTER(IntegrateStats)]
        self.mech = mech
        self.conp = bool(conp)
        h = ctypes.c_void_p()
        check(L.pyjac_hip_integrator_create(ctypes.cast(mech.pointer(), ctypes.c_void_p),
                                            int(self.conp), ctypes.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, '_h', None)
        if h is not None and h.value:
            self.lib.pyjac_hip_integrator_destroy(h)
            self._h = None

    def integrate(self, param, phi, t_end, rtol=1e-6, atol=1e-12, h0=0.0, fixed=False,
                  max_steps=1000000):
        """Integrate in place; returns the step statistics as a dict."""
        pp, dp = _ptr(param)
        ph, dph = _ptr(phi)
        if not (dp and dph):
            raise ValueError('integrate needs device arrays')
        n = _numel(param)
        if _numel(phi) != n * self.mech.NN:
            raise ValueError('phi must hold n * NN values (C order)')
        st = IntegrateStats()
        check(self.lib.pyjac_hip_integrate_dev(self._h, n, pp, ph, float(t_end), float(rtol),
                                               float(atol), float(h0), int(bool(fixed)),
                                               int(max_steps), _current_stream_ptr(),
                                               ctypes.byref(st)))
        return {f: getattr(st, f) for f, _ in IntegrateStats._fields_}
