"""The generator's entry point for the HIP target.

The reference's ``create_jacobian`` (``core/create_jacobian.py:5212-5420``)
parses the mechanism, builds the tables and emits C/OpenCL source plus a
Python wrapper module.  The HIP backend needs no generated code: the one
library takes any mechanism as tables at run time.  So for ``lang='hip'``
this writes into ``build_path``

* ``pyjac_hip.py`` -- the wrapper module, exposing ``PyJacobianKernel``,
  ``PySpeciesRatesKernel`` and ``PyChemUtilsKernel`` with the reference's
  constructor ``(problem_size, work_size, do_not_compile=False)``, bound to
  this mechanism and these options;
* the mechanism (and thermo) file it was made from, and ``pyjac_hip.json``
  with the options;
* ``libpyjac_hip.so``, copied from the package build (built on demand);
* the mechanism-specialised code objects (``pyjac.libgen.lanegen`` species
  rates, ``pyjac.libgen.lanejac`` analytical Jacobian: the per-mechanism
  code the reference generates, compiled for gfx950), for mechanisms of up to
  64 species, which the wrapper module's kernels attach; copied from the
  package's cache when already built for these tables, else compiled
  (minutes for GRI-Mech 3.0; ``specialise=False`` skips them);
* ``tables.pjt`` (the mechanism tables for C and C++ callers),
  ``pyjac_hip_main.cfg`` and the executable ``pyjac_hip_main``, the
  counterpart of the reference's generated main (kernel.cpp.in:645-748):
  ``pyjac_hip_main <build_path> <problem_size> <work_size>`` reads
  ``data_filename``, runs the kernel and prints the timing line.
"""

import json
import logging
import os
import shutil

from .enum_types import (FiniteDifferenceMode, JacobianFormat, JacobianType, KernelType,
                         RateSpecialization)
from .exceptions import InvalidInputSpecificationException
from .mech_tables import MechTables

MODULE = '''"""pyJac kernels for {mech} (HIP target), generated by pyjac.

Options: conp={conp}, order={order}, jac_format={jac_format}, jac_type={jac_type},
rate_specialization={rate_spec}, kernel_type={kernel_type}.
"""
import json as _json
import os as _os

_here = _os.path.dirname(_os.path.abspath(__file__))
with open(_os.path.join(_here, 'pyjac_hip.json')) as _f:
    _opts = _json.load(_f)
if _os.path.exists(_os.path.join(_here, 'libpyjac_hip.so')):
    _os.environ.setdefault('PYJAC_HIP_LIB', _os.path.join(_here, 'libpyjac_hip.so'))

from pyjac.core.mech_tables import MechTables as _MechTables  # noqa: E402
from pyjac.pywrap import bind as _bind  # noqa: E402

mech = _MechTables.from_file(
    _os.path.join(_here, _opts['mech']),
    _os.path.join(_here, _opts['thermo']) if _opts['thermo'] else None,
    last_spec=_opts['last_spec'], rate_spec=_opts['rate_spec'], jac_type=_opts['jac_type'])
globals().update(_bind(mech, conp=_opts['conp'], order=_opts['order'],
                       jac_format=_opts['jac_format'], jac_type=_opts['jac_type'],
                       fd_mode=_opts.get('fd_mode', 0), fd_order=_opts.get('fd_order', 1),
                       module_dir=_here))
species_names = list(mech.species_names)
reaction_strings = list(mech.reaction_strings)
'''


def create_jacobian(lang, mech_name=None, therm_name=None, build_path='./out/',
                    last_spec=None, kernel_type=KernelType.jacobian, data_order='C',
                    rate_specialization=RateSpecialization.hybrid, conp=True,
                    jac_type=JacobianType.exact, jac_format=JacobianFormat.sparse,
                    fd_order=1, fd_mode=FiniteDifferenceMode.forward,
                    data_filename='data.bin', for_validation=False,
                    build_library=True, specialise=True, **ignored):
    """Write the HIP kernel module for one mechanism; returns its path."""
    if lang != 'hip':
        raise NotImplementedError('this build provides the hip target only')
    if mech_name is None:
        raise InvalidInputSpecificationException('mech_name')
    if data_order not in ('C', 'F'):
        raise InvalidInputSpecificationException('data_order')
    jac_type = JacobianType(jac_type)
    # parse and build the tables now, so that a bad mechanism fails here
    mech = MechTables.from_file(mech_name, therm_name, last_spec=last_spec,
                                rate_spec=rate_specialization, jac_type=jac_type)
    os.makedirs(build_path, exist_ok=True)
    mech_base = os.path.basename(mech_name)
    shutil.copy2(mech_name, os.path.join(build_path, mech_base))
    therm_base = None
    if therm_name:
        therm_base = os.path.basename(therm_name)
        shutil.copy2(therm_name, os.path.join(build_path, therm_base))
    opts = dict(mech=mech_base, thermo=therm_base, last_spec=last_spec,
                rate_spec=int(RateSpecialization(rate_specialization)),
                jac_type=int(jac_type), jac_format=int(JacobianFormat(jac_format)),
                fd_mode=int(FiniteDifferenceMode(fd_mode)), fd_order=int(fd_order),
                conp=bool(conp), order=data_order, kernel_type=KernelType(kernel_type).name)
    with open(os.path.join(build_path, 'pyjac_hip.json'), 'w') as f:
        json.dump(opts, f, indent=1)
    # the executable's inputs (kernel.cpp.in:645-748): tables and options
    mech.save_binary(os.path.join(build_path, 'tables.pjt'))
    kt = KernelType(kernel_type)
    with open(os.path.join(build_path, 'pyjac_hip_main.cfg'), 'w') as f:
        f.write('tables tables.pjt\n')
        f.write('kernel {}\n'.format('jacobian' if kt == KernelType.jacobian else 'species_rates'))
        f.write('conp {}\norder {}\njac_format {}\njac_type {}\n'.format(
            int(bool(conp)), data_order, opts['jac_format'], opts['jac_type']))
        f.write('data {}\nfor_validation {}\n'.format(data_filename, int(bool(for_validation))))
    path = os.path.join(build_path, 'pyjac_hip.py')
    with open(path, 'w') as f:
        f.write(MODULE.format(mech=mech.name, conp=opts['conp'], order=data_order,
                              jac_format=opts['jac_format'], jac_type=opts['jac_type'],
                              rate_spec=opts['rate_spec'], kernel_type=opts['kernel_type']))
    if build_library:
        from ..libgen import LIB_PATH, MAIN_NAME, build_library as _build
        if not os.path.exists(LIB_PATH):
            _build()
        shutil.copy2(LIB_PATH, os.path.join(build_path, os.path.basename(LIB_PATH)))
        main = os.path.join(os.path.dirname(LIB_PATH), MAIN_NAME)
        if os.path.exists(main):
            shutil.copy2(main, os.path.join(build_path, MAIN_NAME))
    if specialise:
        _specialise(mech, build_path, jac_type)
    logging.getLogger(__name__).info('wrote %s (%d species, %d reactions)', path, mech.Ns,
                                     mech.Nr)
    return path


def _specialise(mech, build_path, jac_type):
    """The mechanism-specialised modules of these tables into build_path:
    copied from the package's cache, or compiled there."""
    from ..libgen import lanegen, lanejac
    log = logging.getLogger(__name__)
    if mech.Ns > min(lanegen.MAX_SPECIES, lanejac.MAX_SPECIES):
        log.info('%d species: no specialised modules (the table-driven kernels run)', mech.Ns)
        return
    jobs = [(lanegen, [lanegen.module_path(mech)])]
    if jac_type == JacobianType.exact:
        jobs.append((lanejac, lanejac.part_paths(lanejac.module_path(mech))))
    for gen, cached in jobs:
        dest = os.path.join(build_path, os.path.basename(cached[0]))
        if all(os.path.exists(p) for p in cached):
            for p in cached:
                shutil.copy2(p, os.path.join(build_path, os.path.basename(p)))
        else:
            gen.build(mech, dest)
