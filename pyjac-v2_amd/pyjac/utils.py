"""Command line of the generator, restricted to the HIP target.

Mirrors the reference's ``utils.get_parser`` / ``utils.create``
(``pyjac/utils.py:838-1099``): the same flags, names and defaults.  Options
that only shape the reference's generated code (vector width / depth,
unrolling, explicit SIMD, OpenCL platform, atomics, fused / split rate
kernels, unique pointers, memory limits, reaction sorting) are accepted for
compatibility and do not change the HIP library, which is one
mechanism-independent build driven by tables.
"""

import argparse
import logging

from .core.enum_types import (JacobianFormat, JacobianType, KernelType,
                              RateSpecialization, reaction_sorting)

langs = ['hip']


class EnumType(object):
    """argparse type accepting the names of an enum (reference utils.py:800)."""

    def __init__(self, enum):
        self.enum = enum

    def __call__(self, value):
        if isinstance(value, self.enum):
            return value
        try:
            return self.enum[value.lower()]
        except KeyError:
            raise argparse.ArgumentTypeError(
                '{} is not one of {}'.format(value, str(self)))

    def __str__(self):
        return ', '.join(e.name for e in self.enum)


def get_parser(argv=None):
    parser = argparse.ArgumentParser(description='pyJac: analytical chemical Jacobian '
                                     'and species-rate kernels (HIP target).')
    parser.add_argument('-l', '--lang', type=str, choices=langs, required=True,
                        help='Target language; this build provides "hip".')
    parser.add_argument('-i', '--input', type=str, required=True,
                        help='Input mechanism filename (e.g., mech.dat).')
    parser.add_argument('-t', '--thermo', type=str, default=None,
                        help='Thermodynamic database filename, or nothing if in mechanism.')
    parser.add_argument('-w', '--width', type=int, default=None,
                        help='Accepted for compatibility (HIP kernels choose their own layout).')
    parser.add_argument('-d', '--depth', type=int, default=None,
                        help='Accepted for compatibility.')
    parser.add_argument('-se', '--explicit_simd', default=None, action='store_true',
                        help='Accepted for compatibility.')
    parser.add_argument('-si', '--implicit_simd', action='store_false', dest='explicit_simd',
                        help='Accepted for compatibility.')
    parser.add_argument('-unr', '--unroll', type=int, default=None,
                        help='Accepted for compatibility.')
    parser.add_argument('-b', '--build_path', default='./out/',
                        help='The folder to write the kernel module and library to.')
    parser.add_argument('-ls', '--last_species', type=str, default=None,
                        help='The species to put last (bath gas); default: the first of '
                             'N2, AR and HE in the mechanism.')
    parser.add_argument('-k', '--kernel_type', type=EnumType(KernelType),
                        default=KernelType.jacobian,
                        help='Kernel to expose: {}'.format(EnumType(KernelType)))
    parser.add_argument('-p', '--platform', default='', type=str,
                        help='Accepted for compatibility.')
    parser.add_argument('-o', '--data_order', default='C', type=str, choices=['C', 'F'],
                        help="Data ordering, 'C' (row-major) or 'F' (column-major).")
    parser.add_argument('-rs', '--rate_specialization', default=RateSpecialization.hybrid,
                        type=EnumType(RateSpecialization),
                        help='Specialisation of the Arrhenius rates: {}'.format(
                            EnumType(RateSpecialization)))
    parser.add_argument('-rk', '--fused_rate_kernels', default=False, action='store_true',
                        help='Accepted for compatibility.')
    parser.add_argument('-rn', '--split_rop_net_kernels', default=False, action='store_true',
                        help='Accepted for compatibility.')
    parser.add_argument('-conv', '--constant_volume', dest='conp', action='store_false',
                        help='Use the constant-volume assumption (default: constant pressure).')
    parser.add_argument('-nad', '--no_atomic_doubles', dest='use_atomic_doubles',
                        action='store_false', help='Accepted for compatibility.')
    parser.add_argument('-nai', '--no_atomic_ints', dest='use_atomic_ints',
                        action='store_false', help='Accepted for compatibility.')
    parser.add_argument('-jt', '--jac_type', type=EnumType(JacobianType),
                        default=JacobianType.exact,
                        help='Jacobian type: {} (the HIP target computes exact '
                             'Jacobians)'.format(EnumType(JacobianType)))
    parser.add_argument('-jf', '--jac_format', type=EnumType(JacobianFormat),
                        default=JacobianFormat.sparse,
                        help='Jacobian storage: {} (sparse is CRS for C order, CCS for '
                             'F order)'.format(EnumType(JacobianFormat)))
    parser.add_argument('-up', '--unique_pointers', default=False, action='store_true',
                        help='Accepted for compatibility.')
    parser.add_argument('-m', '--memory_limits', type=str, default='',
                        help='Accepted for compatibility.')
    parser.add_argument('--verbose', action='store_const', const=logging.DEBUG,
                        dest='loglevel', default=logging.INFO,
                        help='Increase verbosity.')
    parser.add_argument('--reaction_sorting', type=EnumType(reaction_sorting),
                        default=reaction_sorting.none, help='Accepted for compatibility.')
    return parser.parse_args(argv)


def create(argv=None, **kwargs):
    """Parse the command line and write the kernel module (reference
    utils.create, utils.py:1070-1099)."""
    args = get_parser(argv)
    vars(args).update(kwargs)
    logging.basicConfig(level=args.loglevel)
    from .core.create_jacobian import create_jacobian
    return create_jacobian(lang=args.lang, mech_name=args.input, therm_name=args.thermo,
                           build_path=args.build_path, last_spec=args.last_species,
                           kernel_type=args.kernel_type, data_order=args.data_order,
                           rate_specialization=args.rate_specialization, conp=args.conp,
                           jac_type=args.jac_type, jac_format=args.jac_format)
