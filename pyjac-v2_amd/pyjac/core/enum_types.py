"""Enumerations shared by the mechanism front-end, the table builders and the
HIP runtime.

The integer values are part of the table contract with the device code (the
falloff ``fall_type`` field is ``reaction_type - reaction_type.fall``, the
third-body ``thd_type`` field is ``thd_body_type``), so they follow
``pyjac/core/enum_types.py:4-150`` of the reference exactly.
"""

from enum import Enum, IntEnum


class reaction_type(Enum):
    """Rate-constant family of a reaction (reference enum_types.py:4)."""
    elementary = 1
    thd = 2
    fall = 3
    chem = 4
    plog = 5
    cheb = 6

    def __int__(self):
        return self.value


class thd_body_type(Enum):
    """Form of the third-body concentration (reference enum_types.py:20)."""
    none = 0
    mix = 1
    species = 2
    unity = 3

    def __int__(self):
        return self.value


class falloff_form(Enum):
    """Falloff blending function (reference enum_types.py:37)."""
    none = 0
    lind = 1
    troe = 2
    sri = 3

    def __int__(self):
        return self.value


class reversible_type(Enum):
    """Reversibility of a reaction (reference enum_types.py:51)."""
    non_reversible = 1
    explicit = 2
    non_explicit = 3

    def __int__(self):
        return self.value


class reaction_sorting(Enum):
    """Optional reaction re-ordering (reference enum_types.py:65)."""
    none = 0
    simd = 1


class RateSpecialization(IntEnum):
    """Arrhenius specialisation level (reference enum_types.py:73).

    * ``fixed``: every rate is ``exp(logA + b logT - Ta/T)``
    * ``hybrid``: ``A``, ``A T^b`` (integer b) or the full exponential
    * ``full``: five forms, see :func:`pyjac.core.rate_tables.assign_rates`
    """
    fixed = 0
    hybrid = 1
    full = 2


class JacobianType(IntEnum):
    """Jacobian flavour (reference enum_types.py:85)."""
    exact = 0
    approximate = 1
    finite_difference = 2


class JacobianFormat(IntEnum):
    """Full (dense NN x NN) or sparse (CRS for C order, CCS for F order)."""
    full = 0
    sparse = 1


class FiniteDifferenceMode(IntEnum):
    forward = 0
    central = 1
    backward = 2


class KernelType(Enum):
    """Kernel produced by the generator (reference enum_types.py:129)."""
    chem_utils = 1
    species_rates = 2
    jacobian = 3
    dummy = 4

    def __int__(self):
        return self.value


class DriverType(Enum):
    queue = 0
    lockstep = 1
