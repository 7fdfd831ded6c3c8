"""Error types raised by the front-end and the HIP runtime bindings.

The reference reports most input errors with ``logger.error`` followed by
``sys.exit`` (e.g. ``core/mech_interpret.py:206``); here they are exceptions
so that a caller embedding the library keeps control.
"""


class PyjacError(Exception):
    """Base class of every error raised by this package."""


class MechanismError(PyjacError):
    """The mechanism or thermo input could not be interpreted."""


class InvalidInputSpecificationException(PyjacError):
    """An option combination is not supported (reference
    ``core/exceptions.py``)."""


class HipRuntimeError(PyjacError):
    """A call through the C ABI returned a non-zero status."""

    def __init__(self, status, message):
        super(HipRuntimeError, self).__init__(
            'pyjac_hip status {}: {}'.format(status, message))
        self.status = status
