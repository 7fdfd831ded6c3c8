"""pyJac species-rate / analytical-Jacobian backend for MI355X (gfx950).

Front-end (Chemkin parsing, rate tables, Jacobian sparsity) in
``pyjac.core``; the HIP library is driven through ``pyjac.pywrap``; the
``--lang hip`` command line lives in ``pyjac.__main__``.
"""

__version__ = '2.0.0b0+hip'
