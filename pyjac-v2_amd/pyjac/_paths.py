"""Locations of the bundled mechanism inputs."""
import os

PKG = os.path.dirname(os.path.abspath(__file__))
# package data: the reference's bundled mechanisms (data/h2o2.inp,
# data/grimech30.inp + thermo30.dat, pyjac/tests/test.inp and the .cti files)
# and their Cantera YAML renderings
MECH_DIR = os.path.join(PKG, 'data', 'mechanisms')

MECHANISMS = {
    'h2o2': ('h2o2.inp', None),
    'gri30': ('grimech30.inp', 'thermo30.dat'),
    'test': ('test.inp', None),
}


def mechanism_files(name):
    """(mechanism, thermo) paths of a bundled mechanism, or of a seeded
    synthetic one ('s150', 's500'), generated on first use."""
    if name in ('s150', 's500'):
        from .core.synthetic import workload_file
        return workload_file(name), None
    mech, therm = MECHANISMS[name]
    return (os.path.join(MECH_DIR, mech),
            os.path.join(MECH_DIR, therm) if therm else None)
