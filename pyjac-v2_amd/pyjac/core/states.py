"""Seeded thermochemical state samplers.

A state is ``phi = [T, V (CONP) | P (CONV), n_0 .. n_{Ns-2}]`` with the
parameter array holding ``P`` (CONP) or ``V`` (CONV), reference
core/array_creator.py:2051-2067.  Compositions are drawn as mass fractions
from a symmetric Dirichlet distribution, converted to mole fractions with the
molecular weights and to moles with the ideal-gas law at the sampled T, P and
V.
"""

import numpy as np

from .chem_model import RU, PA


def sample_states(mech, n, seed=0, T_range=(800.0, 2400.0),
                  P_range_atm=(1.0, 50.0), alpha=0.5, conp=True, V=1.0,
                  order='C'):
    """Return ``(param, phi)``; ``phi`` is ``[n, NN]`` for order 'C' and
    ``[NN, n]`` for order 'F'."""
    rng = np.random.default_rng(seed)
    Ns = mech.Ns
    T = rng.uniform(T_range[0], T_range[1], size=n)
    logP = rng.uniform(np.log(P_range_atm[0]), np.log(P_range_atm[1]), size=n)
    P = np.exp(logP) * PA
    Y = rng.dirichlet(np.full(Ns, alpha), size=n)
    X = Y / mech.mw[None, :]
    X /= X.sum(axis=1, keepdims=True)
    ntot = P * V / (RU * T)
    moles = X * ntot[:, None]
    phi = np.empty((n, Ns + 1))
    phi[:, 0] = T
    phi[:, 1] = V if conp else P
    phi[:, 2:] = moles[:, :Ns - 1]
    param = P.copy() if conp else np.full(n, V, dtype=np.float64)
    if order == 'F':
        phi = np.ascontiguousarray(phi.T)
    return param, phi


# The workloads of the benchmark configurations (mechanism, CONP/CONV, data
# order, Jacobian format, number of states, temperature range, Dirichlet
# concentration).
WORKLOADS = {
    'h2o2_conp_c_65536': dict(mech='h2o2', conp=True, order='C', jac_format=0,
                              n=65536, T_range=(800.0, 2400.0), alpha=0.5),
    'gri30_conv_f_1048576': dict(mech='gri30', conp=False, order='F',
                                 jac_format=1, n=1048576,
                                 T_range=(1000.0, 2800.0), alpha=0.3),
}


def write_data_bin(path, param, phi):
    """Write states as a ``data.bin`` initial-conditions file: one C-order
    record ``[T, param, V|P, n_0 .. n_{Ns-2}]`` per state, the layout the
    reference's tests write for its reader (tests/test_kernel_compilation.py
    :212-217, tests/test_utils/__init__.py:1571-1576).  ``phi`` is
    ``[n, NN]`` (C order)."""
    phi = np.asarray(phi, dtype=np.float64).reshape(len(param), -1)
    rec = np.concatenate((phi[:, :1], np.asarray(param, dtype=np.float64).reshape(-1, 1),
                          phi[:, 1:]), axis=1)
    with open(path, 'wb') as f:
        np.ascontiguousarray(rec).tofile(f)
    return path


def states_from_mass_fractions(mech, T, P, Y, species, conp=True, V=1.0):
    """``(param, phi)`` (C order) of states given by temperature [K],
    pressure [Pa] and mass fractions ``Y[n, len(species)]`` of the named
    species (the others zero), at volume V [m^3]: moles n_k = X_k P V / (RU T)
    with X from Y and the molecular weights.  Every named species must be in
    the mechanism."""
    T = np.asarray(T, dtype=np.float64).reshape(-1)
    P = np.asarray(P, dtype=np.float64).reshape(-1)
    Y = np.asarray(Y, dtype=np.float64).reshape(len(T), -1)
    names = list(mech.species_names)
    full = np.zeros((len(T), mech.Ns))
    for c, name in enumerate(species):
        full[:, names.index(str(name))] = Y[:, c]
    X = full / mech.mw[None, :]
    X /= X.sum(axis=1, keepdims=True)
    ntot = P * V / (RU * T)
    phi = np.empty((len(T), mech.NN))
    phi[:, 0] = T
    phi[:, 1] = V if conp else P
    phi[:, 2:] = (X * ntot[:, None])[:, :mech.Ns - 1]
    param = P.copy() if conp else np.full(len(T), V, dtype=np.float64)
    return param, phi
