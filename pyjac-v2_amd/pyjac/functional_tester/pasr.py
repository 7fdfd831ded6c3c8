"""Partially stirred reactor (PaSR) states without Cantera.

Restates the particle method of the reference's
``functional_tester/partially_stirred_reactor.py`` (``run_simulation``
:745-948, ``mix_substep`` :548-576, ``select_pairs`` :639-675, ``inflow``
:678-712, ``equivalence_ratio`` :383-525, ``parse_input_file`` :951-1017):
``num_part`` particles of constant-pressure gas, each a composition vector
``[h, Y_0 .. Y_{Ns-1}]`` (mass-specific enthalpy and mass fractions), are
exchanged with inflow streams at the residence time, paired and mixed
towards each other at the mixing time, and reacted as adiabatic
constant-pressure reactors between mixing sub-steps.

Differences from the reference, by design:

* the random number generator is seeded (``numpy.random.default_rng(seed)``)
  so a run is reproducible; the reference reseeds from the clock;
* the reaction sub-step is a callable ``react(T, P, Y, dt) -> (T, Y)``
  over all particles at once; :func:`gpu_reactor` builds one from this
  package's batched ROS2 integrator (the reference advances one Cantera
  ``IdealGasConstPressureReactor`` per particle);
* the pilot stream's adiabatic equilibrium (Cantera's ``equilibrate('HP')``
  in the reference) is reached by reacting the fresh mixture at constant
  pressure until its temperature stops changing -- the same state, since
  every reaction of a closed adiabatic reactor then balances;
* temperatures follow from (h, Y) by Newton iteration on the NASA-7
  polynomials of the mechanism tables.

Kept as in the reference: the time-step and sub-step rules, the running
inflow selection, the pair selection and its swaps, Python's
round-half-to-even for the particle counts, and the particle rotation after
pairing -- including its effect when no pair is selected (the last particle
is copied into particle 1).
"""

import numpy as np

from ..core.chem_model import PA, RU

__all__ = ['MixtureThermo', 'equivalence_ratio', 'gpu_reactor', 'equilibrate', 'run_simulation',
           'parse_input_file', 'particle_states']


class MixtureThermo(object):
    """Ideal-gas mixture properties from a mechanism's NASA-7 tables
    (``MechTables``): molar enthalpy / heat capacity per species, mass-specific
    mixture values, and the temperature of a given enthalpy."""

    def __init__(self, mech):
        self.mech = mech
        self.Ns = mech.Ns
        self.W = np.asarray(mech.mw, dtype=np.float64)
        self.lo = np.asarray(mech.arrays['a_lo'], dtype=np.float64).reshape(self.Ns, 7)
        self.hi = np.asarray(mech.arrays['a_hi'], dtype=np.float64).reshape(self.Ns, 7)
        self.Tmid = np.asarray(mech.arrays['T_mid'], dtype=np.float64).reshape(self.Ns)
        self.names = list(mech.species_names)

    def _coeffs(self, T):
        T = np.asarray(T, dtype=np.float64).reshape(-1)
        return np.where((T[:, None] < self.Tmid[None, :])[..., None], self.lo[None], self.hi[None])

    def h_molar(self, T):
        """Species enthalpies [J/kmol], ``[n, Ns]``."""
        T = np.asarray(T, dtype=np.float64).reshape(-1, 1)
        a = self._coeffs(T)
        return RU * (T * (a[..., 0] + T * (a[..., 1] / 2 + T * (a[..., 2] / 3 + T * (
            a[..., 3] / 4 + T * a[..., 4] / 5)))) + a[..., 5])

    def cp_molar(self, T):
        """Species heat capacities at constant pressure [J/kmol/K], ``[n, Ns]``."""
        T = np.asarray(T, dtype=np.float64).reshape(-1, 1)
        a = self._coeffs(T)
        return RU * (a[..., 0] + T * (a[..., 1] + T * (a[..., 2] + T * (a[..., 3] + T * a[..., 4]))))

    def h_mass(self, T, Y):
        return (np.asarray(Y) / self.W * self.h_molar(T)).sum(axis=1)

    def cp_mass(self, T, Y):
        return (np.asarray(Y) / self.W * self.cp_molar(T)).sum(axis=1)

    def mean_mw(self, Y):
        return 1.0 / (np.asarray(Y) / self.W).sum(axis=1)

    def temperature(self, h, Y, T0=None, tol=1e-12, max_iter=100):
        """Temperatures at which the mixtures ``Y`` have enthalpy ``h``."""
        h = np.asarray(h, dtype=np.float64).reshape(-1)
        T = np.full(h.shape, 1500.0) if T0 is None else np.array(T0, dtype=np.float64).reshape(-1)
        for _ in range(max_iter):
            dT = (self.h_mass(T, Y) - h) / self.cp_mass(T, Y)
            dT = np.clip(dT, -500.0, 500.0)
            T = np.maximum(T - dT, 100.0)
            if np.all(np.abs(dT) <= tol * T):
                return T
        raise RuntimeError('temperature from enthalpy did not converge')

    def mole_to_mass(self, X):
        X = np.asarray(X, dtype=np.float64).reshape(-1, self.Ns)
        m = X * self.W
        return m / m.sum(axis=1, keepdims=True)

    def composition(self, spec):
        """Mole fractions (``[Ns]``) from a {species: amount} mapping."""
        X = np.zeros(self.Ns)
        for name, amt in spec.items():
            if name not in self.names:
                raise ValueError('species {} not in the mechanism'.format(name))
            X[self.names.index(name)] = float(amt)
        return X / X.sum()

    def n_atoms(self, species, element):
        sp = self.mech.specs[self.names.index(species)]
        return sum(int(n) for e, n in sp.elem if e.lower() == element.lower())


def equivalence_ratio(thermo, eq_ratio, fuel, oxidizer, complete_products):
    """{species: mole fraction} of the fuel-oxidizer mixture at an
    equivalence ratio (partially_stirred_reactor.py:383-525): the oxidizer
    amount supplies exactly the oxygen that burns the fuel's carbon and
    hydrogen to the complete products (the O per C and per H those products
    carry), and the fuel is scaled by ``eq_ratio``.  Fuel and oxidizer
    amounts summing above one are normalised first."""
    fuel = dict(fuel)
    oxidizer = dict(oxidizer)
    for mix in (fuel, oxidizer):
        total = sum(mix.values())
        if total > 1.0:
            for k in mix:
                mix[k] /= total
    atoms = lambda mix, el: sum(thermo.n_atoms(sp, el) * amt for sp, amt in mix.items())
    elements = set(e.upper() for s in thermo.mech.specs for e, n in s.elem if n)
    prod_atoms = {el: {sp: thermo.n_atoms(sp, el) for sp in complete_products} for el in elements}
    for el in elements:
        in_prod = sum(prod_atoms[el].values()) > 0
        in_reac = atoms(fuel, el) > 0 or atoms(oxidizer, el) > 0
        if in_prod != in_reac:
            raise ValueError('element {} must appear both in the fuel + oxidizer and in the '
                             'complete products'.format(el))

    def o_per_atom(el):
        total = sum(prod_atoms.get(el, {}).values())
        if not total:
            return 0.0
        return sum(prod_atoms['O'][sp] for sp in complete_products
                   if prod_atoms[el][sp] > 0) / total
    o_needed = atoms(fuel, 'C') * o_per_atom('C') + atoms(fuel, 'H') * o_per_atom('H') \
        - atoms(fuel, 'O')
    o_mult = o_needed / atoms(oxidizer, 'O')
    amounts = {sp: amt * o_mult for sp, amt in oxidizer.items()}
    for sp, amt in fuel.items():
        amounts[sp] = amounts.get(sp, 0.0) + amt * eq_ratio
    total = sum(amounts.values())
    return {sp: v / total for sp, v in amounts.items()}


def _to_phi(thermo, T, P, Y, V=1.0):
    """CONP state vectors ``[n, NN]`` (T, V, moles of all but the last
    species) of mixtures at (T, P, Y)."""
    T = np.asarray(T, dtype=np.float64).reshape(-1)
    Y = np.asarray(Y, dtype=np.float64).reshape(-1, thermo.Ns)
    rho = P * thermo.mean_mw(Y) / (RU * T)
    moles = rho[:, None] * V * Y / thermo.W
    phi = np.empty((T.size, thermo.Ns + 1))
    phi[:, 0] = T
    phi[:, 1] = V
    phi[:, 2:] = moles[:, :-1]
    return phi


def _from_phi(thermo, P, phi):
    """(T, Y) of CONP state vectors; the last species' moles follow from the
    ideal-gas law, as in the kernels."""
    T, V = phi[:, 0], phi[:, 1]
    n = np.empty((phi.shape[0], thermo.Ns))
    n[:, :-1] = phi[:, 2:]
    n[:, -1] = P * V / (RU * T) - phi[:, 2:].sum(axis=1)
    m = np.maximum(n, 0.0) * thermo.W
    return T.copy(), m / m.sum(axis=1, keepdims=True)


def gpu_reactor(mech, rtol=1e-6, atol=1e-14):
    """``react(T, P, Y, dt) -> (T, Y)``: every particle an adiabatic
    constant-pressure reactor advanced over ``dt`` by the batched ROS2
    integrator on the GPU (``pywrap.PyIntegrator``)."""
    import torch
    from ..pywrap import PyIntegrator
    thermo = MixtureThermo(mech)
    integ = PyIntegrator(mech, conp=True)

    def react(T, P, Y, dt):
        phi = torch.from_numpy(_to_phi(thermo, T, P, Y)).cuda()
        param = torch.full((phi.shape[0],), float(P), dtype=torch.float64, device='cuda')
        stats = integ.integrate(param, phi, dt, rtol=rtol, atol=atol)
        if stats['failed']:
            raise RuntimeError('{} particle(s) did not reach the end of the reaction step'
                               .format(stats['failed']))
        return _from_phi(thermo, P, phi.cpu().numpy())
    return react


def phi_reactor(mech, integrate):
    """``react`` from a host integrator ``integrate(param, phi, dt) -> phi``
    over CONP state vectors."""
    thermo = MixtureThermo(mech)

    def react(T, P, Y, dt):
        phi = _to_phi(thermo, T, P, Y)
        out = integrate(np.full(phi.shape[0], float(P)), phi, dt)
        return _from_phi(thermo, P, np.asarray(out).reshape(phi.shape))
    return react


def _settle(react, T, P, Y, rtol, t_max):
    """React one mixture over growing intervals until its temperature stops
    changing."""
    T = np.array([float(T)])
    Y = np.asarray(Y, dtype=np.float64).reshape(1, -1)
    dt, elapsed = 1e-5, 0.0
    while elapsed < t_max:
        T1, Y = react(T, P, Y, dt)
        elapsed += dt
        settled = abs(T1[0] - T[0]) <= rtol * T1[0] and elapsed >= 1e-2
        T = T1
        if settled:
            return T[0], Y[0]
        dt = min(dt * 4.0, 1.0)
    raise RuntimeError('no equilibrium within {} s'.format(t_max))


def equilibrate(thermo, react, T, P, Y, rtol=1e-12, t_max=50.0, T_kick=2000.0):
    """Adiabatic constant-pressure equilibrium of one mixture (Cantera's
    ``equilibrate('HP')`` in the reference).  The equilibrium depends only on
    the enthalpy, the pressure and the element amounts, so it is reached from
    any composition that shares them and reacts: a cold mixture is first
    burnt from ``T_kick`` (same elements, higher enthalpy), and that burnt
    composition, brought back to the mixture's own enthalpy, is then
    reacted to rest."""
    Y = np.asarray(Y, dtype=np.float64).reshape(-1)
    h0 = thermo.h_mass([T], Y[None])
    if T < T_kick:
        T1, Y1 = _settle(react, T_kick, P, Y, 1e-6, t_max)
        T = thermo.temperature(h0, Y1[None], T0=[T1])[0]
        Y = Y1
    return _settle(react, T, P, Y, rtol, t_max)


class _Stream(object):
    def __init__(self, comp, T, flow):
        self.comp = comp
        self.T = T
        self.flow = flow
        self.xflow = 0.0


def _inflow(streams):
    """Index of the stream of the next inflowing particle: the one whose
    running flow is largest after adding every stream's flow; the total is
    then taken off it (partially_stirred_reactor.py:678-712)."""
    best, fl_max, total = None, 0.0, 0.0
    for i, s in enumerate(streams):
        s.xflow += s.flow
        total += s.flow
        if s.xflow > fl_max:
            fl_max, best = s.xflow, i
    streams[best].xflow -= total
    return best


def _swap(comp, i, j):
    comp[[i, j]] = comp[[j, i]]


def _select_pairs(rng, comp, num_pairs, num_skip=0):
    """Move ``num_pairs`` random pairs (particles 2i, 2i+1) to the end of the
    list, before the last ``num_skip`` pairs, each pair's order flipped with
    probability 1/2 (partially_stirred_reactor.py:639-675)."""
    npart = comp.shape[0]
    for k in range(num_pairs):
        avail = npart // 2 - k - num_skip
        if avail <= 0:
            raise ValueError('more pairs requested than particles')
        i = 2 * int(rng.integers(avail))
        j = i + 1
        if rng.random() > 0.5:
            _swap(comp, i, j)
        last = npart - 2 * (k + num_skip + 1)
        _swap(comp, i, last)
        _swap(comp, j, last + 1)


def run_simulation(mech, case, init_temp, pres, eq_ratio, fuel, oxidizer,
                   complete_products=('CO2', 'H2O', 'N2'), num_part=100, tau_res=10e-3,
                   tau_mix=1e-3, tau_pair=1e-3, num_res=10, seed=0, react=None,
                   pilot=None, log=None):
    """Run the PaSR and return ``[steps + 1, num_part, Ns + 3]`` particle data
    (time, T, P, mass fractions with negatives zeroed), as the reference's
    ``run_simulation``.  ``pres`` is in atm.  ``react`` defaults to
    :func:`gpu_reactor`; ``pilot`` may give the pilot's equilibrium (T, Y) to
    skip its computation; ``log(time, mean_T)`` is called after every step
    (the reference prints the same two columns)."""
    thermo = MixtureThermo(mech)
    if react is None:
        react = gpu_reactor(mech)
    P = pres * PA
    rng = np.random.default_rng(seed)
    Ns = thermo.Ns

    dt_max = 0.1 * min(tau_res, tau_pair)
    num_substeps = 1 + int(dt_max / (0.040 * tau_mix))
    time_end = num_res * tau_res

    reactants = thermo.composition(equivalence_ratio(thermo, eq_ratio, fuel, oxidizer,
                                                     complete_products))

    def stream(X, T):
        Y = thermo.mole_to_mass(X)
        return np.concatenate([thermo.h_mass([T], Y), Y[0]])

    case = case.lower()
    streams = []
    if case == 'premixed':
        streams.append(_Stream(stream(reactants, init_temp), init_temp, 0.95))
        pilot_flow = 0.05
    elif case == 'non-premixed':
        streams.append(_Stream(stream(thermo.composition({'O2': 0.21, 'N2': 0.79}), init_temp),
                               init_temp, 0.85))
        streams.append(_Stream(stream(thermo.composition(fuel), init_temp), init_temp, 0.05))
        pilot_flow = 0.1
    else:
        raise ValueError('case must be "premixed" or "non-premixed"')
    if pilot is None:
        Yr = thermo.mole_to_mass(reactants)[0]
        pilot = equilibrate(thermo, react, init_temp, P, Yr)
    T_pilot, Y_pilot = pilot
    pilot_comp = np.concatenate([thermo.h_mass([T_pilot], Y_pilot[None]), Y_pilot])
    streams.append(_Stream(pilot_comp, T_pilot, pilot_flow))

    comp = np.tile(pilot_comp, (num_part, 1))
    T = np.full(num_part, float(T_pilot))

    def save(t):
        rec = np.empty((num_part, Ns + 3))
        rec[:, 0] = t
        rec[:, 1] = T
        rec[:, 2] = P
        rec[:, 3:] = np.maximum(comp[:, 1:], 0.0)
        return rec

    data = [save(0.0)]
    time, part_out, part_pair = 0.0, 0.0, 0.0
    while time < time_end:
        dt = min(dt_max, time_end - time)
        part_out += num_part * dt / tau_res
        npart_out = int(round(part_out))
        part_out -= npart_out
        num_fl_pairs = 2 * npart_out
        # temperatures travel with their particles through the swaps
        state = np.concatenate([comp, T[:, None]], axis=1)
        _select_pairs(rng, state, num_fl_pairs)
        for i in range(npart_out):
            s = streams[_inflow(streams)]
            state[num_part + 1 - 2 * (i + 1), :-1] = s.comp
            state[num_part + 1 - 2 * (i + 1), -1] = s.T
        part_pair += 0.5 * num_part * dt / tau_pair
        num_pairs = int(round(part_pair))
        part_pair -= num_pairs
        _select_pairs(rng, state, num_pairs, num_fl_pairs)
        last = state[-1].copy()
        for i in range(1, 2 * num_pairs - 2, 2):
            state[num_part - i] = state[num_part - i - 2]
        state[(1 - 2 * num_pairs) % num_part] = last
        comp, T = state[:, :-1], state[:, -1].copy()

        dt_sub = dt / num_substeps
        decay = 0.5 * (1.0 - np.exp(-2.0 * dt_sub / tau_mix))
        for _ in range(num_substeps):
            # pairwise mixing of (h, Y) towards each other
            a, b = comp[0::2][:num_part // 2], comp[1::2][:num_part // 2]
            delt = (a - b) * decay
            comp[0:2 * (num_part // 2):2] -= delt
            comp[1:2 * (num_part // 2):2] += delt
            T = thermo.temperature(comp[:, 0], comp[:, 1:], T0=T)
            T, Y = react(T, P, comp[:, 1:], dt_sub)
            comp = np.concatenate([thermo.h_mass(T, Y)[:, None], Y], axis=1)
        time += dt
        data.append(save(time))
        if log is not None:
            log(time, float(T.mean()))
    return np.stack(data)


def parse_input_file(path):
    """PaSR parameters from a YAML file with the reference's keys
    (partially_stirred_reactor.py:951-1017), defaults filled in."""
    import yaml
    with open(path) as f:
        pars = yaml.safe_load(f)
    if pars.get('case') not in ('premixed', 'non-premixed'):
        raise ValueError('case must be "premixed" or "non-premixed"')
    for key in ('temperature', 'pressure', 'fuel', 'oxidizer', 'complete products'):
        if not pars.get(key):
            raise ValueError('{} must be given'.format(key))
    if not pars.get('equivalence ratio') or pars['equivalence ratio'] < 0.0:
        raise ValueError('equivalence ratio must be given and positive')
    for key, default in (('number of particles', 100), ('residence time', 10e-3),
                         ('mixing time', 1e-3), ('pairing time', 1e-3),
                         ('number of residence times', 5)):
        if not pars.get(key):
            pars[key] = default
    return pars


def run_from_input(mech, pars, seed=0, react=None, log=None):
    """:func:`run_simulation` with the parameters of :func:`parse_input_file`."""
    return run_simulation(mech, pars['case'], pars['temperature'], pars['pressure'],
                          pars['equivalence ratio'], pars['fuel'], pars['oxidizer'],
                          pars['complete products'], pars['number of particles'],
                          pars['residence time'], pars['mixing time'], pars['pairing time'],
                          pars['number of residence times'], seed=seed, react=react, log=log)


def particle_states(mech, data, conp=True, V=1.0, unique=True):
    """Kernel inputs ``(param, phi)`` (C order) from PaSR particle data: the
    distinct (T, P, Y) states of all particles and time steps, as moles at
    volume ``V`` (CONP: param = P, phi[1] = V; CONV: param = V, phi[1] = P)."""
    thermo = MixtureThermo(mech)
    flat = np.asarray(data, dtype=np.float64).reshape(-1, data.shape[-1])[:, 1:]
    if unique:
        _, idx = np.unique(flat, axis=0, return_index=True)
        flat = flat[np.sort(idx)]
    T, P, Y = flat[:, 0], flat[:, 1], flat[:, 2:]
    Y = Y / Y.sum(axis=1, keepdims=True)
    rho = P * thermo.mean_mw(Y) / (RU * T)
    moles = rho[:, None] * V * Y / thermo.W
    phi = np.empty((T.size, thermo.Ns + 1))
    phi[:, 0] = T
    phi[:, 1] = V if conp else P
    phi[:, 2:] = moles[:, :-1]
    param = P.copy() if conp else np.full(T.size, V)
    return param, phi
