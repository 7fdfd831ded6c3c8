"""Seeded synthetic mechanisms for the large-mechanism workloads.

``synthetic_mechanism(n_species, n_reactions, seed, **mix)`` writes a Chemkin
mechanism (elements, species, NASA-7 thermo, reactions with LOW / TROE / SRI /
PLOG / CHEB auxiliary lines) that goes through the same parser as the real
mechanisms:

* species S0001 .. S{n-1} with C/H/O compositions drawn from a grid (C <= 5,
  H <= 12, O <= 5; isomers once the grid is used up), plus an inert AR bath
  gas that is last and never reacts (so it does not densify the Jacobian
  pattern);
* element-balanced reactions (products drawn to match the reactants' atoms),
  so that mass is conserved -- the Jacobian's temperature and extra-variable
  rows eliminate the bath gas through mass conservation;
* NASA-7 heat-capacity coefficients perturbed (5 %) from the GRI-Mech 3.0
  species, with the enthalpy and entropy constants set so that h/RT lies in
  [-3, 3] and s/R in [28, 32] at 1000 K (random reactions then have
  equilibrium constants in physical ranges);
* Arrhenius A log-uniform in [1e6, 1e16] (cgs), b ~ U(-2, 2),
  Ea ~ U(0, 60) kcal/mol, 2-3 species on each side (2 -> 1 for half of the
  pressure-dependent reactions), 90 % reversible;
* the requested fractions of third-body, Troe, SRI, PLOG (pressures
  log-uniform in [0.1, 100] atm) and Chebyshev (6 x 4 over T in [300, 3000] K
  and P in [0.01, 100] atm) reactions.

``WORKLOADS`` holds the two configurations of the benchmark workloads:
S150 (150 species, 900 reactions, 10 % third body, 10 % Troe, 5 % PLOG x 4)
and S500 (500 species, 3000 reactions, 8 % Troe, 4 % SRI, 6 % PLOG x 5,
3 % Chebyshev 6 x 4).
"""

import os
import tempfile

import numpy as np

GENERATOR_VERSION = 3

WORKLOADS = {
    's150': dict(n_species=150, n_reactions=900, seed=150, thd=0.10, troe=0.10, sri=0.0,
                 plog=0.05, plog_n=4, cheb=0.0),
    's500': dict(n_species=500, n_reactions=3000, seed=500, thd=0.0, troe=0.08, sri=0.04,
                 plog=0.06, plog_n=5, cheb=0.03),
}


def _thermo_entry(name, comp, lo, hi, Tlo=300.0, Thi=5000.0, Tmid=1000.0):
    elems = ''.join('{:<2}{:>3}'.format(e, n) for e, n in comp[:4])
    line1 = '{:<18}{:<6}{:<20}G{:>10.3f}{:>10.3f}{:>10.3f}    1'.format(
        name, 'SYN', elems, Tlo, Thi, Tmid)
    f = lambda v: '{:15.8E}'.format(v)
    line2 = ''.join(f(v) for v in hi[:5]) + '    2'
    line3 = ''.join(f(v) for v in (hi[5], hi[6], lo[0], lo[1], lo[2])) + '    3'
    line4 = ''.join(f(v) for v in lo[3:7]) + ' ' * 15 + '    4'
    return '\n'.join((line1, line2, line3, line4))


def _templates():
    """NASA-7 (low, high) coefficient pairs of the GRI-Mech 3.0 species."""
    from .._paths import mechanism_files
    from .mech_interpret import load_mechanism
    _, specs, _ = load_mechanism(*mechanism_files('gri30'))
    return [(np.array(s.lo), np.array(s.hi)) for s in specs if s.name not in ('AR', 'N2', 'HE')]


def synthetic_mechanism(n_species, n_reactions, seed, thd=0.1, troe=0.1, sri=0.0, plog=0.05,
                        plog_n=4, cheb=0.0, **unused):
    """Chemkin text of a seeded synthetic mechanism (thermo included)."""
    rng = np.random.default_rng(seed)
    tmpl = _templates()
    names = ['S{:04d}'.format(i + 1) for i in range(n_species - 1)] + ['AR']
    # element compositions (C, H, O) from a grid, isomers once it is used up
    grid = [(c, h, o) for c in range(6) for h in range(13) for o in range(6) if c + h + o]
    order = rng.permutation(len(grid))
    comps = [grid[order[i % len(grid)]] for i in range(n_species - 1)]
    by_comp = {}
    for i, c in enumerate(comps):
        by_comp.setdefault(c, []).append(i)
    comp_arr = np.array(comps)
    out = ['ELEMENTS', 'C H O AR', 'END', 'SPECIES']
    for i in range(0, len(names), 8):
        out.append(' '.join(names[i:i + 8]))
    out += ['END', 'THERMO ALL', '   300.000  1000.000  5000.000']
    for i, nm in enumerate(names):
        if nm == 'AR':
            lo = hi = np.array([2.5, 0, 0, 0, 0, -745.375, 4.366])
            comp = [('AR', 1)]
        else:
            lo0, hi0 = tmpl[rng.integers(len(tmpl))]
            pert = 1.0 + 0.05 * rng.standard_normal(7)
            lo, hi = lo0 * pert, hi0 * pert
            # enthalpy and entropy constants: h/RT in [-3, 3] and s/R in
            # [28, 32] at 1000 K for both ranges, so that equilibrium
            # constants of random reactions stay in physical ranges
            h_t, s_t = rng.uniform(-3.0, 3.0), rng.uniform(28.0, 32.0)
            for a in (lo, hi):
                T = 1000.0
                a[5] = T * (h_t - (a[0] + a[1] * T / 2 + a[2] * T**2 / 3 + a[3] * T**3 / 4 +
                                   a[4] * T**4 / 5))
                a[6] = s_t - (a[0] * np.log(T) + a[1] * T + a[2] * T**2 / 2 + a[3] * T**3 / 3 +
                              a[4] * T**4 / 4)
            comp = [(e, int(c)) for e, c in zip(('C', 'H', 'O'), comps[i]) if c]
        out.append(_thermo_entry(nm, comp, lo, hi))
    out += ['END', 'REACTIONS']
    reactive = n_species - 1   # AR never reacts
    kinds = rng.choice(['thd', 'troe', 'sri', 'plog', 'cheb', 'plain'], size=n_reactions,
                       p=[thd, troe, sri, plog, cheb, 1.0 - thd - troe - sri - plog - cheb])
    seen = set()

    def text(sp):
        terms = []
        for s in sorted(set(sp)):
            c = sp.count(s)
            terms.append(('{} '.format(c) if c > 1 else '') + names[s])
        return ' + '.join(terms)

    def balanced(nl, nr):
        """Reactants (nl species) and products (nr species) with the same
        element totals, so that every reaction conserves elements and mass."""
        while True:
            reac = sorted(rng.choice(reactive, size=nl, replace=True).tolist())
            rem = comp_arr[reac].sum(axis=0)
            prod = []
            for _ in range(nr - 1):
                for _try in range(60):
                    c = int(rng.integers(reactive))
                    left = rem - comp_arr[c]
                    if (left >= 0).all() and left.any():
                        prod.append(c)
                        rem = left
                        break
                else:
                    break
            if len(prod) != nr - 1:
                continue
            last = by_comp.get(tuple(int(x) for x in rem))
            if not last:
                continue
            prod = sorted(prod + [last[int(rng.integers(len(last)))]])
            if prod != reac:
                return reac, prod

    def efficiencies():
        if rng.random() < 0.2:
            return ''
        pick = rng.choice(reactive, size=rng.integers(2, 7), replace=False)
        return ' '.join('{}/{:.2f}/'.format(names[p], rng.uniform(0.5, 5.0)) for p in pick)

    for kind in kinds:
        while True:
            if kind in ('thd', 'troe', 'sri') and rng.random() < 0.5:
                nl, nr = 2, 1          # recombination
            else:
                nl, nr = int(rng.integers(2, 4)), int(rng.integers(2, 4))
            kl, kr = balanced(nl, nr)
            kl, kr = tuple(kl), tuple(kr)
            if (kl, kr) not in seen and (kr, kl) not in seen:
                break
        seen.add((kl, kr))
        lhs, rhs = text(list(kl)), text(list(kr))
        arrow = '<=>' if rng.random() < 0.9 else '=>'
        A = 10.0 ** rng.uniform(6, 16)
        b = rng.uniform(-2, 2)
        Ea = rng.uniform(0, 60000.0)
        if kind == 'thd':
            out.append('{} + M {} {} + M   {:.4E} {:.4f} {:.2f}'.format(lhs, arrow, rhs, A, b, Ea))
            eff = efficiencies()
            if eff:
                out.append(eff)
        elif kind in ('troe', 'sri'):
            out.append('{} (+M) {} {} (+M)   {:.4E} {:.4f} {:.2f}'.format(lhs, arrow, rhs, A, b, Ea))
            out.append('LOW / {:.4E} {:.4f} {:.2f} /'.format(A * 10.0 ** rng.uniform(2, 6),
                                                              rng.uniform(-3, 0),
                                                              rng.uniform(0, 20000.0)))
            if kind == 'troe':
                pars = [rng.uniform(0.1, 0.9), rng.uniform(50, 500), rng.uniform(500, 5000)]
                if rng.random() < 0.6:
                    pars.append(rng.uniform(1000, 10000))
                out.append('TROE / ' + ' '.join('{:.4f}'.format(p) for p in pars) + ' /')
            else:
                pars = [rng.uniform(0.1, 1.0), rng.uniform(100, 1000), rng.uniform(100, 3000)]
                if rng.random() < 0.5:
                    pars += [rng.uniform(0.5, 2.0), rng.uniform(-0.5, 0.5)]
                out.append('SRI / ' + ' '.join('{:.4f}'.format(p) for p in pars) + ' /')
            eff = efficiencies()
            if eff:
                out.append(eff)
        elif kind == 'plog':
            out.append('{} {} {}   {:.4E} {:.4f} {:.2f}'.format(lhs, arrow, rhs, A, b, Ea))
            for p in np.sort(10.0 ** rng.uniform(-1, 2, size=plog_n)):
                out.append('PLOG / {:.6g} {:.4E} {:.4f} {:.2f} /'.format(
                    p, A * 10.0 ** rng.uniform(-1, 1), b + rng.uniform(-0.3, 0.3),
                    Ea * rng.uniform(0.8, 1.2)))
        elif kind == 'cheb':
            out.append('{} (+M) {} {} (+M)   1.0 0.0 0.0'.format(lhs, arrow, rhs))
            out.append('TCHEB / 300.0 3000.0 /  PCHEB / 0.01 100.0 /')
            out.append('CHEB / 6 4 /')
            coef = rng.uniform(-0.5, 0.5, size=(6, 4)) / (1.0 + np.add.outer(np.arange(6),
                                                                               np.arange(4)))
            coef[0, 0] = rng.uniform(8.0, 13.0)
            flat = coef.reshape(-1)
            for i in range(0, flat.size, 4):
                out.append('CHEB / ' + ' '.join('{:.6e}'.format(v) for v in flat[i:i + 4]) + ' /')
        else:
            out.append('{} {} {}   {:.4E} {:.4f} {:.2f}'.format(lhs, arrow, rhs, A, b, Ea))
    out.append('END')
    return '\n'.join(out) + '\n'


def workload_file(name, directory=None):
    """Path of the Chemkin file of workload ``name`` ('s150', 's500'),
    generated on first use into ``directory`` (default: a cache folder in the
    system temporary directory)."""
    cfg = WORKLOADS[name]
    directory = directory or os.path.join(tempfile.gettempdir(),
                                          'pyjac_synthetic_v{}'.format(GENERATOR_VERSION))
    os.makedirs(directory, exist_ok=True)
    path = os.path.join(directory, '{}.inp'.format(name))
    if not os.path.exists(path):
        text = synthetic_mechanism(**cfg)
        tmp = path + '.{}.tmp'.format(os.getpid())
        with open(tmp, 'w') as f:
            f.write(text)
        os.replace(tmp, path)
    return path
