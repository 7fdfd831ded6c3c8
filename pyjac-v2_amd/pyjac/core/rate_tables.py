"""Flat mechanism tables and the Jacobian sparsity pattern.

``assign_rates`` restates ``pyjac/core/rate_subs.py:74-537`` and
``determine_jac_inds`` restates ``pyjac/core/create_jacobian.py:68-282``: they
turn the parsed reactions into the integer/float tables every kernel reads
(Arrhenius parameters and their hybrid specialisation, falloff/Troe/SRI,
PLOG/Chebyshev, third-body efficiencies, reversible map, per-reaction
(species, nu) lists, NASA-7 coefficients) and into the exact non-zero pattern
of the analytical Jacobian in flat / CRS / CCS form.

:func:`named_tables` then derives the arrays under the names the reference's
generated code uses (``simple_rtype_2_map``, ``thd_mask``, ``jac_row_inds``
...), which is what the golden fixtures taken from the reference's generated
headers are compared against.
"""

import numpy as np

from .enum_types import (RateSpecialization, JacobianType, reaction_type,
                         thd_body_type, falloff_form)

KINT = np.int32


def get_nu(isp, rxn):
    """Net stoichiometric coefficient of species ``isp`` (reference
    utils.py:516-546)."""
    if isp in rxn.prod and isp in rxn.reac:
        nu = rxn.prod_nu[rxn.prod.index(isp)] - rxn.reac_nu[rxn.reac.index(isp)]
        if nu == 0:
            return 0
    elif isp in rxn.prod:
        nu = rxn.prod_nu[rxn.prod.index(isp)]
    elif isp in rxn.reac:
        nu = -rxn.reac_nu[rxn.reac.index(isp)]
    else:
        return 0
    return nu


def _is_integer(val):
    try:
        return float(val).is_integer()
    except (TypeError, ValueError):
        return False


def _select(reacs, matchers):
    hits = [(i, r) for i, r in enumerate(reacs)
            if any(r.match(m) for m in matchers)]
    if not hits:
        return [], np.empty(0, dtype=KINT), 0
    idx, rx = zip(*hits)
    return list(rx), np.array(idx, dtype=KINT), len(rx)


def _specialize(rates, rate_spec, fall=False):
    """Arrhenius specialisation (reference rate_subs.py:245-293).

    Returns ``(rtype, A, b, Ta, fall_types)`` where ``A`` holds ``log(A)``
    except for types 0/1 where it holds ``A`` itself."""
    full = rate_spec == RateSpecialization.full
    fixed = rate_spec == RateSpecialization.fixed
    n = len(rates)
    rtype = np.zeros(n, dtype=KINT)
    ftype = np.zeros(n, dtype=KINT) if fall else None
    A = np.zeros(n)
    b = np.zeros(n)
    Ta = np.zeros(n)
    for i, rxn in enumerate(rates):
        if (rxn.high or rxn.low) and fall:
            if rxn.high:
                Ai, bi, Tai = rxn.high
                ftype[i] = int(reaction_type.chem)
            else:
                Ai, bi, Tai = rxn.low
                ftype[i] = int(reaction_type.fall)
        else:
            Ai, bi, Tai = rxn.A, rxn.b, rxn.E
        A[i] = np.log(Ai)
        b[i] = bi
        Ta[i] = Tai
        if fixed:
            rtype[i] = 0
            continue
        if bi == 0 and Tai == 0:
            A[i] = Ai
            rtype[i] = 0
        elif bi == int(bi) and bi and Tai == 0:
            A[i] = Ai
            rtype[i] = 1
        elif Tai == 0 and bi != 0:
            rtype[i] = 2
        elif bi == 0 and Tai != 0:
            rtype[i] = 3
        else:
            rtype[i] = 4
        if not full:
            rtype[i] = rtype[i] if rtype[i] <= 1 else 2
    if fall:
        ftype -= int(reaction_type.fall)
    return rtype, A, b, Ta, ftype


def assign_rates(reacs, specs, rate_spec=RateSpecialization.hybrid):
    """Build the rate tables, keyed like the reference's return value."""
    rate_spec = RateSpecialization(rate_spec)
    Ns = len(specs)
    rev_map = np.array([i for i, r in enumerate(reacs) if r.rev], dtype=KINT)

    nu_sum, net_num_spec, net_spec, net_nu = [], [], [], []
    reac_has_ns, ns_nu = [], []
    for i_rxn, rxn in enumerate(reacs):
        slist = sorted(set(list(rxn.reac) + list(rxn.prod)))
        net_spec.extend(slist)
        net_num_spec.append(len(slist))
        for sp in slist:
            r = rxn.reac_nu[rxn.reac.index(sp)] if sp in rxn.reac else 0
            p = rxn.prod_nu[rxn.prod.index(sp)] if sp in rxn.prod else 0
            net_nu.extend([p, r])
        nu_sum.append(sum(get_nu(sp, rxn) for sp in slist))
        ns_r = rxn.reac_nu[rxn.reac.index(Ns - 1)] if Ns - 1 in rxn.reac else 0
        ns_p = rxn.prod_nu[rxn.prod.index(Ns - 1)] if Ns - 1 in rxn.prod else 0
        if ns_r or ns_p:
            reac_has_ns.append(i_rxn)
            ns_nu.extend([ns_p, ns_r])
    allint = all(_is_integer(nu) for nu in net_nu)
    ntype = KINT if allint else np.float64
    nu_sum = np.array(nu_sum, dtype=ntype)
    net_nu = np.array(net_nu, dtype=ntype)
    ns_nu = np.array(ns_nu, dtype=ntype)
    reac_has_ns = np.array(reac_has_ns, dtype=KINT)
    net_num_spec = np.array(net_num_spec, dtype=KINT)
    net_spec = np.array(net_spec, dtype=KINT)

    spec_to_reac, spec_nu, spec_reac_count, spec_list = [], [], [], []
    for isp in range(Ns):
        lst = [(ir, get_nu(isp, r)) for ir, r in enumerate(reacs)]
        lst = [x for x in lst if x[1]]
        if lst:
            rl, nl = zip(*lst)
            spec_to_reac.extend(rl)
            spec_nu.extend(nl)
            spec_reac_count.append(len(rl))
            spec_list.append(isp)

    simple_rate, simple_map, num_simple = _select(
        reacs, [reaction_type.elementary, reaction_type.thd,
                reaction_type.fall, reaction_type.chem])
    s_type, s_A, s_b, s_Ta, _ = _specialize(simple_rate, rate_spec)

    plog_reacs, plog_map, num_plog = _select(reacs, [reaction_type.plog])
    num_pressures = np.array([len(p.plog_par) for p in plog_reacs], dtype=KINT)
    plog_params = [p.plog_par for p in plog_reacs]

    cheb_reacs, cheb_map, num_cheb = _select(reacs, [reaction_type.cheb])
    cheb_n_pres = np.array([c.cheb_n_pres for c in cheb_reacs], dtype=KINT)
    cheb_n_temp = np.array([c.cheb_n_temp for c in cheb_reacs], dtype=KINT)
    cheb_plim = np.array([c.cheb_plim for c in cheb_reacs])
    cheb_tlim = np.array([c.cheb_tlim for c in cheb_reacs])
    cheb_coeff = [c.cheb_par for c in cheb_reacs]

    fall_reacs, fall_map, num_fall = _select(
        reacs, [reaction_type.fall, reaction_type.chem])
    f_type, f_A, f_b, f_Ta, fall_types = _specialize(fall_reacs, rate_spec,
                                                     fall=True)
    blend = np.array([next(int(y) for y in r.type
                           if isinstance(y, falloff_form))
                      for r in fall_reacs], dtype=KINT)
    lind_map = np.where(blend == int(falloff_form.lind))[0].astype(KINT)
    sri_map = np.where(blend == int(falloff_form.sri))[0].astype(KINT)
    sri_par = [list(reacs[fall_map[i]].sri_par) for i in sri_map]
    for ps in sri_par:
        if len(ps) != 5:
            ps.extend([1, 0])
    if sri_par:
        sri = [np.array(x, dtype=np.float64) for x in zip(*sri_par)]
    else:
        sri = [np.empty(0) for _ in range(5)]
    troe_map = np.where(blend == int(falloff_form.troe))[0].astype(KINT)
    troe_par = [list(reacs[fall_map[i]].troe_par) for i in troe_map]
    for ps in troe_par:
        if len(ps) != 4:
            ps.append(0)
    if troe_par:
        troe_a, troe_T3, troe_T1, troe_T2 = [
            np.array(x, dtype=np.float64) for x in zip(*troe_par)]
        troe_T1 = 1. / troe_T1
        troe_T3 = 1. / troe_T3
    else:
        troe_a = troe_T3 = troe_T1 = troe_T2 = np.empty(0)

    thd_reacs, thd_map, num_thd = _select(
        reacs, [reaction_type.fall, reaction_type.chem, reaction_type.thd])
    thd_type = np.array([next(int(y) for y in r.type
                              if isinstance(y, thd_body_type))
                         for r in thd_reacs], dtype=KINT)
    last = Ns - 1
    thd_spec_num, thd_spec, thd_eff, thd_has_ns, thd_ns_eff = [], [], [], [], []
    for i, r in enumerate(thd_reacs):
        if r.match(thd_body_type.species):
            thd_spec_num.append(1)
            thd_spec.append(r.pdep_sp)
            thd_eff.append(1)
            if r.pdep_sp == last:
                thd_has_ns.append(i)
                thd_ns_eff.append(1)
        elif r.match(thd_body_type.unity):
            thd_spec_num.append(0)
        else:
            thd_spec_num.append(len(r.thd_body_eff))
            sp, ef = zip(*r.thd_body_eff)
            thd_spec.extend(sp)
            thd_eff.extend(ef)
            k = next((k for k, s in enumerate(sp) if s == last), None)
            if k is not None:
                thd_has_ns.append(i)
                thd_ns_eff.append(ef[k])

    a_lo = np.zeros((Ns, 7))
    a_hi = np.zeros((Ns, 7))
    T_mid = np.zeros(Ns)
    for k, sp in enumerate(specs):
        a_lo[k, :] = sp.lo[:]
        a_hi[k, :] = sp.hi[:]
        T_mid[k] = sp.Trange[1]

    pp_cheb_coeff = np.empty(0)
    pp_cheb_plim = np.empty(0)
    pp_cheb_tlim = np.empty(0)
    if num_cheb:
        pp_cheb_coeff = np.zeros((num_cheb, int(np.max(cheb_n_temp)),
                                  int(np.max(cheb_n_pres))))
        for i, p in enumerate(cheb_coeff):
            pp_cheb_coeff[i, :cheb_n_temp[i], :cheb_n_pres[i]] = p
        pp_cheb_plim = np.log(np.array(cheb_plim, dtype=np.float64))
        pp_cheb_tlim = 1. / np.array(cheb_tlim, dtype=np.float64)

    pp_plog = np.empty(0)
    maxP = None
    if num_plog:
        maxP = int(np.max(num_pressures))
        pp_plog = np.zeros((4, num_plog, maxP))
        for m in range(4):
            for i, nP in enumerate(num_pressures):
                for j in range(nP):
                    pp_plog[m, i, j] = plog_params[i][j][m]
        with np.errstate(divide='ignore'):
            pp_plog[0] = np.log(pp_plog[0])
            pp_plog[1] = np.log(pp_plog[1])
        pp_plog[np.isinf(pp_plog)] = 0

    mws = np.array([sp.mw for sp in specs])
    return {
        'simple': {'A': s_A, 'b': s_b, 'Ta': s_Ta, 'type': s_type,
                   'num': num_simple, 'map': simple_map},
        'plog': {'map': plog_map, 'num': num_plog, 'num_P': num_pressures,
                 'params': plog_params, 'max_P': maxP,
                 'post_process': {'params': pp_plog}},
        'cheb': {'map': cheb_map, 'num': num_cheb, 'num_P': cheb_n_pres,
                 'num_T': cheb_n_temp, 'params': cheb_coeff,
                 'Tlim': cheb_tlim, 'Plim': cheb_plim,
                 'post_process': {'params': pp_cheb_coeff,
                                  'Plim': pp_cheb_plim,
                                  'Tlim': pp_cheb_tlim}},
        'fall': {'map': fall_map, 'num': num_fall, 'ftype': fall_types,
                 'blend': blend, 'A': f_A, 'b': f_b, 'Ta': f_Ta,
                 'type': f_type,
                 'sri': {'map': sri_map, 'num': sri_map.size,
                         'a': sri[0], 'b': sri[1], 'c': sri[2],
                         'd': sri[3], 'e': sri[4]},
                 'troe': {'map': troe_map, 'num': troe_map.size,
                          'a': troe_a, 'T3': troe_T3, 'T1': troe_T1,
                          'T2': troe_T2},
                 'lind': {'map': lind_map, 'num': lind_map.size}},
        'thd': {'map': thd_map, 'num': num_thd, 'type': thd_type,
                'spec_num': np.array(thd_spec_num, dtype=KINT),
                'spec': np.array(thd_spec, dtype=KINT),
                'eff': np.array(thd_eff, dtype=np.float64),
                'has_ns': np.array(thd_has_ns, dtype=KINT),
                'eff_ns': np.array(thd_ns_eff, dtype=np.float64)},
        'fwd': {'map': np.arange(len(reacs)), 'num': len(reacs)},
        'rev': {'map': rev_map, 'num': len(rev_map)},
        'net': {'num_reac_to_spec': net_num_spec, 'nu_sum': nu_sum,
                'nu': net_nu, 'reac_to_spec': net_spec, 'allint': allint},
        'net_per_spec': {'reac_count': np.array(spec_reac_count, dtype=KINT),
                         'nu': np.array(spec_nu, dtype=ntype),
                         'reacs': np.array(spec_to_reac, dtype=KINT),
                         'map': np.array(spec_list, dtype=KINT),
                         'allint': allint},
        'Nr': len(reacs),
        'Ns': Ns,
        'thermo': {'a_lo': a_lo, 'a_hi': a_hi, 'T_mid': T_mid},
        'mws': mws,
        'mw_post': mws[:-1] / mws[-1],
        'reac_has_ns': reac_has_ns,
        'ns_nu': ns_nu,
    }


def _offsets(counts):
    counts = np.asarray(counts)
    return np.concatenate(([0], np.cumsum(counts))).astype(KINT)


def determine_jac_inds(reacs, specs, rate_spec=RateSpecialization.hybrid,
                       jacobian_type=JacobianType.exact):
    """Exact non-zero pattern of the Jacobian (reference
    create_jacobian.py:68-282), returned inside the ``assign_rates`` dict
    under ``'jac_inds'`` with ``flat_C``, ``flat_F``, ``crs`` and ``ccs``."""
    val = assign_rates(reacs, specs, rate_spec)
    NN = len(specs) + 1
    inds = [(0, x) for x in range(NN)] + [(1, x) for x in range(NN)]

    nz_specs = val['net_per_spec']['map']
    rxn_count = _offsets(val['net_per_spec']['reac_count'])
    rxn_maps = val['net_per_spec']['reacs']
    rxn_spec = val['net']['reac_to_spec']
    rxn_off = _offsets(val['net']['num_reac_to_spec'])
    has_ns = set(val['reac_has_ns'].tolist())
    thd_has_ns = set(val['thd']['has_ns'].tolist())
    thd_spec = val['thd']['spec']
    thd_eff = val['thd']['eff']
    thd_map = val['thd']['map']
    rev_set = set(val['rev']['map'].tolist())
    nu = val['net']['nu']
    thd_off = _offsets(val['thd']['spec_num'])
    ns = val['Ns'] - 1
    thd_index = {int(r): i for i, r in enumerate(thd_map)}

    for n_i, spec in enumerate(nz_specs):
        if spec == len(specs) - 1:
            continue
        row = spec + 2
        nonzero = {0, 1}

        def add(slist):
            nonzero.update(x + 2 for x in slist if x + 2 < NN)

        for rxn in rxn_maps[rxn_count[n_i]:rxn_count[n_i + 1]]:
            ti = thd_index.get(int(rxn))
            if (rxn in has_ns or (ti is not None and ti in thd_has_ns)) \
                    and jacobian_type != JacobianType.approximate:
                add(range(NN))
                break
            dspecs = rxn_spec[rxn_off[rxn]:rxn_off[rxn + 1]]
            if rxn not in rev_set:
                rn = nu[2 * rxn_off[rxn]:2 * rxn_off[rxn + 1]]
                is_prod = [rn[2 * i] != 0 for i in range(len(dspecs))]
                is_reac = [rn[2 * i + 1] != 0 for i in range(len(dspecs))]
                si = int(np.where(dspecs == spec)[0][0])
                if not (is_prod[si] and is_reac[si]):
                    dspecs = [x for i, x in enumerate(dspecs) if is_reac[i]]
            add(dspecs)
            if ti is not None:
                last_eff = 1.0
                tsp = thd_spec[thd_off[ti]:thd_off[ti + 1]]
                tef = thd_eff[thd_off[ti]:thd_off[ti + 1]]
                if ns in tsp:
                    last_eff = tef[np.where(tsp == ns)[0][0]]
                add([x for i, x in enumerate(tsp) if tef[i] != last_eff])
        inds.extend((row, x) for x in sorted(nonzero))

    inds = np.asarray(inds, dtype=KINT)
    rows, cols = inds[:, 0], inds[:, 1]
    order_F = np.lexsort((rows, cols))
    inds_F = inds[order_F]
    row_ptr = [np.count_nonzero(rows == i) for i in range(NN)]
    col_ptr = [np.count_nonzero(cols == i) for i in range(NN)]
    col_ind = np.concatenate([cols[rows == i] for i in range(NN)])
    row_ind = np.concatenate([rows[cols == i] for i in range(NN)])
    val['jac_inds'] = {
        'flat_C': inds,
        'flat_F': inds_F,
        'crs': {'col_ind': col_ind.astype(KINT), 'row_ptr': _offsets(row_ptr)},
        'ccs': {'row_ind': row_ind.astype(KINT), 'col_ptr': _offsets(col_ptr)},
    }
    return val


def named_tables(info):
    """Arrays under the names of the reference's generated sources
    (``examples/opencl/*/species_rates.oclh``, ``chem_utils.oclh``,
    ``jacobian.oclh``), for comparison with the golden fixtures."""
    Ns, Nr = info['Ns'], info['Nr']
    out = {}
    s = info['simple']
    out['simple_A'] = s['A']
    out['simple_beta'] = s['b']
    out['simple_Ta'] = s['Ta']
    for t in range(5):
        inds = np.where(s['type'] == t)[0].astype(KINT)
        if inds.size:
            out['simple_rtype_{}_inds'.format(t)] = inds
            out['simple_rtype_{}_map'.format(t)] = s['map'][inds]
    simple_mask = -np.ones(Nr, dtype=KINT)
    simple_mask[s['map']] = np.arange(s['num'], dtype=KINT)
    out['simple_mask'] = simple_mask
    f = info['fall']
    if f['num']:
        out['fall_A'] = f['A']
        out['fall_beta'] = f['b']
        out['fall_Ta'] = f['Ta']
        out['fall_type'] = f['ftype']
        out['fall_map'] = f['map']
        thd_of = {int(r): i for i, r in enumerate(info['thd']['map'])}
        out['fall_to_thd_map'] = np.array([thd_of[int(r)] for r in f['map']],
                                          dtype=KINT)
        for key in ('lind', 'troe', 'sri'):
            if f[key]['num']:
                out[key + '_map'] = f[key]['map']
        if f['troe']['num']:
            for k_out, k_in in (('troe_a', 'a'), ('troe_T1', 'T1'),
                                ('troe_T2', 'T2'), ('troe_T3', 'T3')):
                out[k_out] = f['troe'][k_in]
        if f['sri']['num']:
            for k in 'abcde':
                out['sri_' + k] = f['sri'][k]
    t = info['thd']
    if t['num']:
        out['thd_map'] = t['map']
        thd_mask = -np.ones(Nr, dtype=KINT)
        thd_mask[t['map']] = np.arange(t['num'], dtype=KINT)
        out['thd_mask'] = thd_mask
        out['thd_offset'] = _offsets(t['spec_num'])
        out['thd_spec'] = t['spec']
        out['thd_eff'] = t['eff']
        out['thd_type'] = t['type']
        fall_set = set(f['map'].tolist()) if f['num'] else set()
        out['thd_only_map'] = np.array(
            [i for i, r in enumerate(t['map']) if int(r) not in fall_set],
            dtype=KINT)
    n = info['net']
    out['net_reac_to_spec_offsets'] = _offsets(n['num_reac_to_spec'])
    out['rxn_to_spec'] = n['reac_to_spec']
    out['reac_to_spec_nu'] = n['nu']
    out['nu_sum'] = n['nu_sum']
    out['rev_map'] = info['rev']['map']
    rev_mask = -np.ones(Nr, dtype=KINT)
    rev_mask[info['rev']['map']] = np.arange(info['rev']['num'], dtype=KINT)
    out['rev_mask'] = rev_mask
    out['mw_factor'] = info['mw_post']
    out['T_mid'] = info['thermo']['T_mid']
    out['a_lo'] = info['thermo']['a_lo'].T.reshape(-1)
    out['a_hi'] = info['thermo']['a_hi'].T.reshape(-1)
    nzs = info['net_per_spec']['map']
    out['net_nonzero_spec_no_ns'] = nzs[nzs != Ns - 1]
    out['rxn_has_ns'] = info['reac_has_ns']
    if 'jac_inds' in info:
        out['jac_row_inds'] = info['jac_inds']['ccs']['row_ind']
        out['jac_col_ptr'] = info['jac_inds']['ccs']['col_ptr']
    return out
