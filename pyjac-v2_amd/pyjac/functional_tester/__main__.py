"""``python -m pyjac.functional_tester -i pasr_input.yaml -m mech.cti -o out.npy``:
run the partially stirred reactor of :mod:`.pasr` (its reaction sub-steps
on the GPU) and save the particle data as the reference's
``partially_stirred_reactor.py`` command line does (:1020-1047)."""

import sys
from argparse import ArgumentParser

import numpy as np


def main(argv=None):
    ap = ArgumentParser(description='Runs a partially stirred reactor (PaSR) simulation.')
    ap.add_argument('-i', '--input', required=True, help='PaSR parameters (YAML).')
    ap.add_argument('-m', '--mech', required=True, help='Mechanism (Chemkin or Cantera CTI).')
    ap.add_argument('-t', '--thermo', default=None, help='Thermodynamic data (Chemkin), optional.')
    ap.add_argument('-o', '--output', default='pasr_output.npy', help='Particle data (.npy).')
    ap.add_argument('--seed', type=int, default=0, help='Random seed.')
    ap.add_argument('--rtol', type=float, default=1e-6, help='Integrator relative tolerance.')
    ap.add_argument('--atol', type=float, default=1e-14, help='Integrator absolute tolerance.')
    args = ap.parse_args(argv)
    from ..core.mech_tables import MechTables
    from .pasr import gpu_reactor, parse_input_file, run_from_input
    mech = MechTables.from_file(args.mech, args.thermo)
    pars = parse_input_file(args.input)
    print('pilot equilibrium, then one line per step', flush=True)
    print('Time [ms]  Temperature [K]', flush=True)
    data = run_from_input(mech, pars, seed=args.seed,
                          react=gpu_reactor(mech, rtol=args.rtol, atol=args.atol),
                          log=lambda t, T: print('{:6.2f}  {:9.1f}'.format(t * 1000.0, T),
                                                 flush=True))
    np.save(args.output, data)
    print('{}: {} time steps x {} particles'.format(args.output, data.shape[0], data.shape[1]))
    return 0


if __name__ == '__main__':
    sys.exit(main())
