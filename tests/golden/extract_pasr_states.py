"""Write tests/golden/h2_pasr_states.npz from the reference's partially
stirred reactor output (data/h2_pasr_output.npy: 102 time steps x 10
particles x [t, T, P, Y of the 10 species of data/h2o2.cti]).  Keeps the
distinct thermochemical states (T, P, Y) and the species names in the order
of h2o2.cti (`ideal_gas(..., species = "H2 H O O2 OH H2O HO2 H2O2 AR N2")`).

    python tests/golden/extract_pasr_states.py /root/reference
"""

import os
import sys

import numpy as np

SPECIES = ['H2', 'H', 'O', 'O2', 'OH', 'H2O', 'HO2', 'H2O2', 'AR', 'N2']


def main(ref_root):
    data = np.load(os.path.join(ref_root, 'data', 'h2_pasr_output.npy'), allow_pickle=False)
    flat = data.reshape(-1, data.shape[-1])
    _, idx = np.unique(flat[:, 1:], axis=0, return_index=True)
    states = flat[np.sort(idx), 1:]
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'h2_pasr_states.npz')
    np.savez_compressed(out, T=states[:, 0], P=states[:, 1], Y=states[:, 2:],
                        species=np.array(SPECIES))
    print(out, states.shape)


if __name__ == '__main__':
    main(sys.argv[1] if len(sys.argv) > 1 else '/root/reference')
