"""Write tests/golden/ed_values.json: exact ground-state energies of small
coupling-list Hamiltonians (sparse exact diagonalisation, oracle/ed.py).

Run from the repository root:  python tests/golden/make_ed_golden.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import ed, model  # noqa: E402


def cases():
    yield 'heisenberg_N12_obc', model.chain_couplings(12, 1.0, 1.0)
    yield 'fixture_N12_easystart', model.chain_couplings(12, 1.0, 1.0, convention='easystart')
    yield 'xxz_N10_jz0.5_hz0.3', model.chain_couplings(10, 1.0, 0.5, hz=0.3)
    yield 'xx_N10', model.chain_couplings(10, 1.0, 0.0)
    bonds = [(i, i + 1) for i in range(9)] + [(i, i + 2) for i in range(8)]
    yield 'j1j2_N10_mg', model.couplings_from_bonds(10, bonds, [1.0] * 9 + [0.5] * 8,
                                                    [1.0] * 9 + [0.5] * 8)


def main():
    out = {}
    for name, c in cases():
        e0, _ = ed.ground_state(c['l'], c['op'], c['index1'], c['coeff1'], c['index2'], c['coeff2'])
        out[name] = e0
    with open(os.path.join(HERE, 'ed_values.json'), 'w') as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print(json.dumps(out, indent=1, sort_keys=True))


if __name__ == '__main__':
    main()
