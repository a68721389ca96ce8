"""Write tests/golden/c2_block_fp64.json: the f64 CPU oracle's result on the
batched benchmark chain (XXZ N=64 open, Jxy=1, Jz=1, h=0, chi=128, ncv 20
(ARPACK's), tau 1e-4, eigs_tol 1e-10, 8 sweeps) from the seeded initial MPS of
each seed, by the block-environment restatement ``oracle/block_dmrg`` (same
effective Hamiltonian, gauge moves and ARPACK local solve as the reference,
``algorithms/DMRG_anyH.py:18-101``; about an hour of CPU per seed).

Run from the repository root:  python tests/golden/make_c2_golden.py [seed ...]
(existing seeds in the file are kept; default seed 0).
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import numpy as np  # noqa: E402

from oracle import block_dmrg, dmrg_ref, model  # noqa: E402

OUT = os.path.join(HERE, 'c2_block_fp64.json')


def c2_para():
    c = model.chain_couplings(64, 1.0, 1.0)
    return dmrg_ref.make_para(c, 128, 8, tau=1e-4, eigs_tol=1e-10, break_tol=0.0, dt_ob=1)


def main(seeds):
    out = {}
    if os.path.exists(OUT):
        with open(OUT) as f:
            out = json.load(f)
    out['_provenance'] = ('oracle/block_dmrg (numpy/scipy f64, ARPACK eigsh ncv 20) on the XXZ '
                          'chain N=64 open, Jxy=Jz=1, h=0, chi=128, tau=1e-4, eigs_tol=1e-10, '
                          '8 sweeps, initial MPS from RandomState(seed) in the reference draw '
                          'order; written by tests/golden/make_c2_golden.py')
    for seed in seeds:
        t0 = time.time()
        ob, A, info, _ = block_dmrg.dmrg_block(c2_para(), seed=seed)
        e = float(np.ravel(ob['e_per_site'])[0]) * 64
        out['seed%d' % seed] = {
            'E_total': e, 'energies_per_observation': info['energies'],
            'eb_full_0_6': np.ravel(ob['eb_full'])[:6].tolist(),
            'mz_max_abs': float(np.abs(ob['mz']).max()),
            'bond_dims': [int(x) for x in A.virtual_dim],
            'cpu_seconds': time.time() - t0}
        print('seed %d: E = %.15f (%.0f s)' % (seed, e, time.time() - t0), flush=True)
        with open(OUT, 'w') as f:
            json.dump(out, f, indent=1)


if __name__ == '__main__':
    main([int(a) for a in sys.argv[1:]] or [0])
