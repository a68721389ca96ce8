"""The ``para`` configuration of finite-size DMRG (host side of the drop-in).

Same public names, keys and coupling-list conventions as the reference's
``library/Parameters.py:8-329,860-896`` and the bond/operator helpers of
``library/HamiltonianModule.py:8-44,242-370``, so scripts that build a
``para`` dict, override entries and re-call ``make_consistent_parameter_dmrg``
keep working.  Additions (not in the reference): the snake-mapped
J1-J2 square lattice (``lattice='snakeJ1J2'``) and the XXZ chain with a
uniform field used by the serving benchmark.

Conventions: ``op = [id, sx, sy, sz, su, sd, -hx*sx - hz*sz]``; an
``index2`` row ``(i, j, a, b)`` with ``coeff2`` c is ``c op[a]_i op[b]_j``;
Heisenberg bonds expand to rows ``(4,5), (5,4), (3,3)`` with coefficients
``(jxy/2, jxy/2, jz)`` (``is_pauli``: ``(2jxy, 2jxy, 4jz)``, ``coeff1 = 2``).
``para['coeff_convention'] = 'easystart'`` switches chain and square lattices
to the ``(jxy, jxy, jz)`` rows of ``EasyStart_DMRG.py:142-150``.
"""
import os
from math import comb

import numpy as np

LATTICES = ('chain', 'square', 'arbitrary', 'jigsaw', 'full', 'longRange', 'husimi',
            'snakeJ1J2')


# --------------------------------------------------------------- operators
def spin_operators(spin='half'):
    """Spin operators (``library/HamiltonianModule.py:8-44``)."""
    op = {}
    if spin == 'half':
        op['id'] = np.eye(2)
        op['sx'] = np.array([[0.0, 0.5], [0.5, 0.0]])
        op['sy'] = np.array([[0.0, 0.5j], [-0.5j, 0.0]])
        op['sz'] = np.diag([0.5, -0.5])
        op['su'] = np.array([[0.0, 1.0], [0.0, 0.0]])
        op['sd'] = np.array([[0.0, 0.0], [1.0, 0.0]])
    elif spin == 'one':
        r = 1.0 / np.sqrt(2.0)
        op['id'] = np.eye(3)
        op['sx'] = r * np.array([[0, 1, 0], [1, 0, 1], [0, 1, 0]], dtype=float)
        op['sy'] = r * np.array([[0, -1j, 0], [1j, 0, -1j], [0, 1j, 0]])
        op['sz'] = np.diag([1.0, 0.0, -1.0])
        op['su'] = np.real(op['sx'] + 1j * op['sy'])
        op['sd'] = np.real(op['sx'] - 1j * op['sy'])
    else:
        raise ValueError("spin must be 'half' or 'one', got %r" % (spin,))
    return op


def physical_dim_from_spin(spin):
    """``library/Parameters.py:890-896``."""
    return {'half': 2, 'one': 3}.get(spin, False)


# ------------------------------------------------------------- bond lists
def positions_nearest_neighbor_1d(l, bound_cond='open'):
    """``library/HamiltonianModule.py:242-253``."""
    pos = [(n, n + 1) for n in range(l - 1)]
    if bound_cond == 'periodic':
        pos.append((0, l - 1))
    return np.array(pos, dtype=int).reshape(-1, 2)


def positions_jigsaw_1d(length, bound_cond='open'):
    """``library/HamiltonianModule.py:256-291`` (odd length for OBC)."""
    if bound_cond == 'open':
        length += (length % 2 == 0)
        n_half = (length - 1) // 2
    else:
        length += (length % 2 == 1)
        n_half = length // 2
    pos = [(n, n + 1) for n in range(length - 1)]
    if bound_cond == 'periodic':
        pos.append((0, length - 1))
        pos += [(2 * n, 2 * n + 2) for n in range(n_half - 1)]
        pos.append((0, length - 2))
    else:
        pos += [(2 * n, 2 * n + 2) for n in range(n_half)]
    return np.array(pos, dtype=int)


def positions_nearest_neighbor_square(width, height, bound_cond='open'):
    """Row-major square lattice (``library/HamiltonianModule.py:294-318``):
    horizontal bonds row by row, then vertical bonds column by column."""
    pos = [(n * width + i, n * width + i + 1) for n in range(height) for i in range(width - 1)]
    pos += [(i * width + n, (i + 1) * width + n) for n in range(width) for i in range(height - 1)]
    if bound_cond == 'periodic':
        pos += [(n * width, (n + 1) * width - 1) for n in range(height)]
        pos += [(n, (height - 1) * width + n) for n in range(width)]
    return np.array(pos, dtype=int)


def positions_fully_connected(n_site):
    """All pairs ``i < j`` (``library/HamiltonianModule.py:321-330``)."""
    pos = [(a, b) for a in range(n_site - 1) for b in range(a + 1, n_site)]
    assert len(pos) == comb(n_site, 2)
    return np.array(pos, dtype=int).reshape(-1, 2)


def positions_husimi(depth):
    """Husimi cactus of triangles (``library/HamiltonianModule.py:333-348``)."""
    nb = sum(2 ** (n + 1) for n in range(depth))

    def tri(a, b, c):
        return [[a, b], [b, c], [a, c]]

    pos = tri(0, nb + 1, 2 * nb + 2)
    inner = nb - 2 ** depth
    for br in range(3):
        off = (nb + 1) * br
        for n in range(1, inner + 1):
            pos += tri(n + off, 2 * n + 1 + off, 2 * n + 2 + off)
    return np.array(pos, dtype=int)


def snake_index(x, y, width):
    """Site number of lattice point (x, y) on a row-by-row snake (boustrophedon)."""
    return y * width + (x if y % 2 == 0 else width - 1 - x)


def positions_snake_j1j2(width, height):
    """J1 (nearest) and J2 (diagonal) bonds of an open ``width x height``
    square lattice numbered along a snake path.  Returns ``(pos_j1, pos_j2)``
    with ``i < j`` in every row (not in the reference, whose square lattice is
    row-major nearest-neighbour only)."""
    j1, j2 = [], []
    for y in range(height):
        for x in range(width):
            a = snake_index(x, y, width)
            if x + 1 < width:
                j1.append(sorted((a, snake_index(x + 1, y, width))))
            if y + 1 < height:
                j1.append(sorted((a, snake_index(x, y + 1, width))))
                if x + 1 < width:
                    j2.append(sorted((a, snake_index(x + 1, y + 1, width))))
                if x - 1 >= 0:
                    j2.append(sorted((a, snake_index(x - 1, y + 1, width))))
    return np.array(sorted(j1), dtype=int), np.array(sorted(j2), dtype=int)


def interactions_position2full_index_heisenberg_two_body(index_pos):
    """Each bond -> rows ``(i,j,4,5), (i,j,5,4), (i,j,3,3)``
    (``library/HamiltonianModule.py:351-370``)."""
    index_pos = np.asarray(index_pos, dtype=int).reshape(-1, 2)
    out = np.empty((3 * index_pos.shape[0], 4), dtype=int)
    out[:, :2] = np.repeat(index_pos, 3, axis=0)
    out[:, 2:] = np.tile(np.array([[4, 5], [5, 4], [3, 3]]), (index_pos.shape[0], 1))
    return out


def sort_positions(pos):
    """Rows sorted by first site, ties by second (``algorithms/DMRG_anyH.py:318-335``)."""
    pos = np.asarray(pos, dtype=int)
    order = np.lexsort((pos[:, 1], pos[:, 0]))
    return pos[order]


def from_index2_to_positions_h2(index2):
    """Distinct bonds of ``index2`` in sorted order (``library/Parameters.py:860-868``)."""
    pos = sort_positions(np.asarray(index2)[:, :2])
    keep = np.ones(pos.shape[0], dtype=bool)
    keep[1:] = np.any(pos[1:] != pos[:-1], axis=1)
    return np.array(pos[keep], dtype=int)


def check_continuity_pos_h2(pos_h2):
    """Sites must be numbered 0..max without gaps (``library/Parameters.py:871-884``)."""
    pos_h2 = np.asarray(pos_h2)
    if pos_h2.min() != 0:
        raise ValueError('site numbering must start with 0, not %d' % pos_h2.min())
    missing = sorted(set(range(int(pos_h2.max()) + 1)) - set(np.unique(pos_h2).tolist()))
    if missing:
        raise ValueError('site numbers missing from the couplings: %s' % missing)


# ------------------------------------------------------------- para dicts
def common_parameters_dmrg():
    """``library/Parameters.py:8-26``."""
    return {'chi': 30, 'sweep_time': 100, 'if_print_detail': False, 'tau': 1e-4,
            'eigs_tol': 1e-5, 'break_tol': 1e-8, 'is_real': True, 'dt_ob': 4,
            'ob_position': 0, 'eigWay': 1, 'isParallel': False, 'isParallelEnvLMR': False,
            'is_save_op': True}


def parameter_dmrg_arbitrary():
    """``library/Parameters.py:29-64``."""
    return {'spin': 'half', 'jxy': 0, 'jz': 0, 'hx': 0, 'hz': 0,
            'index1': np.array([[0, 6], [1, 6], [2, 6]]),
            'coeff1': np.array([1, 1, 1]).reshape(-1, 1),
            'index2': np.array([[0, 1, 3, 3], [1, 2, 3, 3]]),
            'coeff2': np.array([1, 1]).reshape(-1, 1),
            'lattice': 'arbitrary', 'data_exp': 'Put_here_your_file_name_to_save_data'}


def parameter_dmrg_chain():
    """``library/Parameters.py:68-80``."""
    return {'spin': 'half', 'bound_cond': 'open', 'l': 18, 'jxy': 1, 'jz': 1, 'hx': 0, 'hz': 0,
            'lattice': 'chain'}


def parameter_dmrg_jigsaw():
    """``library/Parameters.py:84-97``."""
    return {'spin': 'one', 'bound_cond': 'open', 'l': 21, 'jxy': 1, 'jz': 1, 'jxy1': 1, 'jz1': 1,
            'hx': 0, 'hz': 0, 'lattice': 'jigsaw'}


def parameter_dmrg_square():
    """``library/Parameters.py:101-115``."""
    para = {'bound_cond': 'open', 'square_width': 4, 'square_height': 4, 'spin': 'half',
            'jxy': 1, 'jz': 1, 'hx': 0, 'hz': 0, 'lattice': 'square'}
    para['l'] = para['square_width'] * para['square_height']
    return para


def parameter_dmrg_full():
    """``library/Parameters.py:119-131``."""
    return {'spin': 'half', 'l': 6, 'jxy': 0, 'jz': 1, 'hx': 0.5, 'hz': 0, 'bound_cond': 'open',
            'is_pauli': True, 'lattice': 'full'}


def parameter_dmrg_long_range():
    """``library/Parameters.py:134-147``."""
    return {'spin': 'half', 'alpha': 1, 'l': 6, 'jxy': 0, 'jz': 1, 'hx': 0.5, 'hz': 0,
            'bound_cond': 'open', 'is_pauli': True, 'lattice': 'longRange'}


def parameter_dmrg_husimi():
    """``library/Parameters.py:150-162``."""
    return {'spin': 'half', 'depth': 2, 'jxy': 0, 'jz': 1, 'hx': 0.5, 'hz': 0, 'is_pauli': True,
            'lattice': 'husimi'}


def parameter_dmrg_snake_j1j2():
    """Snake-mapped open square lattice with J1 and J2 Heisenberg couplings."""
    return {'spin': 'half', 'square_width': 6, 'square_height': 6, 'j1': 1.0, 'j2': 0.5,
            'jxy': 1, 'jz': 1, 'hx': 0, 'hz': 0, 'lattice': 'snakeJ1J2', 'bound_cond': 'open'}


_DEFAULTS = {'chain': parameter_dmrg_chain, 'square': parameter_dmrg_square,
             'arbitrary': parameter_dmrg_arbitrary, 'jigsaw': parameter_dmrg_jigsaw,
             'full': parameter_dmrg_full, 'longRange': parameter_dmrg_long_range,
             'husimi': parameter_dmrg_husimi, 'snakeJ1J2': parameter_dmrg_snake_j1j2}


def generate_parameters_dmrg(lattice='chain'):
    """``library/Parameters.py:166-193``: lattice defaults + common parameters,
    made consistent."""
    if lattice not in _DEFAULTS:
        raise ValueError('unknown lattice %r; choose one of %s' % (lattice, LATTICES))
    para = dict(_DEFAULTS[lattice](), **common_parameters_dmrg())
    para = make_consistent_parameter_dmrg(para)
    para['project_path'] = os.getcwd()
    para['data_path'] = os.path.join(para['project_path'], 'data_dmrg')
    return para


COEFF_CONVENTIONS = ('parameters', 'easystart')


def _heisenberg_coeffs(jxy, jz, is_pauli=False, convention='parameters'):
    """coeff2 rows ``(S+S-, S-S+, SzSz)`` per bond.

    ``convention='parameters'`` (``library/Parameters.py:208-211``):
    ``(jxy/2, jxy/2, jz)``, i.e. ``H = sum jxy (SxSx + SySy) + jz SzSz``;
    ``is_pauli`` lattices use ``(2jxy, 2jxy, 4jz)`` (``Parameters.py:267-280``).
    ``convention='easystart'`` (``EasyStart_DMRG.py:146-150``): ``(jxy, jxy, jz)``,
    the convention of the reference's saved runs in ``data_dmrg``."""
    jxy = np.atleast_1d(np.asarray(jxy, dtype=float))
    jz = np.atleast_1d(np.asarray(jz, dtype=float))
    if convention not in COEFF_CONVENTIONS:
        raise ValueError('coeff_convention must be one of %s (got %r)'
                         % (COEFF_CONVENTIONS, convention))
    if convention == 'easystart':
        trip = np.stack([jxy, jxy, jz], axis=1)
    elif is_pauli:
        trip = np.stack([2 * jxy, 2 * jxy, 4 * jz], axis=1)
    else:
        trip = np.stack([jxy / 2, jxy / 2, jz], axis=1)
    return trip.reshape(-1, 1)


def _field_rows(l, value=1.0):
    index1 = np.stack([np.arange(l), np.full(l, 6)], axis=1).astype(int)
    return index1, np.full((l, 1), float(value))


def make_consistent_parameter_dmrg(para):
    """Derive ``index1/coeff1/index2/coeff2/positions_h2/op/d/nh/data_exp``
    from the lattice keys (``library/Parameters.py:197-329``)."""
    lat = para['lattice']
    if lat == 'chain':
        para['positions_h2'] = positions_nearest_neighbor_1d(para['l'], para['bound_cond'])
        para['index1'], para['coeff1'] = _field_rows(para['l'])
        para['index2'] = interactions_position2full_index_heisenberg_two_body(para['positions_h2'])
        nb = para['positions_h2'].shape[0]
        para['coeff2'] = _heisenberg_coeffs(np.full(nb, para['jxy']), np.full(nb, para['jz']),
                                            convention=para.get('coeff_convention', 'parameters'))
        para['data_exp'] = 'chainN%d_j(%g,%g)_h(%g,%g)_chi%d' % (
            para['l'], para['jxy'], para['jz'], para['hx'], para['hz'], para['chi']) + \
            para['bound_cond']
    elif lat == 'square':
        para['l'] = para['square_width'] * para['square_height']
        para['positions_h2'] = positions_nearest_neighbor_square(
            para['square_width'], para['square_height'], para['bound_cond'])
        para['index1'], para['coeff1'] = _field_rows(para['l'])
        para['index2'] = interactions_position2full_index_heisenberg_two_body(para['positions_h2'])
        nb = para['positions_h2'].shape[0]
        para['coeff2'] = _heisenberg_coeffs(np.full(nb, para['jxy']), np.full(nb, para['jz']),
                                            convention=para.get('coeff_convention', 'parameters'))
        para['data_exp'] = 'square(%d,%d)N%d_j(%g,%g)_h(%g,%g)_chi%d' % (
            para['square_width'], para['square_height'], para['l'], para['jxy'], para['jz'],
            para['hx'], para['hz'], para['chi']) + para['bound_cond']
    elif lat == 'arbitrary':
        para['index1'] = np.asarray(para['index1'], dtype=int).reshape(-1, 2)
        para['index2'] = np.asarray(para['index2'], dtype=int).reshape(-1, 4)
        para['l'] = int(max(para['index1'][:, 0].max(), para['index2'][:, 0].max(),
                            para['index2'][:, 1].max())) + 1
        para['positions_h2'] = from_index2_to_positions_h2(para['index2'])
        check_continuity_pos_h2(para['positions_h2'])
    elif lat == 'jigsaw':
        if para['bound_cond'] == 'open' and para['l'] % 2 == 0:
            para['l'] += 1
        elif para['bound_cond'] != 'open' and para['l'] % 2 == 1:
            para['l'] += 1
        para['positions_h2'] = positions_jigsaw_1d(para['l'], para['bound_cond'])
        para['index1'], para['coeff1'] = _field_rows(para['l'])
        para['index2'] = interactions_position2full_index_heisenberg_two_body(para['positions_h2'])
        nb = para['positions_h2'].shape[0]
        n1 = para['l'] - (para['bound_cond'] == 'open')
        jxy = np.array([para['jxy']] * n1 + [para['jxy1']] * (nb - n1), dtype=float)
        jz = np.array([para['jz']] * n1 + [para['jz1']] * (nb - n1), dtype=float)
        para['coeff2'] = _heisenberg_coeffs(jxy, jz)
        para['data_exp'] = 'JigsawN%d_j(%g,%g,%g,%g)_h(%g,%g)_chi%d' % (
            para['l'], para['jxy'], para['jz'], para['jxy1'], para['jz1'], para['hx'], para['hz'],
            para['chi']) + para['bound_cond']
    elif lat in ('full', 'longRange'):
        para['positions_h2'] = positions_fully_connected(para['l'])
        para['index1'], para['coeff1'] = _field_rows(para['l'], 2.0 if para['is_pauli'] else 1.0)
        para['index2'] = interactions_position2full_index_heisenberg_two_body(para['positions_h2'])
        nb = para['positions_h2'].shape[0]
        scale = np.ones(nb)
        if lat == 'longRange':
            dist = np.abs(para['positions_h2'][:, 0] - para['positions_h2'][:, 1]).astype(float)
            if para['bound_cond'] != 'open':
                dist = np.minimum(dist, para['l'] - dist)
            scale = 1.0 / dist ** para['alpha']
        para['coeff2'] = _heisenberg_coeffs(para['jxy'] * scale, para['jz'] * scale,
                                            para['is_pauli'])
        if lat == 'full':
            para['data_exp'] = 'fullConnectedN%d_j(%g,%g)_h(%g,%g)_chi%d' % (
                para['l'], para['jxy'], para['jz'], para['hx'], para['hz'], para['chi'])
        else:
            para['data_exp'] = 'longRange' + para['bound_cond'] + \
                'N%d_j(%g,%g)_h(%g,%g)_chi%d_alpha%g' % (para['l'], para['jxy'], para['jz'],
                                                        para['hx'], para['hz'], para['chi'],
                                                        para['alpha'])
    elif lat == 'husimi':
        para['positions_h2'] = positions_husimi(para['depth'])
        para['l'] = int(para['positions_h2'].max()) + 1
        para['index1'], para['coeff1'] = _field_rows(para['l'])
        para['coeff1'] = para['coeff1'].reshape(-1)
        para['index2'] = interactions_position2full_index_heisenberg_two_body(para['positions_h2'])
        nb = para['positions_h2'].shape[0]
        para['coeff2'] = _heisenberg_coeffs(np.full(nb, para['jxy']),
                                            np.full(nb, para['jz'])).reshape(-1)
        para['data_exp'] = 'HusimiDepth%d_j(%g,%g)_h(%g,%g)_chi%d' % (
            para['depth'], para['jxy'], para['jz'], para['hx'], para['hz'], para['chi'])
    elif lat == 'snakeJ1J2':
        w, h = para['square_width'], para['square_height']
        para['l'] = w * h
        p1, p2 = positions_snake_j1j2(w, h)
        para['positions_h2'] = sort_positions(np.vstack([p1, p2]))
        is_j2 = np.array([tuple(r) in set(map(tuple, p2.tolist()))
                          for r in para['positions_h2'].tolist()])
        jb = np.where(is_j2, para['j2'], para['j1'])
        para['index1'], para['coeff1'] = _field_rows(para['l'])
        para['index2'] = interactions_position2full_index_heisenberg_two_body(para['positions_h2'])
        para['coeff2'] = _heisenberg_coeffs(para['jxy'] * jb, para['jz'] * jb)
        para['data_exp'] = 'snakeJ1J2(%d,%d)N%d_j(%g,%g)_h(%g,%g)_chi%d' % (
            w, h, para['l'], para['j1'], para['j2'], para['hx'], para['hz'], para['chi'])
    else:
        raise ValueError('unknown lattice %r' % (lat,))
    para['d'] = physical_dim_from_spin(para['spin'])
    para['nh'] = para['index2'].shape[0]
    op = spin_operators(para['spin'])
    para['op'] = [op['id'], op['sx'], op['sy'], op['sz'], op['su'], op['sd'],
                  -para['hx'] * op['sx'] - para['hz'] * op['sz']]
    return para


# ---------------------------------------------------------------- TEBD
def generate_parameters_standard_tebd(lattice='chain'):
    """``library/Parameters.py:334-357``: imaginary-time TEBD of the XXZ
    chain (``tau = tau0 dtau^nt`` for ``nt < taut``, ``iterate_time`` sweeps
    of gates per tau, an energy test every ``dt_ob`` sweeps)."""
    para = {'spin': 'half', 'jxy': 1, 'jz': 1, 'hx': 0, 'hz': 0, 'l': 12, 'chi': 32,
            'bound_cond': 'open', 'tau0': 1e-1, 'dtau': 0.1, 'taut': 3, 'dt_ob': 10,
            'iterate_time': 5000, 'lattice': lattice, 'save_mode': 'final', 'if_break': True,
            'break_tol': 1e-7}
    para['data_path'] = os.path.join(os.getcwd(), 'data_tebd')
    return make_para_consistent_tebd(para)


def make_para_consistent_tebd(para):
    """``library/Parameters.py:360-368``."""
    para['positions_h2'] = positions_nearest_neighbor_1d(para['l'], para['bound_cond'])
    para['num_h2'] = para['positions_h2'].shape[0]
    para['d'] = physical_dim_from_spin(para['spin'])
    op = spin_operators(para['spin'])
    para['op'] = [op['id'], op['sx'], op['sy'], op['sz'], op['su'], op['sd'],
                  -para['hx'] * op['sx'] - para['hz'] * op['sz']]
    para['ob_time'] = int(para['iterate_time'] / para['dt_ob'])
    return para


def hamiltonian_heisenberg(spin, jx, jy, jz, hx, hz):
    """Two-site bond Hamiltonian (``library/HamiltonianModule.py:141-147``):
    ``jx SxSx + jy Re(SySy) + jz SzSz - hx (1 Sx + Sx 1) - hz (1 Sz + Sz 1)``,
    a d^2 x d^2 matrix on the merged index ``s_1 d + s_2``."""
    op = spin_operators(spin)
    h = jx * np.kron(op['sx'], op['sx']) + jy * np.kron(op['sy'], op['sy']).real + \
        jz * np.kron(op['sz'], op['sz'])
    h = h - hx * (np.kron(op['id'], op['sx']) + np.kron(op['sx'], op['id']))
    h = h - hz * (np.kron(op['id'], op['sz']) + np.kron(op['sz'], op['id']))
    return np.real(h)


# ---------------------------------------------------------------- iDMRG
def generate_parameters_infinite_dmrg():
    """``library/Parameters.py:430-452``: two-site infinite-size DMRG of the
    XXZ chain in a field, 'mpo' type (cell tensor of the shift gate
    I - tau h), chi = 16, up to 1000 iterations with an energy test every
    ``dt_ob``."""
    para = {'dmrg_type': 'mpo', 'spin': 'half', 'jxy': 0, 'jz': 1, 'hx': 0.5, 'hz': 0,
            'n_site': 2, 'chi': 16, 'sweep_time': 1000, 'tau': 1e-4, 'eigs_tol': 1e-15,
            'break_tol': 2e-10, 'is_symme_env': False, 'is_real': True, 'form': 'center_ort',
            'dt_ob': 10}
    para['data_path'] = os.path.join(os.getcwd(), 'data_idmrg')
    return make_para_consistent_idmrg(para)


def make_para_consistent_idmrg(para):
    """``library/Parameters.py:455-467``: 'white' forces symmetric
    environments; ``hamilt_index`` rows (op1, op2, coefficient) of the
    Heisenberg bond with the field split over the bond's two sites."""
    if para['dmrg_type'] not in ('mpo', 'white'):
        print("Bad para['dmrg_type']. Set to 'white'")
        para['dmrg_type'] = 'white'
    if para['dmrg_type'] == 'white':
        print("Warning: dmrg_type=='white' only suits for nearest-neighbor chains")
        print('In this mode, self.is_symme_env is set as True')
        para['is_symme_env'] = True
    para['model'] = 'heisenberg'
    para['hamilt_index'] = hamiltonian_indexes(
        para['model'], (para['jxy'], para['jz'], -para['hx'] / 2, -para['hz'] / 2))
    para['d'] = physical_dim_from_spin(para['spin'])
    return para


def hamiltonian_indexes(model, parameters):
    """``library/HamiltonianModule.py:97-138``: rows (op1, op2, coefficient)
    over the operators [id, sx, sy, sz, su, sd]; 'heisenberg' takes
    (jud, jzz, hx, hz), 'q-ising' (jzz, hx)."""
    if model == 'heisenberg':
        jud, jzz, hx, hz = parameters
        rows = [(4, 5, jud), (5, 4, jud), (3, 3, jzz), (0, 1, hx), (1, 0, hx), (0, 3, hz),
                (3, 0, hz)]
    elif model == 'q-ising':
        jzz, hx = parameters
        rows = [(3, 3, jzz), (0, 1, hx), (1, 0, hx)]
    else:
        raise ValueError('unknown model %r' % (model,))
    return np.array(rows, dtype=float)


def hamiltonian2cell_tensor(h, tau, way='shift'):
    """Cell tensor of a two-site Hamiltonian (``library/HamiltonianModule.py:177-200``):
    the gate g = I - tau h ('shift') or expm(-tau h) is split by an SVD over
    (s1 s1') | (s2 s2') into vl (s1, s1', k) and vr (k, s2, s2'), each with
    the square roots of the singular values, and
    ``T[k, (x y), (c e), q] = sum_ab vr[k, x, a] g[(a b), (c e)] vl[y, b, q]``:
    one cell of two spins with the intra-cell gate inside and the halves of
    the gates to the neighbouring cells on its MPO bonds k (left) and q
    (right).  Shape (d^2, d^2, d^2, d^2)."""
    dd = h.shape[0]
    d = int(round(dd ** 0.5))
    if way == 'shift':
        g = np.eye(dd) - tau * h
    else:
        from scipy.linalg import expm
        g = expm(-tau * h)
    g4 = g.reshape(d, d, d, d)
    u, s, vt = np.linalg.svd(g4.transpose(0, 2, 1, 3).reshape(dd, dd))
    root = np.sqrt(s)
    vl = (u * root[None, :]).reshape(d, d, dd)
    vr = (root[:, None] * vt).reshape(dd, d, d)
    t = np.einsum('kxa,abce,ybq->kxyceq', vr, g4, vl)
    return t.reshape(dd, dd, dd, dd)
