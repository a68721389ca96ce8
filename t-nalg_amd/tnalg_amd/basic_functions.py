"""``.pr`` files: the reference's result files (``library/BasicFunctions.py:29-95``).

``save_pr(path, file, data, names)`` pickles ``{names[i]: data[i]}`` into
``path/file`` (creating ``path``); ``load_pr(path_file, names=None)`` returns
the dict, one entry (``names`` a string) or a tuple of entries (``names`` a
list or tuple), or ``False`` when the file does not exist.  Callers cache
``dmrg_finite_size`` results this way (``ScriptRun/DeepMPS/VaryChi.py:43-49``,
``EasyStart_DMRG.py:156-165``): ``save_pr(p, f, (ob, A, info, para), ('ob', 'A',
'info', 'para'))``.  ``A`` (``dmrg.MpsOpenBoundaryClass``) pickles as host
data and rebuilds its device state on first use after loading.

Unlike the reference, ``load_pr`` unpickles through an allow-list: numpy's
array and scalar reconstructors, a few builtins and this package's MPS
class.  Files written by ``save_pr`` load; a file that names anything else
raises ``pickle.UnpicklingError`` instead of running it.
"""
import builtins
import collections
import os
import pickle

import numpy as np

_NUMPY = {
    ('numpy._core.multiarray', '_reconstruct'), ('numpy.core.multiarray', '_reconstruct'),
    ('numpy._core.multiarray', 'scalar'), ('numpy.core.multiarray', 'scalar'),
    ('numpy', 'ndarray'), ('numpy', 'dtype'),
    ('numpy.matrixlib.defmatrix', 'matrix'), ('numpy', 'matrix'),
}
_BUILTINS = {'complex', 'set', 'frozenset', 'slice', 'range', 'bytearray'}


class _Loader(pickle.Unpickler):
    def find_class(self, module, name):
        if (module, name) in _NUMPY:
            if name == 'matrix':
                return np.matrix
            mod = 'numpy._core.multiarray' if module.startswith('numpy.core') else module
            return pickle.Unpickler.find_class(self, mod, name)
        if module == 'builtins' and name in _BUILTINS:
            return getattr(builtins, name)
        if (module, name) == ('collections', 'OrderedDict'):
            return collections.OrderedDict
        if (module, name) == ('tnalg_amd.dmrg', 'MpsOpenBoundaryClass'):
            from .dmrg import MpsOpenBoundaryClass
            return MpsOpenBoundaryClass
        raise pickle.UnpicklingError('load_pr: %s.%s is not an allowed type' % (module, name))


def mkdir(path):
    """``BasicFunctions.mkdir``: create ``path`` if it does not exist."""
    if path and not os.path.isdir(path):
        os.makedirs(path, exist_ok=True)


def save_pr(path, file, data, names):
    """``{names[i]: data[i]}`` pickled into ``path/file`` (``BasicFunctions.py:29-55``)."""
    if len(data) != len(names):
        raise ValueError('save_pr: %d data items for %d names' % (len(data), len(names)))
    mkdir(path)
    with open(os.path.join(path, file), 'wb') as s:
        pickle.dump({names[i]: data[i] for i in range(len(names))}, s)


def load_pr(path_file, names=None):
    """The dict saved by ``save_pr``, one entry (``names`` a str) or a tuple of
    entries (a list or tuple of names); ``False`` if there is no such file
    (``BasicFunctions.py:58-92``)."""
    if not os.path.isfile(path_file):
        return False
    with open(path_file, 'rb') as s:
        tmp = _Loader(s).load()
    if names is None:
        return tmp
    if isinstance(names, str):
        return tmp[names]
    if isinstance(names, (list, tuple)):
        return tuple(tmp[n] for n in names)
    raise TypeError('load_pr: names must be None, a str, a list or a tuple')
