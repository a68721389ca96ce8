"""tnalg_amd - MI355X-native finite-size DMRG, a drop-in for T-Nalg's
``dmrg_finite_size(para)`` path (``algorithms/DMRG_anyH.py:18-101``).

Public API (mirrors the reference's names):
    dmrg_finite_size(para=None, ...) -> (ob, A, info, para)
    parameters.generate_parameters_dmrg / make_consistent_parameter_dmrg / ...
    MpsOpenBoundaryClass  (the returned ``A``)
    tebd_standard(para) -> (A, ob, info)             (algorithms/TEBDalgo.py:11-87)
    dmrg_infinite_size(para, A, hamilt) -> (A, ob, info)   (DMRG_anyH.py:106-188)
    save_pr / load_pr     (``.pr`` result files, library/BasicFunctions.py:29-95)
    Engine / Solve        (batched device engine)
    serving.SolveServer   (continuous-batching request loop)
The compute runs in ``libtnalg_hip.so`` (HIP, gfx950); there is no CPU path.
"""
from . import parameters  # noqa: F401
from .basic_functions import load_pr, save_pr  # noqa: F401
from .parameters import generate_parameters_dmrg, make_consistent_parameter_dmrg  # noqa: F401

__all__ = ['parameters', 'generate_parameters_dmrg', 'make_consistent_parameter_dmrg',
           'dmrg_finite_size', 'Engine', 'MpsOpenBoundaryClass', 'save_pr', 'load_pr',
           'tebd_standard', 'dmrg_infinite_size', 'MpsInfinite']


def __getattr__(name):
    # the engine imports torch and loads the HIP library lazily
    if name in ('Engine', 'Solve'):
        from . import engine
        return getattr(engine, name)
    if name in ('dmrg_finite_size', 'MpsOpenBoundaryClass', 'run_dmrg'):
        from . import dmrg
        return getattr(dmrg, name)
    if name in ('tebd_standard', 'MpsStandardTEBD'):
        from . import tebd
        return getattr(tebd, name)
    if name in ('dmrg_infinite_size', 'dmrg_infinite_size_batch', 'MpsInfinite'):
        from . import idmrg
        return getattr(idmrg, name)
    raise AttributeError(name)
