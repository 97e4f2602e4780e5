"""MI355X-native implementation of goSPL's per-timestep surface-process hot path.

The compute path is the HIP library ``libgospl_hip.so`` (built in-tree from
``gospl_amd/csrc``) called through a C ABI; PyTorch provides device memory and
streams.  ``gospl_amd.model.Model`` keeps goSPL's driver and method names
(gospl/model.py:179-466).
"""

__all__ = ["Model", "load_library"]


def load_library():
    from ._lib import load
    return load()


def __getattr__(name):
    if name == "Model":
        from .model import Model
        return Model
    raise AttributeError(name)
