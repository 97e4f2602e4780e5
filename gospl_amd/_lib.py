"""ctypes binding of the C-ABI library ``libgospl_hip.so``.

Every exported symbol is declared here with its argument types.  There is no
fallback: if the library is missing the import fails with an error that says
how to build it.
"""

from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgospl_hip.so")

_lib = None

P = ctypes.c_void_p
I64 = ctypes.c_int64
I32 = ctypes.c_int32
F64 = ctypes.c_double

class FlowParams(ctypes.Structure):
    """gospl_flow_params."""
    _fields_ = [("nrcv", I32), ("flowexp", F64), ("sealevel", F64), ("floor_level", F64), ("add_eps", I32),
                ("dt", F64), ("maxarea", F64), ("max_passes", I32), ("patience", I32), ("rel_improve", F64)]


class FlowBuffers(ctypes.Structure):
    """gospl_flow_buffers (device pointers)."""
    _fields_ = [(n, P) for n in ("h", "rain", "lfill", "pitids", "flatdirs", "ptdir", "lspill", "rcv0", "dist0",
                                 "wgt0", "lsink0", "rcv1", "dist1", "wgt1", "lsink1", "b", "fal", "fillfal",
                                 "water", "hl", "q", "nfa", "rhs", "fa")] + \
        [("cap", I64)] + [(n, P) for n in ("spill", "pitparams", "flvl", "fvol", "inv")]


class FlowResult(ctypes.Structure):
    """gospl_flow_result."""
    _fields_ = [("npits", I64), ("cascade_passes", I32), ("flow_solves", I32), ("undrained", I64),
                ("last_set", I32)]


# name -> (restype, argtypes).  Mirrors include/gospl_hip.h.
SIGNATURES = {
    "gospl_last_error": (ctypes.c_char_p, []),
    "gospl_fv_geometry": (I32, [P, I64, P, P, I64, P, I64, P, I32, P, P, P, P, P]),
    "gospl_mesh_create": (I32, [I64, I32, I32, P, P, P, P, P, P, P, ctypes.POINTER(P)]),
    "gospl_mesh_destroy": (I32, [P]),
    "gospl_mesh_touch": (I32, [P]),
    "gospl_mfdreceivers": (I32, [P, I32, F64, P, F64, P, P, P, P]),
    "gospl_flow_direction": (I32, [P, I32, F64, P, F64, P, I32, P, P, P, P, P, P, P]),
    "gospl_fill_depressions": (I32, [P, P, F64, F64, P, P]),
    "gospl_fill_elevation": (I32, [P, P, I32, F64, F64, P, P]),
    "gospl_pit_info": (I32, [P, P, P, F64, P, P, P, P, I64, P, P, P, P, ctypes.POINTER(I64), P]),
    "gospl_pit_tables": (I32, [P, I64, I64, P, P, P, P, P, P]),
    "gospl_flowacc": (I32, [P, I32, P, P, P, P, ctypes.POINTER(I32), ctypes.POINTER(I64), P]),
    "gospl_spl": (I32, [P, I32, P, P, P, P, P, F64, F64, F64, F64, P, P, P, P, P, P,
                        ctypes.POINTER(I32), P]),
    "gospl_flowacc_refine": (I32, [P, I32, P, P, P, P, F64, I32, ctypes.POINTER(I32), ctypes.POINTER(F64),
                                   ctypes.POINTER(I64), P]),
    "gospl_spl_refine": (I32, [P, I32, P, P, P, P, P, F64, F64, F64, F64, P, P, P, P, P, P, F64, I32,
                               ctypes.POINTER(I32), ctypes.POINTER(F64), P]),
    "gospl_hillslope": (I32, [P, F64, F64, F64, F64, I32, P, P, P, P]),
    "gospl_hillslope_coeff": (I32, [P, P, P, P]),
    "gospl_hillslope_refine": (I32, [P, F64, F64, F64, F64, I32, F64, I32, P, P, P, ctypes.POINTER(I32),
                                     ctypes.POINTER(F64), P]),
    "gospl_spl_nl": (I32, [P, I32, P, P, P, P, P, F64, F64, F64, F64, F64, P, P, P, P, P, P, F64, I32,
                           ctypes.POINTER(I32), ctypes.POINTER(F64), P]),
    "gospl_spl_dep": (I32, [P, I32, P, P, P, P, P, F64, F64, F64, F64, F64, P, P, P, P, P, P, F64, I32,
                            ctypes.POINTER(I32), ctypes.POINTER(F64), P, P]),
    "gospl_spl_nl_dep": (I32, [P, I32, P, P, P, P, P, F64, F64, F64, F64, F64, F64, P, P, P, P, P, P, F64, I32,
                               ctypes.POINTER(I32), ctypes.POINTER(I32), ctypes.POINTER(F64), P]),
    "gospl_coast_distance": (I32, [P, P, P, F64, P, P, ctypes.POINTER(I64), P]),
    "gospl_eps_fill": (I32, [P, P, F64, P, ctypes.POINTER(I32), P]),
    "gospl_diffuse_fixed": (I32, [P, P, P, P, I32, ctypes.POINTER(I32), P]),
    "gospl_marine_receivers": (I32, [P, P, P, P, P, P]),
    "gospl_marine_prep": (I32, [P, P, P, P, P, F64, F64, F64, P, P, P]),
    "gospl_marine_route": (I32, [P, P, P, P, P, P, P, ctypes.POINTER(I32), P]),
    "gospl_marine_dep": (I32, [P, P, P, P, P, F64, F64, F64, P, F64, I32, ctypes.POINTER(I32),
                               ctypes.POINTER(F64), P]),
    "gospl_marine_thickness": (I32, [P, P, F64, P, P]),
    "gospl_marine_apply": (I32, [P, P, F64, P, P, P, P]),
    "gospl_flowacc_adjoint": (I32, [P, I32, P, P, P, P, P]),
    "gospl_spl_adjoint": (I32, [P, I32, P, P, P, P, P, F64, F64, F64, F64, P, P, P, P, P, P, P, P, P]),
    "gospl_hillslope_frozen": (I32, [P, F64, F64, F64, P, I32, P, P, P]),
    "gospl_hillslope_adjoint": (I32, [P, F64, F64, F64, P, I32, P, P, P]),
    "gospl_rain_source": (I32, [P, P, P, F64, P, P]),
    "gospl_pit_inflow": (I32, [P, P, P, P, I64, P, P]),
    "gospl_pit_water": (I32, [P, P, P, P, P, P, P]),
    "gospl_cascade_seed": (I32, [P, P, I64, I64, P, P, P, F64, P, P]),
    "gospl_cascade_add": (I32, [P, P, P, P, F64, P, P, P]),
    "gospl_flow_finalize": (I32, [P, P, P, P, I32, P, P, P]),
    "gospl_mesh_set_tiebreak": (I32, [P, P]),
    "gospl_dag_set_shape": (I32, [P, I32, I32, I32]),
    "gospl_dag_get_shape": (I32, [P, P]),
    "gospl_fill_elevation_seg": (I32, [P, P, I32, F64, F64, P, I32, P, P]),
    "gospl_flow_finalize_seg": (I32, [P, P, P, P, P, I32, P, P, P, P]),
    "gospl_spl_batch": (I32, [P, I32, P, P, P, P, P, F64, P, F64, F64, P, P, P, P, P, P, ctypes.POINTER(I32), P]),
    "gospl_hillslope_refine_seg": (I32, [P, F64, F64, F64, F64, P, I32, P, P, F64, I32, P, P, P, P, P]),
    "gospl_flow_finalize_noalloc": (I32, [I64, P, P, P, I32, P, P, P]),
    "gospl_flow_accumulation": (I32, [P, ctypes.POINTER(FlowParams), ctypes.POINTER(FlowBuffers),
                                      ctypes.POINTER(FlowResult), P]),
    "gospl_flow_tables": (I32, [P, ctypes.POINTER(P), ctypes.POINTER(I64)]),
    "gospl_np_sum": (F64, [P, I64]),
    "gospl_pit_info_sed": (I32, [P, P, P, F64, F64, P, P, P, P, I64, P, P, P, P, ctypes.POINTER(I64), P]),
    "gospl_sed_source": (I32, [P, P, P, P]),
    "gospl_sed_stray": (I32, [P, P, P, I32, P, P, P]),
    "gospl_sed_add": (I32, [P, P, P, P, F64, P, P, P]),
    "gospl_sed_closed": (I32, [P, P, P, P, P]),
    "gospl_sed_bottomup": (I32, [P, P, P, P, P, P, P]),
    "gospl_sed_tilt": (I32, [P, P, P, I64, P, P, P, P, P]),
    "gospl_sed_inlets": (I32, [P, I32, P, P, I64, P, P, P, P]),
    "gospl_sed_lake_depth": (I32, [P, P, I64, P, P, P, P, P, P, P]),
    "gospl_sed_lake_init": (I32, [P, P, P, P, P, P, P, P, P, P, P]),
    "gospl_sed_pit_volume": (I32, [P, P, P, P, I64, P, P]),
    "gospl_sed_scale": (I32, [P, P, P, P, P, P]),
    "gospl_sed_apply": (I32, [P, P, P, P, P]),
    "gospl_nl_diffusion": (I32, [P, P, P, F64, F64, F64, F64, I32, I32, P, P, ctypes.POINTER(I64), P]),
    "gospl_sweep_timing": (I32, [P, I32, ctypes.POINTER(F64), ctypes.POINTER(I64),
                                 ctypes.POINTER(F64), ctypes.POINTER(I64)]),
}


def check(rc: int, what: str = "") -> None:
    """Raise RuntimeError with the library's message when rc != 0."""
    if rc != 0:
        msg = load().gospl_last_error()
        raise RuntimeError("%s failed (%d): %s" % (what, rc, msg.decode() if msg else ""))


def _declare(lib):
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


def load(path: str | None = None):
    """Load (once) and return the library handle."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    # torch first: the library takes torch's streams and device pointers, so it
    # must bind to the HIP runtime torch loads (its bundled libamdhip64.so.7).
    # Loading this library first would bring in the system runtime of the
    # same soname, and torch would then resolve to it, beside its own HSA
    # runtime, and find no device.
    import torch  # noqa: F401
    if not os.path.exists(p):
        raise RuntimeError(
            "libgospl_hip.so not found at %s: build it with "
            "`python -c 'import __graft_entry__ as g; g.build()'` "
            "(or `python -m gospl_amd.build`)" % p)
    lib = ctypes.CDLL(p)
    _declare(lib)
    if path is None:
        _lib = lib
    return lib
