"""ctypes binding of libdgami.so (the C ABI declared in include/dgami.h).

The library is loaded from this package directory. There is no CPU fallback:
if the library is missing or no GPU is present, the first call raises.

Error mapping follows the reference's behaviour: argument/shape errors raise
ValueError (four_point.py:119-120, 266-272), singular matrices raise
numpy.linalg.LinAlgError (raised by np.linalg.inv/solve in the reference),
HIP runtime failures raise RuntimeError.
"""

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdgami.so")

PLAIN_FP64 = 0
MIXED_FP32_IR = 1
MIXED_BF16X3_IR = 2

_c = ctypes.c_void_p
_i = ctypes.c_int
_d = ctypes.c_double
_i64 = ctypes.c_int64

# name -> argument types (every function returns int except where noted)
SIGNATURES = {
    "dgami_last_error": ([], ctypes.c_char_p),
    "dgami_version": ([], _i),
    "dgami_profile_enable": ([_i], _i),
    "dgami_profile_read": ([_c, _c, _c, _c], _i),
    "dgami_fft_grid": ([_c, _c, _i64, _i, _i, _i, _i, _d, _c], _i),
    "dgami_bubble_irrq": ([_i, _i, _i, _i, _i, _i, _i, _d, _c, _c, _i, _c, _c], _i),
    "dgami_chi0_prepare": ([_i, _i, _i, _i, _i, _d, _c, _c, _c, _c, _c], _i),
    "dgami_ladder_solve_sum": ([_i, _i, _i, _i, _i, _d, _c, _c, _c, _c, _i, _c, _c, _c], _i),
    "dgami_solve_sum_dense": ([_i, _i, _i, _d, _c, _i, _c, _c, _c], _i),
    "dgami_block_invert": ([_i, _i, _c, _c, _c, _c], _i),
    "dgami_ladder_invert": ([_i, _i, _i, _i, _d, _c, _c, _c, _c, _c, _c, _c], _i),
    "dgami_full_vertex": ([_i, _i, _i, _i, _d, _c, _c, _c, _c, _c, _c, _c, _c, _c], _i),
    "dgami_full_vertex_prec": ([_i, _i, _i, _i, _d, _i, _i, _i, _c, _c, _c, _c, _c, _c, _c, _c, _c], _i),
    "dgami_pairing_vertex_pp": ([_i, _i, _i, _i, _i, _c, _c, _c, _c, _c], _i),
    "dgami_g_full": ([_i, _i, _i, _d, _d, _c, _c, _i, _i, _c, _c], _i),
    "dgami_vrg_chi_phys": ([_i, _i, _i, _i, _d, _c, _c, _c, _c, _c, _c, _c, _c, _c, _c], _i),
    "dgami_kernel_r": ([_i, _i, _i, _i, _i, _d, _c, _c, _c, _c, _i, _c, _c], _i),
    "dgami_dc_kernel": ([_i, _i, _i, _i, _i, _d, _c, _c, _c, _c, _c], _i),
    "dgami_sigma_from_kernel": ([_i, _i, _i, _i, _i, _i, _i, _d, _c, _c, _i, _c, _c, _c], _i),
    "dgami_hartree_fock": ([_i, _i, _i, _i, _c, _c, _c, _c, _c, _c, _c], _i),
    "dgami_sigma_assemble": ([_i, _i, _i, _i, _c, _c, _c, _c, _c, _i, _c, _d, _c, _c], _i),
    "dgami_fill_sums": ([_i, _i, _i, _d, _d, _c, _c, _i, _c, _c, _c, _c, _c], _i),
    "dgami_kmean": ([_i, _i64, _c, _c, _c], _i),
    "dgami_chi0_pp": ([_i, _i, _i, _i, _c, _c, _c], _i),
    "dgami_pairing_setup": ([_i, _i, _i, _i, _i, _i, _c, _c, _i, _c, _c, _c], _i),
    "dgami_pairing_matvec": ([_i, _i, _i, _i, _i, _i, _d, _c, _c, _c, _i, _c, _c, _c], _i),
    "dgami_pairing_setup_c64": ([_i, _i, _i, _i, _i, _i, _c, _c, _i, _c, _c, _c], _i),
    "dgami_pairing_matvec_c64": ([_i, _i, _i, _i, _i, _i, _d, _c, _c, _c, _i, _c, _c, _c], _i),
}

class LadderModel(ctypes.Structure):
    """dgami_ladder_model (include/dgami.h)."""
    _fields_ = [("nchan", _i), ("nq", _i), ("nb", _i), ("nw", _i), ("niv", _i), ("beta", _d),
                ("chi0_inv", _c), ("gamma", _c), ("u_r", _c), ("v_r", _c), ("chi_aux_sum", _c), ("info", _c)]


SIGNATURES["dgami_ladder_solve_sum_multi"] = ([_i, _c, _c], _i)
SIGNATURES["dgami_ladder_solve_sum_multi_prec"] = ([_i, _c, _i, _c], _i)
SIGNATURES["dgami_kernel_orbital_expand"] = ([_i, _i, _i, _i, _i, _c, _c, _c, _c, _c, _c], _i)
SIGNATURES["dgami_ladder_tune_candidates"] = ([_i, _i, _i, _c, _i], _i)
SIGNATURES["dgami_ladder_tune_set"] = ([_i, _i, _i, _i], _i)
SIGNATURES["dgami_ladder_tune_get"] = ([_i, _i, _i], _i)
SIGNATURES["dgami_ladder_tune_clear"] = ([], _i)
SIGNATURES["dgami_bubble_irrq_noalloc"] = SIGNATURES["dgami_bubble_irrq"]
SIGNATURES["dgami_invert_dense"] = ([_i, _i, _c, _c, _c, _c], _i)
SIGNATURES["dgami_pairing_vertex_pp_range"] = ([_i, _i, _i, _i, _i, _i, _c, _c, _c, _c, _c], _i)
SIGNATURES["dgami_generalized_chi"] = ([_i, _i, _i, _i, _d, _i, _c, _c, _c, _c], _i)
SIGNATURES["dgami_local_gamma"] = ([_i, _i, _i, _i, _d, _c, _c, _c, _c, _c, _c], _i)
SIGNATURES["dgami_local_sigma"] = ([_i, _i, _i, _i, _d] + [_c] * 11, _i)
SIGNATURES["dgami_g_full_dmu"] = ([_i, _i, _i, _d, _c, _c, _c, _i, _i, _c, _c], _i)
SIGNATURES["dgami_fit_smom"] = ([_i, _i, _i, _d, _c, _c, _c, _c], _i)
SIGNATURES["dgami_update_mu"] = ([_i, _i, _i, _d, _d, _d, _c, _c, _i, _c, _c, _c, _c, _c], _i)
SIGNATURES["dgami_occupations"] = ([_i, _i, _i, _d, _c, _c, _i, _c, _c, _c, _c, _c], _i)
SIGNATURES["dgami_workspace_hold"] = ([], _i)
SIGNATURES["dgami_workspace_release"] = ([], _i)

_lib = None


def load():
    """Load libdgami.so (once). Raises ImportError if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m dgamore_amd.build`")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (args, res) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def last_error():
    return load().dgami_last_error().decode()


def check(rc, what):
    """Map a C-ABI status code to the reference's exception types."""
    if rc == 0:
        return
    msg = last_error()
    if rc == -1:
        raise ValueError(f"{what}: {msg}")
    if rc > 0:
        raise np.linalg.LinAlgError(f"{what}: {rc} singular matrices")
    raise RuntimeError(f"{what}: {msg} (status {rc})")


def call(name, *args):
    check(getattr(load(), name)(*args), name)


def profile_enable(on=True):
    """Start (and reset) or stop the library's trailing-update kernel timing."""
    call("dgami_profile_enable", int(bool(on)))


def profile_read():
    """Totals since profile_enable(True): {'ms', 'bytes', 'flops', 'launches'} (waits for the events)."""
    ms, by, fl, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_double(), ctypes.c_int64()
    call("dgami_profile_read", ctypes.addressof(ms), ctypes.addressof(by), ctypes.addressof(fl), ctypes.addressof(n))
    return {"ms": ms.value, "bytes": by.value, "flops": fl.value, "launches": n.value}


def hold_workspaces(graph):
    """Keep the library's workspace buffers alive while `graph` (a captured HIP graph that names
    them) exists: outgrown buffers are retired instead of freed until the graph is collected."""
    import weakref

    n = load().dgami_workspace_hold()
    if n < 1:
        raise RuntimeError(f"dgami_workspace_hold: {last_error()}")
    fin = weakref.finalize(graph, _release_workspaces)
    fin.atexit = False
    return fin


def _release_workspaces():
    lib = load()
    if lib.dgami_workspace_release() < 0:
        raise RuntimeError(f"dgami_workspace_release: {last_error()}")
