"""ctypes binding of librichmol_amd.so (the C ABI of include/richmol_amd.h).

The library is built in-tree by ``python -m richmol_amd._build``.  Loading
fails loudly when it is missing: there is no CPU fallback.
"""
import atexit
import ctypes
import os
import weakref

from ._build import LIB_PATH

RM_OK = 0
RM_NOT_CONVERGED = 1
RM_DEVICE_TIMEOUT = 2
RM_STALE_DOMAIN = 4
# rm_plan_set_tuning knobs and their defaults
TUNE_KNOBS = {"pair": 0, "coop_max": 1, "coop_grid": 2, "fork": 3, "wave_max_rows": 4, "coop_min_entries": 5}
TUNE_DEFAULTS = {"pair": 1, "coop_max": 8, "coop_grid": 0, "fork": 1, "wave_max_rows": 64,
                 "coop_min_entries": 6400}
RM_MAX_FIELD = 32
RM_MAX_LEAF = 8
RM_MAX_KRYLOV = 128

_c_i64p = ctypes.POINTER(ctypes.c_int64)
_c_i32p = ctypes.POINTER(ctypes.c_int32)
_c_dblp = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p

# name -> (restype, argtypes); every symbol declared in include/richmol_amd.h
SIGNATURES = {
    "rm_last_error": (ctypes.c_char_p, []),
    "rm_version": (ctypes.c_int, []),
    "rm_plan_create": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int64, _c_i64p, _c_i32p, _c_i64p,
                                      _c_i32p, _c_dblp, ctypes.c_int64, _c_i64p, _c_i32p, _c_dblp,
                                      _c_i32p, ctypes.c_int32, ctypes.c_int32,
                                      ctypes.POINTER(_vp)]),
    "rm_plan_destroy": (ctypes.c_int, [_vp]),
    "rm_plan_info": (ctypes.c_int, [_vp, _c_i64p, _c_i64p]),
    "rm_plan_set_field": (ctypes.c_int, [_vp, _c_dblp, _c_dblp, _vp]),
    "rm_matvec": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int64, ctypes.c_int64, _vp]),
    "rm_phase_init": (ctypes.c_int, [ctypes.c_int64, _vp, ctypes.c_double, ctypes.c_double, _vp, _vp]),
    "rm_step": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int64, ctypes.c_int64, ctypes.c_double,
                               ctypes.c_double, _vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_double,
                               _vp, _vp, _vp]),
    "rm_propagate": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                    _vp, _c_dblp, ctypes.c_double, ctypes.c_double, _vp,
                                    ctypes.c_int32, ctypes.c_int32, ctypes.c_double, _vp, _vp, _vp]),
    "rm_propagate_expect": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                           _vp, _c_dblp, ctypes.c_double, ctypes.c_double, _vp,
                                           ctypes.c_int32, ctypes.c_int32, ctypes.c_double, _vp, _vp,
                                           _vp, _vp, _vp, _vp]),
    "rm_plan_set_propagator": (ctypes.c_int, [_vp, ctypes.c_int32]),
    "rm_rotdens": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, _vp, ctypes.c_int32, _vp, ctypes.c_int32,
                                  _vp, _vp, ctypes.c_int32, _vp, ctypes.c_int32, _vp, _vp]),
    "rm_plan_timing": (ctypes.c_int, [_vp, ctypes.c_int32]),
    "rm_plan_kernel_time": (ctypes.c_int, [_vp, _c_dblp, _c_i64p]),
    "rm_plan_set_kernel": (ctypes.c_int, [_vp, ctypes.c_int32]),
    "rm_ell_config": (ctypes.c_int, [ctypes.c_int32, _c_i32p, _c_i32p, _c_i32p, _c_i32p]),
    "rm_phase_apply": (ctypes.c_int, [ctypes.c_int64, _vp, _vp, ctypes.c_int64, ctypes.c_int64, _vp, _vp]),
    "rm_expect": (ctypes.c_int, [_vp, _vp, ctypes.c_int64, ctypes.c_int64, _vp, _vp]),
    "rm_plan_subspaces": (ctypes.c_int, [_vp, _c_i64p, _c_i64p]),
    "rm_plan_set_precision": (ctypes.c_int, [_vp, ctypes.c_int32]),
    "rm_plan_set_domain_reuse": (ctypes.c_int, [_vp, ctypes.c_int32]),
    "rm_plan_set_tuning": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32]),
    "rm_plan_set_coop_abort": (ctypes.c_int, [_vp, ctypes.c_int32]),
    "rm_plan_domains": (ctypes.c_int, [_vp, _vp, ctypes.c_int64, ctypes.c_int64, _c_i32p, _vp]),
    "rm_plan_step_flags": (ctypes.c_int, [_vp, ctypes.c_int64, _vp, _vp, _vp]),
    "rm_batch_create": (ctypes.c_int, [ctypes.POINTER(_vp)]),
    "rm_batch_destroy": (ctypes.c_int, [_vp]),
    "rm_batch_set_block_rows": (ctypes.c_int, [_vp, ctypes.c_int32]),
    "rm_batch_propagate": (ctypes.c_int, [_vp, _vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                          ctypes.c_int32, ctypes.c_double, ctypes.c_int32, _vp]),
    "rm_batch_propagate_static": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                                 ctypes.c_int32, ctypes.c_double, ctypes.c_int32, _vp]),
    "rm_kplan_create": (ctypes.c_int, [ctypes.c_int32, _c_i32p, ctypes.c_int32, ctypes.c_int32, _c_i32p,
                                       _c_i32p, _c_i32p, _c_i32p, _c_i64p, _c_dblp, _c_i64p, _c_i32p,
                                       _c_i64p, _c_dblp, ctypes.POINTER(_vp)]),
    "rm_kplan_destroy": (ctypes.c_int, [_vp]),
    "rm_kplan_set_kernel": (ctypes.c_int, [_vp, ctypes.c_int32]),
    "rm_kplan_info": (ctypes.c_int, [_vp, _c_i64p, _c_dblp]),
    "rm_kplan_matvec": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int64, ctypes.c_int32, _vp, ctypes.c_int32,
                                       _vp]),
    "rm_kplan_propagate": (ctypes.c_int, [_vp, _vp, ctypes.c_int64, ctypes.c_int32, ctypes.c_int64, _c_dblp,
                                          ctypes.c_double, ctypes.c_double, _vp, ctypes.c_int32,
                                          ctypes.c_int32, ctypes.c_double, _vp, _vp, ctypes.c_int32, _vp]),
    "rm_kplan_propagate_expect": (ctypes.c_int, [_vp, _vp, ctypes.c_int64, ctypes.c_int32, ctypes.c_int64,
                                                 _c_dblp, ctypes.c_double, ctypes.c_double, _vp, ctypes.c_int32,
                                                 ctypes.c_int32, ctypes.c_double, _vp, _vp, ctypes.c_int32,
                                                 _vp, _c_dblp, _vp, _vp, _vp]),
}


class RmJob(ctypes.Structure):
    """rm_job of include/richmol_amd.h (one job of rm_batch_propagate)."""
    _fields_ = [("plan", _vp), ("psi", _vp), ("ld", ctypes.c_int64), ("col", ctypes.c_int32),
                ("lo", ctypes.c_int32), ("len", ctypes.c_int32), ("fprod_dev", _vp),
                ("sflag_dev", _vp), ("nsteps", ctypes.c_int32), ("f_re", ctypes.c_double),
                ("f_im", ctypes.c_double), ("phase", _vp), ("kused", _vp), ("status", _vp)]

RM_SUBSPACE_MAX_ROWS = 1024
PRECISIONS = {"fp64": 0, "fp32": 1}

_LIB = None


class RichmolAmdError(RuntimeError):
    pass


# Objects holding library handles (plans, batches, Kronecker plans), closed
# newest first by an exit hook while the HIP runtime is still up -- not by
# finalizers that interpreter shutdown runs in an arbitrary order.
_LIVE = []


def track(obj):
    """Register an object with a close() method for the exit-time release."""
    _LIVE.append(weakref.ref(obj))
    if len(_LIVE) > 4096:
        _LIVE[:] = [r for r in _LIVE if r() is not None]


def release_all():
    """Close every tracked handle (newest first) and drop the module caches."""
    import sys
    for name in ("richmol_amd.plan", "richmol_amd.tdse"):
        mod = sys.modules.get(name)
        if mod is not None:
            for attr in ("_CACHE", "_KRON_CACHE"):
                cache = getattr(mod, attr, None)
                if cache is not None:
                    cache.clear()
    while _LIVE:
        obj = _LIVE.pop()()
        if obj is not None:
            try:
                obj.close()
            except Exception:
                pass


def load():
    """Load the library (once) and attach the ctypes signatures."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RichmolAmdError(
            f"{LIB_PATH} is missing: build it with `python -m richmol_amd._build` "
            "(the HIP path has no CPU fallback)")
    # torch first: its HIP runtime and rocBLAS are then the ones the library binds to
    import torch  # noqa: F401
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    # registered after torch's import: runs before torch's own exit hooks
    atexit.register(release_all)
    return lib


def check(rc, what=""):
    if rc < 0:
        msg = load().rm_last_error().decode(errors="replace")
        raise RichmolAmdError(f"{what}: {msg} (status {rc})")
    return rc


def ptr(arr, ctype):
    """ctypes pointer to the data of a C-contiguous numpy array."""
    return arr.ctypes.data_as(ctypes.POINTER(ctype))
