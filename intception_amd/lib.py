"""ctypes binding of libintc.so (include/intc.h).

The library is the product path: there is no CPU fallback.  ``load()`` raises
when libintc.so has not been built; device entry points return an error status
(turned into ``IntcError``) when no GPU is present.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libintc.so")

_lib = None

c_int, c_ll, c_double, c_void_p = ctypes.c_int, ctypes.c_longlong, ctypes.c_double, ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_llp = ctypes.POINTER(ctypes.c_longlong)

# (name, restype, argtypes) for every symbol include/intc.h declares
_SHELL = [_dp, _dp, c_int, c_int, c_int, c_int, c_int, c_int, _dp]
SIGNATURES = [
    ("intc_last_error", ctypes.c_char_p, []),
    ("intc_version", c_int, []),
    ("intc_init", c_int, []),
    ("intc_set_boys_mode", c_int, [c_int]),
    ("intc_class_supported", c_int, [c_int] * 5),
    ("intc_eri_cont_sph_work_array_size", c_int, [c_int] * 12),
    ("intc_eri_cont_cart_work_array_size", c_int, [c_int] * 12),
    ("intc_eri_cont_sph_main", c_int, (_SHELL + [_dp]) * 4 + [_dp, _dp]),
    ("intc_eri_cont_sph_main_noalloc", c_int, (_SHELL + [_dp]) * 4 + [_dp, _dp]),
    ("intc_eri_cont_cart_main", c_int, _SHELL * 4 + [_dp, _dp]),
    ("intc_plan_create", c_int, [ctypes.POINTER(c_void_p), c_int, c_int, _ip, _ip, _ip, _dp, _dp, _dp, c_ll, _ip,
                                 _llp]),
    ("intc_plan_create_pairs", c_int, [ctypes.POINTER(c_void_p), c_int, c_int, _ip, _ip, _ip, _dp, _dp, _dp, c_int,
                                       _ip, c_ll, _ip, _llp]),
    ("intc_plan_create_general", c_int, [ctypes.POINTER(c_void_p), c_int, c_int, _ip, _ip, _ip, _ip, _ip, _dp, _dp,
                                         _dp, c_ll, _ip, _llp, _ip, _ip]),
    ("intc_plan_num_bins", c_int, [c_void_p]),
    ("intc_plan_bin_info", c_int, [c_void_p, c_int, c_int, _ip, _ip, _ip, _ip, _ip]),
    ("intc_plan_set_bin_tuning", c_int, [c_void_p, c_int, c_int, c_int, c_int]),
    ("intc_plan_build_pairs", c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    ("intc_plan_run_bin", c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p]),
    ("intc_plan_execute", c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    ("intc_plan_execute_routed", c_int, [c_void_p] * 8),
    ("intc_plan_graph_create", c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                       c_int, c_void_p]),
    ("intc_plan_graph_launch", c_int, [c_void_p, c_void_p]),
    ("intc_plan_set_route_fp32", c_int, [c_void_p, c_int]),
    ("intc_plan_set_fp32_check", c_int, [c_void_p, c_double]),
    ("intc_plan_redo_counts", c_int, [c_void_p, _ip]),
    ("intc_plan_set_persistent", c_int, [c_void_p, c_int, c_int, c_int]),
    ("intc_plan_persistent_info", c_int, [c_void_p, _ip, _ip]),
    ("intc_plan_set_persistent_bins", c_int, [c_void_p, c_int, _ip]),
    ("intc_plan_set_prim_screen", c_int, [c_void_p, c_double]),
    ("intc_plan_pair_prims", c_int, [c_void_p, c_int, _ip]),
    ("intc_plan_route_counts", c_int, [c_void_p, _ip]),
    ("intc_plan_quartet_order", c_int, [c_void_p, _ip]),
    ("intc_plan_destroy", c_int, [c_void_p]),
    ("intc_overlap_prim_cart_main", c_int, [c_double, _dp, c_int, c_int, c_int, c_double, _dp, c_int, c_int, c_int,
                                            _dp, _dp]),
    ("intc_nuclear_prim_cart_main", c_int, [c_double, _dp, c_int, c_int, c_int, c_double, _dp, c_int, c_int, c_int,
                                            _dp, _dp, _dp]),
    ("intc_overlap_cont_cart_main", c_int, _SHELL * 2 + [_dp, _dp]),
    ("intc_overlap_cont_sph_main", c_int, (_SHELL + [_dp]) * 2 + [_dp, _dp]),
    ("intc_nuclear_cont_cart_main", c_int, _SHELL * 2 + [_dp, _dp, _dp]),
    ("intc_nuclear_cont_sph_main", c_int, (_SHELL + [_dp]) * 2 + [_dp, _dp, _dp]),
    ("intc_onee_plan_create", c_int, [ctypes.POINTER(c_void_p), c_int, c_int, c_int, _ip, _ip, _ip, _dp, _dp, _dp,
                                      c_int, c_ll, _ip, _llp]),
    ("intc_onee_plan_execute", c_int, [c_void_p] * 8),
    ("intc_onee_plan_destroy", c_int, [c_void_p]),
]


class IntcError(RuntimeError):
    pass


def load():
    """Load libintc.so; raises if it has not been built (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise IntcError("libintc.so is not built (run intception_amd.build.build()); "
                        "the ERI path has no CPU fallback")
    # torch first: libintc then binds to the HIP runtime torch has loaded, the
    # one that owns the device memory and streams it hands over.  Loaded the
    # other way round, the system HIP runtime takes the shared soname and
    # torch's own device initialisation fails ("No HIP GPUs are available").
    import torch  # noqa: F401
    L = ctypes.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(rc):
    if rc != 0:
        raise IntcError("libintc status %d: %s" % (rc, load().intc_last_error().decode()))
    return rc
