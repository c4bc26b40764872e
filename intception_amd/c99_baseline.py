"""The generator's C99 backend as a CPU library (libintc_c99.so, built by
intception_amd.build.build_c99): OpenMP over the quartets of a batch, each
through the generated <full>_main of the spherical ERI class.  This is the
CPU baseline that runs this build's own generated C beside the GPU path."""
import ctypes
import os

import numpy as np

from . import build as _build

_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(_build.C99_LIB):
            raise RuntimeError("libintc_c99.so is not built (intception_amd.build.build_c99())")
        L = ctypes.CDLL(_build.C99_LIB)
        ip, dp = ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)
        L.intc_c99_batch.restype = ctypes.c_int
        L.intc_c99_batch.argtypes = [ip, ip, ip, dp, dp, dp, dp, ctypes.c_longlong, ip,
                                     ctypes.POINTER(ctypes.c_longlong), dp, ctypes.c_int]
        _lib = L
    return _lib


def eri_batch(batch, nthreads=None):
    """Every integral of a spherical QuartetBatch (built-in S_l) as a host
    array, dense blocks at batch.out_off."""
    from .runtime import sph_blob
    assert batch.spherical
    L = load()
    nthreads = nthreads or min(16, os.cpu_count() or 1)
    ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    arrs = [np.ascontiguousarray(batch.shell_l, dtype=np.int32), np.ascontiguousarray(batch.shell_nprim, dtype=np.int32),
            np.ascontiguousarray(batch.shell_off, dtype=np.int32)]
    cen = np.ascontiguousarray(batch.shell_cen, dtype=np.float64).reshape(-1)
    ex = np.ascontiguousarray(batch.exps, dtype=np.float64)
    co = np.ascontiguousarray(batch.coefs, dtype=np.float64)
    S = np.ascontiguousarray(sph_blob(), dtype=np.float64)
    q = np.ascontiguousarray(batch.quartets, dtype=np.int32).reshape(-1)
    oo = np.ascontiguousarray(batch.out_off, dtype=np.int64)
    out = np.empty(batch.out_size)
    rc = L.intc_c99_batch(ip(arrs[0]), ip(arrs[1]), ip(arrs[2]), dp(cen), dp(ex), dp(co), dp(S), batch.n_quartet,
                          ip(q), oo.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)), dp(out), int(nthreads))
    if rc:
        raise RuntimeError("intc_c99_batch: invalid shell l")
    return out
