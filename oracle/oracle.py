"""ctypes wrapper of the C oracle (eri_oracle.c).  TEST INFRASTRUCTURE ONLY.

``build()`` compiles liberi_oracle.so with the Makefile beside this file.
``eri_quartet`` calls the per-quartet restatement of the generated
``<full>_main`` (argument layout of
src/intception/generator/callables/main_source_algorithms/contracted_general.py:31-73);
``eri_batch`` runs it over a batch with OpenMP, the way the CPU baseline drives it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liberi_oracle.so")
_GOLDEN = os.path.join(os.path.dirname(_HERE), "tests", "golden", "boys_reference_tables.npz")
_lib = None
_tables = None  # keep the table arrays alive while the library points at them

BOYS_REFERENCE = 0
BOYS_EXACT = 1

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_long)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib, _tables
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    L = ctypes.CDLL(_LIB_PATH)
    L.oracle_boys_ref.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double]
    L.oracle_boys_exact.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_double]
    L.oracle_set_boys_tables.argtypes = [_dp, _dp]
    L.oracle_set_boys_mode.argtypes = [ctypes.c_int]
    L.oracle_get_boys_mode.restype = ctypes.c_int
    L.oracle_eri_work_array_size.argtypes = [ctypes.c_int] * 13
    L.oracle_eri_work_array_size.restype = ctypes.c_int
    shell = [_dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
             ctypes.c_int, _dp]
    L.oracle_eri_cont_sph_main.argtypes = (shell + [_dp]) * 4 + [_dp, _dp]
    L.oracle_eri_cont_cart_main.argtypes = shell * 4 + [_dp, _dp]
    L.oracle_eri_batch.argtypes = [ctypes.c_int, ctypes.c_int, _ip, _ip, _ip, _dp, _dp, _dp,
                                   ctypes.POINTER(_dp), ctypes.c_long, _ip, _lp, _dp, ctypes.c_int]
    L.oracle_eri_batch.restype = ctypes.c_int
    L.oracle_onee_prim.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp, ctypes.c_double,
                                   _dp, _dp, _dp]
    L.oracle_onee_batch.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _ip, _ip, _ip, _dp, _dp, _dp,
                                    ctypes.POINTER(_dp), _dp, ctypes.c_long, _ip, _lp, _dp, ctypes.c_int]
    L.oracle_onee_batch.restype = ctypes.c_int
    t = np.load(_GOLDEN)
    small = np.ascontiguousarray(t["smallx"], dtype=np.float64)
    medium = np.ascontiguousarray(t["mediumx"], dtype=np.float64)
    _tables = (small, medium)
    L.oracle_set_boys_tables(small.ctypes.data_as(_dp), medium.ctypes.data_as(_dp))
    _lib = L
    return L


def set_boys_mode(mode):
    lib().oracle_set_boys_mode(int(mode))


def boys_ref(nmax, x, nmin=0):
    f = np.zeros(nmax + 1)
    lib().oracle_boys_ref(f.ctypes.data_as(_dp), 1, nmin, nmax, float(x))
    return f


def boys_exact(nmax, x):
    f = np.zeros(nmax + 1)
    lib().oracle_boys_exact(f.ctypes.data_as(_dp), 0, nmax, float(x))
    return f


def work_array_size(spherical, ls, nprim, ncont=(1, 1, 1, 1)):
    return lib().oracle_eri_work_array_size(int(spherical), *[int(v) for v in ls],
                                            *[int(v) for v in nprim], *[int(v) for v in ncont])


def _ptr(a):
    return a.ctypes.data_as(_dp)


def eri_quartet(shells, spherical, sph=None, strides=None, ncont_strides=None, out_size=None):
    """One contracted quartet through the per-quartet restatement.

    shells: 4 dicts with keys l, exps (n_prim), coefs (n_prim x n_cont, column
    major, i.e. Fortran order), cen (3).  sph: dict l -> S (ncc x nsph,
    column-major) when spherical.  Output is row-major [a][b][c][d] for
    n_cont = 1 unless strides are given.
    """
    L = lib()
    ns = [(2 * s["l"] + 1) if spherical else (s["l"] + 1) * (s["l"] + 2) // 2 for s in shells]
    if strides is None:
        strides = [ns[1] * ns[2] * ns[3], ns[2] * ns[3], ns[3], 1]
    if ncont_strides is None:
        ncont_strides = [0, 0, 0, 0]
    if out_size is None:
        out_size = int(np.prod(ns))
    out = np.zeros(out_size)
    keep = []
    args = []
    for k, s in enumerate(shells):
        x = np.ascontiguousarray(s["exps"], dtype=np.float64)
        c = np.asfortranarray(np.asarray(s["coefs"], dtype=np.float64).reshape(len(x), -1))
        cen = np.ascontiguousarray(s["cen"], dtype=np.float64)
        keep += [x, c, cen]
        args += [_ptr(x), _ptr(cen), s["l"], ns[k], int(strides[k]), len(x), c.shape[1],
                 int(ncont_strides[k]), c.ctypes.data_as(_dp)]
        if spherical:
            S = np.asfortranarray(sph[s["l"]], dtype=np.float64)
            keep.append(S)
            args.append(S.ctypes.data_as(_dp))
    work = np.zeros(1)
    if spherical:
        L.oracle_eri_cont_sph_main(*args, _ptr(work), _ptr(out))
    else:
        L.oracle_eri_cont_cart_main(*args, _ptr(work), _ptr(out))
    return out


def eri_batch(batch, spherical, sph_by_l=None, nthreads=None):
    """Whole batch (an intception_amd.basis.QuartetBatch-like object) with OpenMP."""
    L = lib()
    if nthreads is None:
        nthreads = os.cpu_count() or 1
    out = np.zeros(int(batch.out_size))
    sph_list = (_dp * 8)()
    keep = []
    if spherical:
        for l in range(8):
            if l in sph_by_l:
                S = np.asfortranarray(sph_by_l[l], dtype=np.float64)
                keep.append(S)
                sph_list[l] = S.ctypes.data_as(_dp)
    ip = lambda a: np.ascontiguousarray(a, dtype=np.int32)
    sl, snp, soff = ip(batch.shell_l), ip(batch.shell_nprim), ip(batch.shell_off)
    q = ip(batch.quartets.reshape(-1))
    oo = np.ascontiguousarray(batch.out_off, dtype=np.int64)
    cen = np.ascontiguousarray(batch.shell_cen, dtype=np.float64)
    ex = np.ascontiguousarray(batch.exps, dtype=np.float64)
    co = np.ascontiguousarray(batch.coefs, dtype=np.float64)
    rc = L.oracle_eri_batch(int(spherical), len(sl), sl.ctypes.data_as(_ip), snp.ctypes.data_as(_ip),
                            soff.ctypes.data_as(_ip), _ptr(cen), _ptr(ex), _ptr(co), sph_list,
                            len(oo), q.ctypes.data_as(_ip), oo.ctypes.data_as(_lp), _ptr(out),
                            int(nthreads))
    assert rc == 0
    return out


def onee_prim(kind, la, lb, xa, A, xb, B, C=None):
    """Primitive Cartesian one-electron block (kind 0 overlap, 1 nuclear
    attraction at C), row-major [a][b] (onee_oracle.c)."""
    out = np.zeros(((la + 1) * (la + 2) // 2) * ((lb + 1) * (lb + 2) // 2))
    arr = [np.ascontiguousarray(v, dtype=np.float64) for v in (A, B, C if C is not None else np.zeros(3))]
    lib().oracle_onee_prim(int(kind), int(la), int(lb), float(xa), _ptr(arr[0]), float(xb), _ptr(arr[1]),
                           _ptr(arr[2]), _ptr(out))
    return out


def onee_batch(kind, spherical, shells, items, out_off, out_size, centres=None, sph_by_l=None, nthreads=None):
    """Contracted one-electron blocks over an SoA shell table.  shells: object
    with shell_l, shell_nprim, shell_off, shell_cen, exps, coefs; items (n, 2)
    shell pairs or (n, 3) (a, b, centre); centres (n_centre, 3)."""
    L = lib()
    nthreads = nthreads or (os.cpu_count() or 1)
    out = np.zeros(int(out_size))
    sph_list = (_dp * 8)()
    keep = []
    if spherical:
        for l in range(8):
            if l in sph_by_l:
                S = np.asfortranarray(sph_by_l[l], dtype=np.float64)
                keep.append(S)
                sph_list[l] = S.ctypes.data_as(_dp)
    ip = lambda a: np.ascontiguousarray(a, dtype=np.int32)
    sl, snp, soff = ip(shells.shell_l), ip(shells.shell_nprim), ip(shells.shell_off)
    it = ip(np.asarray(items).reshape(-1))
    oo = np.ascontiguousarray(out_off, dtype=np.int64)
    cen = np.ascontiguousarray(shells.shell_cen, dtype=np.float64)
    ex = np.ascontiguousarray(shells.exps, dtype=np.float64)
    co = np.ascontiguousarray(shells.coefs, dtype=np.float64)
    cc = np.ascontiguousarray(centres if centres is not None else np.zeros((1, 3)), dtype=np.float64)
    rc = L.oracle_onee_batch(int(kind), int(spherical), len(sl), sl.ctypes.data_as(_ip), snp.ctypes.data_as(_ip),
                             soff.ctypes.data_as(_ip), _ptr(cen), _ptr(ex), _ptr(co), sph_list, _ptr(cc),
                             len(oo), it.ctypes.data_as(_ip), oo.ctypes.data_as(_lp), _ptr(out), int(nthreads))
    assert rc == 0
    return out
