"""GPU: the allocation-free per-quartet twin intc_eri_cont_sph_main_noalloc
against intc_eri_cont_sph_main, through the C-ABI: two calls in a row (the
second with a smaller input) give the same output to the FP64 parity bar
(bit-identical in practice: same classes, orientation, team size and
kernels), the output layout honours the caller's strides and contraction
strides, and over-limit inputs are rejected on the host."""
import ctypes

import numpy as np
import pytest

from conftest import assert_parity
from intception_amd import basis, lib

pytestmark = pytest.mark.gpu

_dp = ctypes.POINTER(ctypes.c_double)


def _fn(L, name):
    f = getattr(L, name)
    f.restype = ctypes.c_int
    f.argtypes = ([_dp, _dp] + [ctypes.c_int] * 6 + [_dp, _dp]) * 4 + [_dp, _dp]
    return f


def _call(f, shells, out, strides, cont_strides):
    keep, args = [], []
    for k, s in enumerate(shells):
        x = np.ascontiguousarray(s["exps"], dtype=np.float64)
        c = np.asfortranarray(np.asarray(s["coefs"], dtype=np.float64).reshape(len(x), -1))
        cen = np.ascontiguousarray(s["cen"], dtype=np.float64)
        S = np.asfortranarray(basis.SPH[s["l"]], dtype=np.float64)
        keep += [x, c, cen, S]
        args += [x.ctypes.data_as(_dp), cen.ctypes.data_as(_dp), s["l"], basis.nsph(s["l"]), int(strides[k]),
                 len(x), c.shape[1], int(cont_strides[k]), c.ctypes.data_as(_dp), S.ctypes.data_as(_dp)]
    work = np.zeros(1)
    return f(*args, work.ctypes.data_as(_dp), out.ctypes.data_as(_dp))


def _shells(rng, cls, kmax, ncont):
    out = []
    for l, nc in zip(cls, ncont):
        k = int(rng.integers(1, kmax + 1))
        e = 10 ** rng.uniform(-1, 2, size=k)
        c = np.stack([basis.normalize_contraction(e, rng.uniform(0.1, 1, size=k), l) for _ in range(nc)], 1)
        out.append({"l": l, "exps": e, "coefs": c, "cen": rng.uniform(0, 4, size=3)})
    return out


def test_noalloc_matches_per_quartet_twin(gpu):
    L = gpu
    ref_f = _fn(L, "intc_eri_cont_sph_main")
    na_f = _fn(L, "intc_eri_cont_sph_main_noalloc")
    rng = np.random.default_rng(71)
    # first call: (dp|dd) with two contractions on two indices and padded
    # strides; second: a smaller (ps|ss) block, dense
    cases = [((2, 1, 2, 2), 6, (2, 1, 2, 1), 1), ((1, 0, 0, 0), 3, (1, 1, 1, 1), 0)]
    for cls, kmax, ncont, pad in cases:
        shells = _shells(rng, cls, kmax, ncont)
        ns = [basis.nsph(l) for l in cls]
        # pad = 1: one unused element after every (c, d) block row
        strides = [ns[1] * (ns[2] * ns[3] + pad), ns[2] * ns[3] + pad, ns[3], 1]
        blk = strides[0] * ns[0]
        cstr = [blk, 2 * blk, 4 * blk, 8 * blk]
        size = 16 * blk
        want = np.full(size, 7.0)
        got = np.full(size, 7.0)
        lib.check(_call(ref_f, shells, want, strides, cstr))
        lib.check(_call(na_f, shells, got, strides, cstr))
        print("noalloc %s K<=%d ncont %s: max |d| %.3e, identical %s" % (cls, kmax, ncont,
                                                                         np.abs(got - want).max(),
                                                                         np.array_equal(got, want)), flush=True)
        assert_parity(got, want, what="noalloc vs per-quartet twin %s" % (cls,))
        # untouched positions keep the caller's values
        assert np.all(got[want == 7.0] == 7.0)


def test_noalloc_rejects_over_limit(gpu):
    L = gpu
    na_f = _fn(L, "intc_eri_cont_sph_main_noalloc")
    rng = np.random.default_rng(5)
    shells = _shells(rng, (1, 0, 0, 0), 1, (1, 1, 1, 1))
    e = 10 ** rng.uniform(-1, 1, size=9)
    shells[0] = {"l": 1, "exps": e, "coefs": basis.normalize_contraction(e, np.ones(9), 1)[:, None],
                 "cen": np.zeros(3)}
    out = np.zeros(3)
    rc = _call(na_f, shells, out, [1, 1, 1, 1], [0, 0, 0, 0])
    assert rc == -1
    assert b"at most 8 primitives" in L.intc_last_error()
