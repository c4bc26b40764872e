"""CPU: the workgroup kernels (intception_amd/csrc/intc_eri_wg.hip, classes
with an f shell and (g s|g s)) compiled with g++ -fsanitize=address and run on
the host (tests/emu/wg_driver.cpp), against the oracle.

Each case is one template instantiation intc_eri_wg_k<VT, BM, LA, LB, LC, LD>
run as 256 OS threads per workgroup: __syncthreads is a barrier of the 256,
the FP64 MFMA builtin an operand exchange between the 64 threads of a wave in
the instruction's lane layout.  The LDS region has exactly the size the host
layout function gives the launch (filled with NaN bytes first) and every
device array exactly its size, so AddressSanitizer reports any access outside
them and a read of an LDS entry no thread wrote shows as NaN.  Two
workgroups stride over the quartets; routed cases read the queue (q_idx,
device-side length).  FP64 results meet the oracle to rounding, FP32 ones to
the recurrence precision.  By default a set of classes covering every ket
and bra level count runs; INTC_EMU_ALL=1 runs every class of the kernel in
both precisions."""
import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from intception_amd import basis, codegen, runtime
from intception_amd import build as bld
from oracle import oracle
from test_kernel_emulation import _inputs

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "intception_amd", "csrc")
GXX = shutil.which("g++")
LDS_DECL = "extern __shared__ __attribute__((aligned(16))) unsigned char wg_lds[];"


def _wg_all():
    return codegen.wg_classes(3, codegen.class_set(2, bld.EXTRA_CLASSES))


# (class, prec, spherical, routed, Boys mode)
DEFAULT = [
    ((0, 0, 3, 1), "d", 1, 1, 0),
    ((1, 0, 3, 3), "f", 1, 0, 0),
    ((2, 1, 3, 2), "d", 0, 1, 1),
    ((2, 2, 3, 3), "d", 1, 0, 0),
    ((3, 0, 3, 0), "d", 1, 1, 1),
    ((3, 2, 3, 2), "f", 0, 1, 0),
    ((3, 3, 3, 3), "d", 1, 1, 0),
    ((3, 3, 3, 3), "f", 1, 0, 1),
    ((4, 0, 4, 0), "d", 1, 0, 0),
]


def _cases():
    if os.environ.get("INTC_EMU_ALL"):
        return [(c, p, s, 1, 0) for c in _wg_all() for p in ("d", "f") for s in (1, 0)]
    return DEFAULT


def _source(tmp):
    """The kernel source up to its launch code, LDS array as a pointer."""
    src = open(os.path.join(CSRC, "intc_eri_wg.hip")).read()
    assert LDS_DECL in src
    src = src.replace(LDS_DECL, "unsigned char *wg_lds = emu_wg_lds;")
    src = src.split("int intc_wg_prepare()")[0]
    inc = os.path.join(tmp, "wg_kernel.inc")
    with open(inc, "w") as fh:
        fh.write(src)
    return inc


def _build(tmp, inc, cls, prec):
    exe = os.path.join(tmp, "wg_%s_%s" % ("".join(map(str, cls)), prec))
    la, lb, lc, ld = cls
    cmd = [GXX, "-std=c++20", "-O1", "-g", "-fsanitize=address", "-fno-omit-frame-pointer",
           "-I", os.path.join(HERE, "emu", "include"), "-I", CSRC, '-DWG_FILE="%s"' % inc,
           "-DEMU_VT=" + ("double" if prec == "d" else "float"),
           "-DEMU_LA=%d" % la, "-DEMU_LB=%d" % lb, "-DEMU_LC=%d" % lc, "-DEMU_LD=%d" % ld,
           "-Wno-attributes", "-o", exe, os.path.join(HERE, "emu", "wg_driver.cpp"), "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    return exe


def _run(tmp, exe, case, nq=12, seed=7):
    cls, prec, sph, routed, bm = case
    b, parts = _inputs(cls, nq, seed, 1, 0, bool(sph))
    tag = "%s_%s_%d%d%d" % ("".join(map(str, cls)), prec, sph, routed, bm)
    inp, outp = os.path.join(tmp, "in_" + tag), os.path.join(tmp, "out_" + tag)
    with open(inp, "wb") as fh:
        for a in parts:
            a.tofile(fh)
        runtime.sph_blob().astype(np.float64).tofile(fh)
    r = subprocess.run([exe, inp, outp, str(sph), "2", str(routed), str(bm)], capture_output=True, text=True,
                       env=dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1"))
    assert r.returncode == 0, (case, r.stdout[-500:], r.stderr[-3000:])
    got = np.fromfile(outp)
    ref = oracle.eri_batch(b, bool(sph), basis.SPH)
    return got, ref


@pytest.mark.skipif(GXX is None, reason="needs g++")
def test_workgroup_kernels_on_the_host(tmp_path):
    tmp = str(tmp_path)
    inc = _source(tmp)
    cases = _cases()
    keys = sorted({(c[0], c[1]) for c in cases})
    nproc = min(8, os.cpu_count() or 1)
    with ThreadPoolExecutor(nproc) as ex:
        exes = dict(zip(keys, ex.map(lambda k: _build(tmp, inc, *k), keys)))
    oracle.set_boys_mode(oracle.BOYS_EXACT)
    # each run is 256 threads meeting at barriers: a few at a time
    with ThreadPoolExecutor(max(1, nproc // 4)) as ex:
        res = list(ex.map(lambda c: _run(tmp, exes[(c[0], c[1])], c), cases))
    bad = []
    for case, (got, ref) in zip(cases, res):
        assert not np.any(got == 12345.0), ("output element not written", case)
        rel = np.abs(got - ref).max() / np.abs(ref).max()
        print(case, "max err / block max %.2e" % rel)
        if not rel < (1e-11 if case[1] == "d" else 1e-3):
            bad.append((case, rel))
    assert not bad, bad
