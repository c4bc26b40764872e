"""CPU: generated class-kernel bodies compiled with g++ -fsanitize=address and
run on the host (tests/emu), against the oracle.

Each body (the code the gfx950 class kernels run) is built against a host
stand-in for the device builtins it uses: one thread per lane of a team, a
barrier of the team's lanes at every cross-lane read and wave barrier.  The
LDS region is exactly the size the launcher gives a wave (NAME_reg(T)
doubles) and every device array is exactly its size, so AddressSanitizer
reports any read or write outside them.  Cases: several teams per wave, team
size 2 (the cross-group reduction), routed queues (q_idx, device-side queue
length), the persistent kernel's form of the body (INTC_BODIES_ONLY), FP64
and FP32 recurrences, and the three-centre (ab|P s) form with a unit s-shell
of exponent 0.  FP64 results meet the oracle to rounding; the FP32 ones to
the recurrence precision.  This is how the kernels' index arithmetic is
checked without a GPU: GPU AddressSanitizer is not available."""
import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from intception_amd import basis, codegen
from intception_amd import build as bld
from oracle import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "intception_amd", "csrc")
GXX = shutil.which("g++")

# (class, prec, T, routed, bodies_only, three_centre_unit_s)
CASES = [
    ((0, 0, 0, 0), "d", 1, 1, 0, 0),
    ((1, 1, 2, 1), "d", 1, 1, 0, 0),
    ((1, 1, 2, 1), "f", 2, 0, 0, 0),
    ((1, 0, 1, 0), "f", 1, 1, 1, 0),
    ((0, 0, 1, 1), "d", 2, 1, 0, 0),
    ((2, 2, 2, 2), "d", 1, 0, 0, 0),
    ((2, 1, 4, 0), "d", 1, 1, 0, 1),
    ((2, 1, 2, 2), "f", 1, 1, 1, 0),
]


def _build(tmp, cls, prec, bodies):
    name = codegen.kernel_name(cls, True, prec)
    src, meta = codegen.emit_class_source(cls, True, prec, None, bld.BATCH.get(name))
    inc = os.path.join(tmp, name + ".inc")
    with open(inc, "w") as fh:
        fh.write(src.replace('#include "../intc_common.h"', '#include "intc_common.h"'))
    exe = os.path.join(tmp, "emu_%s_%d" % (name, bodies))
    cmd = [GXX, "-std=c++20", "-O1", "-g", "-fsanitize=address", "-fno-omit-frame-pointer",
           "-I", os.path.join(HERE, "emu", "include"), "-I", CSRC, '-DCLASS_FILE="%s"' % inc, "-DKNAME=" + name,
           "-Wno-attributes", "-o", exe, os.path.join(HERE, "emu", "lane_driver.cpp"), "-lpthread"]
    if bodies:
        cmd.insert(1, "-DINTC_BODIES_ONLY")
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    return exe, meta


def _inputs(cls, nq, seed, T, unit, sph=True):
    """Random quartets in canonical form, one bra and one ket pair each, and
    the pair builder's arrays (intc_build_pairs, computed on the host)."""
    la, lb, lc, ld = cls
    rng = np.random.default_rng(seed)
    b = basis.make_random_quartets(seed, np.array([cls] * nq, dtype=np.int32), [2] if T > 1 else [1, 2, 3], sph,
                                   box=4.0)
    q = b.quartets
    if unit:
        for i in range(nq):
            sc, sd = q[i, 2], q[i, 3]
            o, n = b.shell_off[sd], b.shell_nprim[sd]
            b.exps[o:o + n] = 0.0
            b.coefs[o:o + n] = 0.0
            b.coefs[o] = 1.0
            b.shell_cen[sd] = b.shell_cen[sc]
            o, n = b.shell_off[sc], b.shell_nprim[sc]
            b.exps[o:o + n] = 10 ** rng.uniform(np.log10(0.05), np.log10(200.0), size=n)
            b.coefs[o:o + n] = basis.normalize_contraction(b.exps[o:o + n], np.ones(n), lc)
    pr = np.empty((2 * nq, 2), dtype=np.int64)
    pr[0::2], pr[1::2] = q[:, :2], q[:, 2:]
    sa, sb = pr[:, 0], pr[:, 1]
    npr = b.shell_nprim[sa] * b.shell_nprim[sb]
    poff = np.concatenate([[0], np.cumsum(npr)[:-1]])
    A, B = b.shell_cen[sa], b.shell_cen[sb]
    cols = [[] for _ in range(6)]
    for p in range(2 * nq):
        kb = b.shell_nprim[sb[p]]
        for loc in range(npr[p]):
            ia, ib = divmod(loc, kb)
            al, be = b.exps[b.shell_off[sa[p]] + ia], b.exps[b.shell_off[sb[p]] + ib]
            ca, cb = b.coefs[b.shell_off[sa[p]] + ia], b.coefs[b.shell_off[sb[p]] + ib]
            z = al + be
            P = (al * A[p] + be * B[p]) / z
            kap = 5.914967172795612878 * np.exp(-al * be / z * np.sum((A[p] - B[p]) ** 2)) * ca * cb / z
            for c, v in zip(cols, (z, 0.5 / z, P[0], P[1], P[2], kap)):
                c.append(v)
    ns = [basis.nsph(l) if sph else basis.ncc(l) for l in cls]
    blk = int(np.prod(ns))
    st = [ns[1] * ns[2] * ns[3], ns[2] * ns[3], ns[3], 1]
    qstr = sum(int(s) << (16 * k) for k, s in enumerate(st))
    parts = [np.array([nq, 2 * nq, int(npr.sum()), nq * blk], dtype=np.int32),
             np.arange(0, 2 * nq, 2, dtype=np.int32), np.arange(1, 2 * nq, 2, dtype=np.int32),
             np.arange(nq, dtype=np.int64) * blk, np.full(nq, qstr, dtype=np.uint64),
             poff.astype(np.int32), npr.astype(np.int32), b.shell_nprim[sb].astype(np.int32),
             A.astype(np.float64).ravel(), (A - B).astype(np.float64).ravel()]
    parts += [np.array(c, dtype=np.float64) for c in cols]
    parts.append(rng.permutation(nq).astype(np.int32))
    return b, parts


def _run(tmp, exes, case, nq=24, seed=5):
    cls, prec, T, routed, bodies, unit = case
    exe = exes[(cls, prec, bodies)]
    b, parts = _inputs(cls, nq, seed, T, unit)
    tag = "%s_%s_%d%d%d%d" % ("".join(map(str, cls)), prec, T, routed, bodies, unit)
    inp, outp = os.path.join(tmp, "in_" + tag), os.path.join(tmp, "out_" + tag)
    with open(inp, "wb") as fh:
        for a in parts:
            a.tofile(fh)
    r = subprocess.run([exe, inp, outp, str(T), "64", str(routed)], capture_output=True, text=True,
                       env=dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1"))
    assert r.returncode == 0, (case, r.stderr[-3000:])
    got = np.fromfile(outp)
    ref = oracle.eri_batch(b, True, basis.SPH)
    return got, ref


@pytest.mark.skipif(GXX is None, reason="needs g++")
def test_fp32_error_check_queue_on_the_host(tmp_path):
    """FP32 bodies with the error check on: a tolerance below every estimate
    queues every quartet exactly once, one above every estimate queues none,
    and the output is the same either way (the FP64 kernel's redo is the
    runtime's, not the body's)."""
    tmp = str(tmp_path)
    oracle.set_boys_mode(oracle.BOYS_EXACT)
    for cls, T in (((1, 1, 2, 1), 1), ((1, 0, 1, 0), 2), ((2, 1, 2, 1), 1)):
        exe, _ = _build(tmp, cls, "f", 0)
        b, parts = _inputs(cls, 20, 9, T, 0)
        nout = int(parts[0][3])
        inp = os.path.join(tmp, "in_chk")
        with open(inp, "wb") as fh:
            for arr in parts:
                arr.tofile(fh)
        outs = []
        for tol in ("1e-300", "1e300"):
            outp = os.path.join(tmp, "out_chk")
            r = subprocess.run([exe, inp, outp, str(T), "64", "1", tol], capture_output=True, text=True,
                               env=dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1"))
            assert r.returncode == 0, r.stderr[-3000:]
            raw = np.fromfile(outp, dtype=np.uint8)
            out = raw[:8 * nout].view(np.float64)
            n = int(raw[8 * nout:8 * nout + 4].view(np.int32)[0])
            q = raw[8 * nout + 4:].view(np.int32)
            outs.append(out)
            if tol == "1e-300":
                assert n == 20 and sorted(q.tolist()) == list(range(20)), (cls, n, q)
            else:
                assert n == 0 and len(q) == 0, (cls, n)
        assert np.array_equal(outs[0], outs[1])


@pytest.mark.skipif(GXX is None, reason="needs g++")
def test_class_bodies_on_the_host(tmp_path):
    tmp = str(tmp_path)
    keys = sorted({(c[0], c[1], c[4]) for c in CASES})
    with ThreadPoolExecutor(min(8, os.cpu_count() or 1)) as ex:
        built = list(ex.map(lambda k: _build(tmp, *k), keys))
    exes = {k: e for k, (e, _) in zip(keys, built)}
    oracle.set_boys_mode(oracle.BOYS_EXACT)
    with ThreadPoolExecutor(min(8, os.cpu_count() or 1)) as ex:
        res = list(ex.map(lambda c: _run(tmp, exes, c), CASES))
    for case, (got, ref) in zip(CASES, res):
        assert not np.any(got == 12345.0), ("output element not written", case)
        rel = np.abs(got - ref).max() / np.abs(ref).max()
        print(case, "max err / block max %.2e" % rel)
        assert rel < (1e-11 if case[1] == "d" else 1e-3), case


@pytest.mark.skipif(GXX is None, reason="needs g++")
def test_queue_entries_outside_the_bin_store_nothing(tmp_path):
    """Routed queue entries outside the bin (negative or >= nq) are passed over:
    their teams store nothing and queue nothing for the FP64 redo, and every
    other quartet of the queue is evaluated as usual (FP64 and FP32 bodies, one
    and two groups per team, the class kernels' pipelined loop and the
    persistent kernel's loop)."""
    tmp = str(tmp_path)
    oracle.set_boys_mode(oracle.BOYS_EXACT)
    nq = 16
    for cls, prec, T, bodies in (((1, 1, 2, 1), "d", 1, 0), ((1, 0, 1, 0), "f", 2, 0), ((1, 1, 2, 1), "d", 1, 1)):
        exe, _ = _build(tmp, cls, prec, bodies)
        b, parts = _inputs(cls, nq, 3, T, 0)
        blk = int(parts[0][3]) // nq
        queue = parts[-1].copy()
        dropped = queue[[1, 6, 7, 12]].copy()
        queue[[1, 6, 7, 12]] = [-1, nq, -5, nq + 3]
        parts = parts[:-1] + [queue]
        inp, outp = os.path.join(tmp, "in_bad"), os.path.join(tmp, "out_bad")
        with open(inp, "wb") as fh:
            for arr in parts:
                arr.tofile(fh)
        args = [exe, inp, outp, str(T), "64", "1"] + (["1e-300"] if prec == "f" else [])
        r = subprocess.run(args, capture_output=True, text=True,
                           env=dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1"))
        assert r.returncode == 0, r.stderr[-3000:]
        raw = np.fromfile(outp, dtype=np.uint8)
        got = raw[:8 * nq * blk].view(np.float64).reshape(nq, blk)
        ref = oracle.eri_batch(b, True, basis.SPH).reshape(nq, blk)
        keep = np.setdiff1d(np.arange(nq), dropped)
        assert np.all(got[dropped] == 12345.0), (cls, prec, bodies)
        rel = np.abs(got[keep] - ref[keep]).max() / np.abs(ref[keep]).max()
        assert rel < (1e-11 if prec == "d" else 1e-3), (cls, prec, bodies, rel)
        if prec == "f":
            n = int(raw[8 * nq * blk:8 * nq * blk + 4].view(np.int32)[0])
            q = raw[8 * nq * blk + 4:].view(np.int32)
            assert n == nq - len(dropped) and sorted(q.tolist()) == sorted(keep.tolist()), (n, q)
