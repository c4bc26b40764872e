"""Build libintc.so (gfx950) from the hand-written runtime and the generated
per-class kernels.

    python -m intception_amd.build [--lmax 2] [--jobs 8]

Steps: Boys tables (mpmath) -> class kernels (codegen) -> hipcc -c per file
(parallel) -> link.  Objects are reused when their source is unchanged.
"""
import argparse
import hashlib
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
GEN = os.path.join(CSRC, "gen")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libintc.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
FLAGS = ["-O3", "-fPIC", "-std=c++17", "--offload-arch=" + ARCH,
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-command-line-argument"]
# generated class kernels: no machine-level loop-invariant code motion.  It
# hoists 64-bit constants (the exp polynomial of the Boys setup) and the
# epilogue's LDS addresses out of the quartet loop into registers that stay
# live across the primitive loop; without it the same kernels need fewer
# VGPRs, e.g. (pp|dp) 216 -> 162 (2 -> 3 waves per SIMD), (dp|dp) 220 -> 164,
# with no spills, for 3 % more instructions in the recurrence loop
CLASS_FLAGS = ["-mllvm", "-disable-machine-licm"]

# occupancy targets (waves per SIMD) for class kernels that reach them without
# spilling (no scratch memory; tests/test_build_artifacts.py checks every class
# kernel) and run faster there (config-3 bins on MI355X: (ps|dp) 3.32 -> 3.06
# ms, (ds|dp) 1.75 -> 1.64 ms); the compiler otherwise settles at one wave fewer
OCCUPANCY = {"intc_eri_d1021_s": 3, "intc_eri_d2021_s": 3, "intc_eri_d2121_s": 3}
# batched-setup choice per kernel where it differs from the default rule
# (codegen.BATCH_MIN_G, BATCH_LK: every group of >= 4 lanes); none at present
BATCH = {}

# classes compiled into the library: every canonical class with shells up to
# l = 2 (Cartesian and spherical), plus the three-centre set (ab|P s) with
# a, b <= d and auxiliary P up to g (the unit s-shell has exponent 0)
EXTRA_CLASSES = [(la, lb, lp, 0) for la in range(3) for lb in range(3) for lp in range(3, 5)]


def _digest(paths, extra=""):
    h = hashlib.sha1(extra.encode())
    for p in paths:
        with open(p, "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()


# the network's kernels: MFMA results in VGPRs (not AGPRs), so the router's
# activations need no accumulator-to-VGPR copies between its layers
MLP_FLAGS = ["-mllvm", "-amdgpu-mfma-vgpr-form"]


# class kernels that keep machine-level loop-invariant motion: with it they
# stay at the same waves per SIMD and without scratch (the hoisted constants
# of the Boys setup then cost no instructions per primitive batch)
LICM_CLASSES = {"intc_eri_d1011_s", "intc_eri_d1020_s", "intc_eri_d1022_s", "intc_eri_d1111_s",
                "intc_eri_d1120_s", "intc_eri_d1122_s", "intc_eri_d0011_s", "intc_eri_d2022_s"}


def _flags(src):
    """compile flags of one source: the generated class kernels add
    CLASS_FLAGS (except LICM_CLASSES), the network's source MLP_FLAGS"""
    base = os.path.splitext(os.path.basename(src))[0]
    gen = os.path.dirname(os.path.abspath(src)) == GEN and base.startswith("intc_eri_")
    mlp = base == "intc_mlp"
    return FLAGS + (CLASS_FLAGS if gen and base not in LICM_CLASSES else []) + (MLP_FLAGS if mlp else [])


def _compile(src, deps, force=False):
    os.makedirs(OBJ, exist_ok=True)
    base = os.path.splitext(os.path.basename(src))[0]
    obj = os.path.join(OBJ, base + ".o")
    stamp = obj + ".sha"
    flags = _flags(src)
    dig = _digest([src] + deps, " ".join(flags))
    if not force and os.path.exists(obj) and os.path.exists(stamp) and open(stamp).read() == dig:
        return obj, False
    cmd = [HIPCC] + flags + ["-I", CSRC, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s" % (src, r.stderr[-6000:]))
    with open(stamp, "w") as fh:
        fh.write(dig)
    return obj, True


C99_DIR = os.path.join(HERE, "_c99")
C99_LIB = os.path.join(HERE, "libintc_c99.so")
C99_FLAGS = "-O3 -march=x86-64-v3"


def c99_full_name(lmax=2):
    """<full> name the generator gives the spherical ERI class."""
    from . import eri_script
    from .generator import generator, generator_options
    gen = generator(eri_script.make_eri(), opt=generator_options(spherical_transformed=True, global_user_lmax=lmax,
                                                                 backend="c99"))
    (name,) = gen._bindings
    return name


def build_c99(lmax=2, jobs=None, verbose=True):
    """The generator's C99 backend for the spherical ERI class (shells up to
    lmax), written as the reference's file set into _c99/ and built with its
    own Makefile (libintception.a), then linked with the OpenMP batch driver
    (csrc/intc_c99_batch.c) into libintc_c99.so: the CPU baseline that runs
    this build's generated C beside the GPU path."""
    from . import c99gen
    name = c99_full_name(lmax)
    c99gen.write(C99_DIR, name, True, lmax)
    jobs = jobs or min(16, os.cpu_count() or 4)
    r = subprocess.run(["make", "-s", "-j%d" % jobs, "-C", C99_DIR, "COPT=" + C99_FLAGS], capture_output=True,
                       text=True)
    if r.returncode != 0:
        raise RuntimeError("C99 backend build failed:\n" + r.stderr[-4000:])
    cmd = ["gcc", "-std=c99", "-fPIC", "-shared", "-fopenmp"] + C99_FLAGS.split() + [
        "-I", C99_DIR, '-DINTC_C99_HEADER="%s.h"' % name, "-DINTC_C99_MAIN=%s_main" % name,
        "-o", C99_LIB + ".tmp", os.path.join(CSRC, "intc_c99_batch.c"),
        "-Wl,--whole-archive", os.path.join(C99_DIR, "libintception.a"), "-Wl,--no-whole-archive", "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("C99 batch driver link failed:\n" + r.stderr[-4000:])
    os.replace(C99_LIB + ".tmp", C99_LIB)
    if verbose:
        print("libintc_c99.so: %s (C99 backend, l <= %d, %s)" % (name, lmax, C99_FLAGS))
    return C99_LIB


def class_sources(files, metas):
    """(source, dependencies) of every generated kernel file.  The persistent
    multi-class source includes every member's class source, so those are its
    dependencies too: else a changed body would stay stale in it while the
    runtime sizes its LDS from the new class table."""
    from . import codegen
    common = [os.path.join(CSRC, "intc_common.h")]
    members = [os.path.join(GEN, m["name"] + ".hip") for m in codegen.multi_members(metas)]
    return [(f, common + members if os.path.basename(f) == "intc_eri_multi.hip" else common) for f in files]


def build(lmax=2, jobs=None, verbose=True, force=False):
    """Generate the class kernels, compile every object (all of them when
    ``force``, else only those whose sources or flags changed) and link
    libintc.so."""
    from . import boys_tables, codegen
    inc = os.path.join(CSRC, "boys_tables.inc")
    if not os.path.exists(inc):
        boys_tables.write_inc(inc)
    classes = codegen.class_set(lmax, EXTRA_CLASSES)
    wg = codegen.wg_classes(3, classes)
    files, metas = codegen.generate(GEN, classes, wg=wg, occupancy=OCCUPANCY, batch=BATCH)
    common = [os.path.join(CSRC, "intc_common.h")]
    srcs = class_sources(files, metas)
    srcs.append((os.path.join(CSRC, "intc_runtime.hip"),
                 common + [inc, os.path.join(GEN, "class_table.inc")]))
    srcs.append((os.path.join(CSRC, "intc_mlp.hip"), []))
    srcs.append((os.path.join(CSRC, "intc_eri_wg.hip"), common + [os.path.join(GEN, "wg_classes.inc")]))
    srcs.append((os.path.join(CSRC, "intc_onee.hip"), common + [os.path.join(GEN, "onee_kernels.inc")]))
    jobs = jobs or min(16, os.cpu_count() or 4)
    with ThreadPoolExecutor(jobs) as ex:
        res = list(ex.map(lambda a: _compile(*a, force=force), srcs))
    objs = [r[0] for r in res]
    rebuilt = sum(1 for r in res if r[1])
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", LIB + ".tmp"] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr[-4000:])
    os.replace(LIB + ".tmp", LIB)
    if verbose and wg:
        print("workgroup kernels: %d classes x 2 outputs x 2 precisions" % len(wg))
    if verbose:
        print("libintc.so: %d classes x 2 outputs x 2 precisions, %d objects rebuilt" % (len(classes), rebuilt))
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--lmax", type=int, default=2)
    ap.add_argument("--jobs", type=int, default=None)
    a = ap.parse_args()
    build(a.lmax, a.jobs)
