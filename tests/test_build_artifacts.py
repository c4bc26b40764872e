"""CPU: properties of the built gfx950 code objects (needs a prior build()).

Every generated register class kernel and the persistent multi-class kernel
run without scratch memory: no register spills, also where build.OCCUPANCY
asks the compiler for more waves per SIMD than it would choose by itself.
The kernel descriptors are read from the code objects this repository built
(intception_amd/_build), with the ROCm LLVM tools."""
import glob
import os
import re
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
OBJ = os.path.join(os.path.dirname(HERE), "intception_amd", "_build")
LLVM = "/opt/rocm/llvm/bin"


def _kernel_scratch(obj, tmp):
    fb, co = os.path.join(tmp, "fb.bin"), os.path.join(tmp, "co.o")
    subprocess.run([os.path.join(LLVM, "llvm-objcopy"), "--dump-section=.hip_fatbin=" + fb, obj], check=True,
                   capture_output=True)
    subprocess.run([os.path.join(LLVM, "clang-offload-bundler"), "--unbundle", "--type=o", "--input=" + fb,
                    "--targets=hipv4-amdgcn-amd-amdhsa--gfx950", "--output=" + co], check=True, capture_output=True)
    notes = subprocess.run([os.path.join(LLVM, "llvm-readelf"), "--notes", co], capture_output=True,
                           text=True, check=True).stdout
    names = re.findall(r"^\s+\.name:\s+(\S+)", notes, re.M)
    priv = [int(x) for x in re.findall(r"\.private_segment_fixed_size:\s+(\d+)", notes)]
    assert len(names) == len(priv) and names
    return dict(zip(names, priv))


@pytest.mark.skipif(not glob.glob(os.path.join(OBJ, "intc_eri_d*.o")) or not os.path.exists(
    os.path.join(LLVM, "llvm-readelf")), reason="needs build() and the ROCm LLVM tools")
def test_register_class_kernels_use_no_scratch(tmp_path):
    from intception_amd import build
    objs = sorted(glob.glob(os.path.join(OBJ, "intc_eri_[df]*.o")) + glob.glob(os.path.join(OBJ, "intc_eri_multi.o")))
    seen, bad = 0, []
    for o in objs:
        for name, priv in _kernel_scratch(o, str(tmp_path)).items():
            seen += 1
            if priv:
                bad.append((name, priv))
    print("%d kernels in %d objects" % (seen, len(objs)))
    assert seen > 100 and not bad, bad
    forced = {k + "_bs" for k in build.OCCUPANCY}
    assert forced <= {n for o in objs for n in _kernel_scratch(o, str(tmp_path))}
