"""Write tests/golden/config4_arbitration.npz: 40-digit McMurchie-Davidson
values (oracle/mp_eri.py) of the three-centre (ab|P s) elements on which the
FP64 GPU path and the oracle differed by more than 1e-14 + 1e-12 |ref|.

Workload: the 100-atom three-centre molecule of seed 202604
(molecule.ThreeCenterMolecule), every quartet with Q_ab Q_P >= 1e-12; its
class-stratified sample of 1998 quartets (27 classes) held 24 such elements
in six quartets, listed below by shell indices.  At 40 digits the GPU values
agree to <= 6e-16 and the oracle (the reference's a->c->d scheme) is off by
up to 2.4e-14.  The fixture holds those quartets' shells, the 24 elements,
the block maximum and 8 random further elements of each block, and their
40-digit values.

    python tests/golden/make_config4_arbitration.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from intception_amd import basis, molecule  # noqa: E402
from oracle import mp_eri, oracle  # noqa: E402

DISPUTED = {
    (319, 324, 1794, 1795): [15, 42, 61, 62, 67, 69, 71, 78, 89, 103, 107, 116, 193, 197, 224],
    (214, 454, 2706, 2707): [211, 215, 220, 224],
    (364, 384, 2466, 2467): [139, 224],
    (303, 489, 1818, 1819): [130],
    (358, 409, 2898, 2899): [15],
    (358, 484, 2274, 2275): [71],
}


def main():
    mol = molecule.ThreeCenterMolecule(100, seed=202604)
    quartets = np.array(list(DISPUTED), dtype=np.int32)
    b = basis.QuartetBatch(True, mol.shell_l, mol.shell_nprim, mol.shell_off, mol.shell_cen, mol.exps, mol.coefs,
                           quartets)
    oracle.set_boys_mode(oracle.BOYS_EXACT)
    ref = oracle.eri_batch(b, True, basis.SPH, 8)
    rng = np.random.Generator(np.random.PCG64(44))
    rec = {k: [] for k in ("l", "nprim", "exps", "coefs", "cen", "elem", "elem_q", "disputed", "val")}
    sz = b.sizes()
    for k, els in enumerate(DISPUTED.values()):
        blk = ref[b.out_off[k]:b.out_off[k] + sz[k]]
        extra = [int(np.argmax(np.abs(blk)))] + [int(v) for v in rng.choice(len(blk), size=8, replace=False)]
        allel = list(els) + [e for e in dict.fromkeys(extra) if e not in els]
        shells = b.quartet_shells(k)
        vals = mp_eri.contracted_eri_elements(shells, allel, True, basis.SPH)
        for sh in shells:
            rec["l"].append(sh["l"])
            rec["nprim"].append(len(sh["exps"]))
            rec["exps"].append(np.pad(sh["exps"], (0, 6 - len(sh["exps"]))))
            rec["coefs"].append(np.pad(sh["coefs"], (0, 6 - len(sh["coefs"]))))
            rec["cen"].append(sh["cen"])
        rec["elem"] += allel
        rec["elem_q"] += [k] * len(allel)
        rec["disputed"] += [e in els for e in allel]
        rec["val"] += [float(v) for v in vals]
        print("quartet %s: %d elements" % (quartets[k].tolist(), len(allel)), flush=True)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "config4_arbitration.npz")
    np.savez(path, l=np.array(rec["l"], dtype=np.int32), nprim=np.array(rec["nprim"], dtype=np.int32),
             exps=np.array(rec["exps"]), coefs=np.array(rec["coefs"]), cen=np.array(rec["cen"]),
             elem=np.array(rec["elem"], dtype=np.int64), elem_q=np.array(rec["elem_q"], dtype=np.int32),
             disputed=np.array(rec["disputed"], dtype=bool), val=np.array(rec["val"]))
    print("wrote", path)


if __name__ == "__main__":
    main()
