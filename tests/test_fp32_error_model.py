"""CPU: the FP32 error model behind the FP32 route (tests/fp32_emulation.py,
a numpy restatement of the register kernels' recurrence arithmetic).

Quartets of config-3 molecule shells (the 200-atom molecule of seed 202603;
each quartet's second and fourth shells within 5 bohr of the first and
third), against the oracle's accurate Boys scheme:

* the FP64 restatement meets the oracle to rounding, which pins it;
* FP32 recurrences lose up to ~1e-4 of a block's maximum;
* computing every recurrence value exactly from its FP32 inputs and
  rounding it once (the limit of compensated sums inside a recurrence step)
  cuts the largest error by less than 3x: the error comes from rounding the
  stored FP32 values, amplified by the later steps, the contraction, the HRR
  and the transforms, not from the summation inside a step;
* the FP32 kernels' a-posteriori estimate C 2^-24 max|(e 0|f 0)|
  (codegen.fp32_err_c) is at least twice the measured error of every
  quartet."""
import numpy as np
import pytest

from intception_amd import basis, codegen, molecule
from oracle import oracle

import fp32_emulation as emu

U = 2.0 ** -24


def _molecule_quartets(cls, nq, seed, rmax=5.0):
    mol = molecule.Molecule(200, 202603)
    rng = np.random.default_rng(seed)
    by_l = {l: np.nonzero(mol.shell_l == l)[0] for l in range(3)}
    rows = []
    while len(rows) < nq:
        a, c = rng.choice(by_l[cls[0]]), rng.choice(by_l[cls[2]])
        bs = by_l[cls[1]][np.linalg.norm(mol.shell_cen[by_l[cls[1]]] - mol.shell_cen[a], axis=1) < rmax]
        ds = by_l[cls[3]][np.linalg.norm(mol.shell_cen[by_l[cls[3]]] - mol.shell_cen[c], axis=1) < rmax]
        if len(bs) and len(ds):
            rows.append((a, rng.choice(bs), c, rng.choice(ds)))
    return basis.QuartetBatch(True, mol.shell_l, mol.shell_nprim, mol.shell_off, mol.shell_cen, mol.exps,
                              mol.coefs, np.array(rows, dtype=np.int32))


@pytest.mark.parametrize("cls", [(1, 1, 2, 1), (1, 1, 2, 2), (2, 0, 2, 0)])
def test_fp32_error_model(cls):
    nq = 400
    b = _molecule_quartets(cls, nq, seed=11)
    oracle.set_boys_mode(oracle.BOYS_EXACT)
    ref = oracle.eri_batch(b, True, basis.SPH).reshape(nq, -1)
    bm = np.abs(ref).max(1)
    sel = bm > 1e-13

    def err(got):
        return np.abs(got - ref).max(1)

    d, _ = emu.evaluate(b, cls, "d")
    e64 = err(d)
    assert np.all(e64[sel] <= 1e-11 * bm[sel] + 1e-15)
    f, inter = emu.evaluate(b, cls, "f")
    e32 = err(f)
    f_step, _ = emu.evaluate(b, cls, "f", mode="step_rounded")
    e_step = err(f_step)
    f_in, _ = emu.evaluate(b, cls, "f", mode="inputs_only")
    e_in = err(f_in)
    rel = e32[sel] / bm[sel]
    print("%s: FP32 err / block max median %.1e max %.1e; largest error: plain %.2e, steps rounded once %.2e, "
          "exact on FP32 inputs %.2e" % (cls, np.median(rel), rel.max(), e32.max(), e_step.max(), e_in.max()))
    assert rel.max() > 1e-7
    assert e_step.max() > e32.max() / 3.0
    assert e_in.max() <= e32.max()
    c = codegen.fp32_err_c(cls)
    ratio = e32 / (U * np.maximum(inter, 1e-300))
    print("  error / (2^-24 max|(e0|f0)|): max %.1f, estimate constant C = %.0f" % (ratio.max(), c))
    assert np.all(e32 <= 0.5 * c * U * inter + 1e-300)
