"""GPU: the batched twin with general contraction and caller strides
(intc_plan_create_general): per quartet the reference's <full>_main fields,
n_cont per shell, iskip and iskip_cont per index
(contracted_general.py:44-65, 1808-1875).  Every element of the output
array, touched or not, against the oracle's per-quartet restatement with the
same strides (untouched positions keep the caller's values)."""
import numpy as np
import pytest

from conftest import assert_parity
from intception_amd import basis, runtime
from oracle import oracle

pytestmark = pytest.mark.gpu


def _case(seed, spherical, nq=40):
    rng = np.random.default_rng(seed)
    shells = []
    for _ in range(12):
        l = int(rng.integers(0, 3))
        k = int(rng.integers(1, 4))
        nc = int(rng.integers(1, 3))
        e = 10 ** rng.uniform(-1, 1.5, size=k)
        c = np.stack([basis.normalize_contraction(e, rng.uniform(0.1, 1, size=k), l) for _ in range(nc)], axis=1)
        shells.append({"l": l, "exps": e, "coefs": c, "cen": rng.uniform(0, 5, size=3), "sph": basis.SPH[l]})
    quartets = rng.integers(0, len(shells), size=(nq, 4))
    n = lambda s: (2 * s["l"] + 1) if spherical else (s["l"] + 1) * (s["l"] + 2) // 2
    out_off, isk, iskc = [], [], []
    pos = 0
    for q in quartets:
        ns = [n(shells[s]) for s in q]
        nk = [shells[s]["coefs"].shape[1] for s in q]
        # padded rows: one unused element after every (c, d) row, contractions outermost
        st = [ns[1] * (ns[2] * ns[3] + 1), ns[2] * ns[3] + 1, ns[3], 1]
        blk = ns[0] * st[0]
        cst = [blk * nk[1] * nk[2] * nk[3], blk * nk[2] * nk[3], blk * nk[3], blk]
        out_off.append(pos)
        isk.append(st)
        iskc.append(cst)
        pos += blk * int(np.prod(nk)) + 3
    return shells, quartets, np.array(out_off), np.array(isk), np.array(iskc), pos


@pytest.mark.parametrize("spherical", [False, True])
def test_general_contraction_and_strides(gpu, spherical):
    import torch
    shells, quartets, out_off, isk, iskc, size = _case(40 + spherical, spherical)
    plan = runtime.GeneralEriPlan(spherical, shells, quartets, out_off, isk, iskc, size)
    out = torch.full((size,), 7.0, dtype=torch.float64, device=plan.device)
    plan.run(out)
    got = out.cpu().numpy()
    oracle.set_boys_mode(oracle.BOYS_EXACT)
    want = np.full(size, 7.0)
    for q, sq in enumerate(quartets):
        sh = [shells[s] for s in sq]
        ref = oracle.eri_quartet(sh, spherical, basis.SPH if spherical else None, strides=list(isk[q]),
                                 ncont_strides=list(iskc[q]), out_size=size - out_off[q])
        # the addressed elements: sum_X (iX iskip_X + kX iskip_cont_X)
        nk = [shells[s]["coefs"].shape[1] for s in sq]
        ns = [(2 * shells[s]["l"] + 1) if spherical else (shells[s]["l"] + 1) * (shells[s]["l"] + 2) // 2
              for s in sq]
        idx = np.zeros(1, dtype=np.int64)
        for X in range(4):
            idx = (idx[:, None] + (np.arange(ns[X]) * isk[q][X])[None, :]).reshape(-1)
        cidx = np.zeros(1, dtype=np.int64)
        for X in range(4):
            cidx = (cidx[:, None] + (np.arange(nk[X]) * iskc[q][X])[None, :]).reshape(-1)
        el = (cidx[:, None] + idx[None, :]).reshape(-1)
        want[out_off[q] + el] = ref[el]
    print("general plan: %d quartets, %d elements, spherical=%s" % (len(quartets), size, spherical), flush=True)
    assert_parity(got, want, what="general contraction plan")
    plan.close()
