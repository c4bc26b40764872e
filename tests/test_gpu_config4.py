"""GPU: the FP64 path and the learned routing on the full three-centre
workload: (ab|P s) quartets of the 100-atom molecule of seed 202604 with 12
auxiliary shells per atom (l <= 4, K = 1) and a unit s-shell in the fourth
slot, every quartet with Q_ab Q_P >= 1e-12, in plan order.

* FP64 parity: ~2000 quartets drawn evenly from the 27 canonical classes of
  the full plan's output, against the oracle (accurate Boys scheme).  Bar:
  3e-14 + 1e-12 |ref| + 5e-11 of the block maximum.  The oracle's own a->c->d
  scheme is the less accurate side here: on the 24 sample elements outside
  the plain FP64 bar (1e-14 + 1e-12 |ref|) the GPU agrees with the 40-digit
  McMurchie-Davidson values to <= 6e-16 and the oracle is off by up to
  2.4e-14 absolute, about 1.2e-11 of the block maximum
  (tests/golden/config4_arbitration.npz).
* Those elements against their 40-digit values: the GPU meets the plain FP64
  bar there.
* Learned routing at full size: a router trained on three-centre molecules
  of other seeds; the routed output stays within 1e-10 of plain FP64, alone
  and with the primitive-pair screen (eps = 1e-11) on top.
"""
import time

import numpy as np
import pytest

from arbitration import arbitration_batch
from conftest import assert_parity
from intception_amd import runtime, screening
from molecule_checks import block_bar_check, gather_blocks, max_abs_diff, oracle_blocks, stratified_sample

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mol4(gpu):
    t0 = time.perf_counter()
    mol, pf = screening.molecule_setup(100, 202604, three_center=True)
    pp = screening.plan_sorted(mol, mol.screened_quartets(1e-12))
    batch = mol.pair_batch(pp)
    plan = runtime.EriPlan(batch)
    out = plan.run()
    print("100-atom three-centre molecule: %d quartets, %d integrals, FP64 output in %.1f s"
          % (batch.n_quartet, batch.out_size, time.perf_counter() - t0), flush=True)
    yield mol, pf, pp, batch, plan, out
    plan.close()


def test_full_three_center_fp64_vs_oracle(mol4):
    _, _, _, batch, _, out = mol4
    idx, n_cls = stratified_sample(batch, 2000, seed=4)
    got = gather_blocks(out, batch, idx)
    ref, sub = oracle_blocks(batch, idx)
    n_bad, worst, emax = block_bar_check(got, ref, sub, atol=3e-14, block_rel=5e-11)
    n_strict, worst_strict, _ = block_bar_check(got, ref, sub)
    print("FP64 vs oracle: %d quartets over %d classes, %d elements, max abs err %.3e, worst err/bar %.3f; "
          "plain FP64 bar: %d elements outside (worst err/bar %.2f)"
          % (len(idx), n_cls, len(ref), emax, worst, n_strict, worst_strict), flush=True)
    assert n_cls == 27
    assert n_bad == 0


def test_full_three_center_learned_routing(mol4):
    import torch
    mol, pf, pp, batch, plan, out64 = mol4
    t0 = time.perf_counter()
    # the benchmark's router (training and calibration sets as there)
    router, stats = screening.build_router(n_train=2_000_000, train_seeds=(202621, 202622),
                                           val_seed=(202629, 202628, 202627), n_val=1_000_000, n_atoms=100,
                                           epochs=25, under_weight=100.0, tau_skip=2.5e-11, three_center=True)
    print("router built in %.1f s: val skip %.3f, val fp32 %.3f" % (time.perf_counter() - t0,
                                                                    stats["val_skip_fraction"],
                                                                    stats["val_fp32_fraction"]), flush=True)
    dev = plan.device
    route = torch.empty(batch.n_quartet, dtype=torch.int8, device=dev)
    router.route(torch.from_numpy(pf).to(dev), torch.from_numpy(np.ascontiguousarray(pp, dtype=np.int32)).to(dev),
                 route)
    codes = torch.bincount(route.to(torch.int64), minlength=3).cpu().numpy() / batch.n_quartet
    plan.set_route_fp32(True)
    # FP32 error check at the router's FP32 budget: FP32-routed quartets whose
    # estimate exceeds it are recomputed by the FP64 kernels
    plan.set_fp32_check(router.fp32_budget)
    got = plan.run_routed(route)
    err = max_abs_diff(got, out64)
    print("routed fp64/fp32/skip = %s, %d FP32 quartets redone in FP64, max abs err vs FP64 %.3e"
          % (np.round(codes, 4).tolist(), plan.redo_counts().sum(), err), flush=True)
    assert err <= 1e-10
    plan.set_prim_screen(plan.prim_screen_cut(1e-11))
    try:
        plan.run_routed(route, got)
        err_ps = max_abs_diff(got, out64)
    finally:
        plan.set_prim_screen(0.0)
    print("routed + primitive screen: max abs err vs FP64 %.3e" % err_ps, flush=True)
    assert err_ps <= 1e-10
    del got


def test_three_center_arbitrated_elements(gpu):
    b, g = arbitration_batch("config4")
    got = runtime.compute(b)
    pos = b.out_off[g["elem_q"]] + g["elem"]
    print("config4 arbitration: %d elements (%d disputed), max |gpu - mp| %.3e" % (
        len(pos), int(g["disputed"].sum()), np.abs(got[pos] - g["val"]).max()), flush=True)
    assert_parity(got[pos], g["val"], what="three-centre elements vs 40-digit MD")
