"""GPU: the FP64 path and the learned routing on the full screened-molecule
workload: the 200-atom molecule of seed 202603 (random sequential packing,
shells s,s,p,p,d with K = 4,1,3,1,1 per atom), every unique shell quartet
with Q_P Q_P' >= 1e-12 (68.5 M quartets, 4.17e9 integrals), in plan order.

* FP64 parity: ~2000 quartets drawn evenly from the 21 canonical classes of
  the full plan's output, against the oracle (accurate Boys scheme) at
  |d| <= 1e-14 + 1e-12 |ref|.
* Learned routing at full size: a router trained on molecules of other seeds
  routes every quartet (skip / FP32 recurrences / FP64); the routed output
  stays within 1e-10 of the plain FP64 output, alone and with the
  primitive-pair screen (eps = 1e-11) on top.
"""
import time

import numpy as np
import pytest

from intception_amd import runtime, screening
from molecule_checks import block_bar_check, gather_blocks, max_abs_diff, oracle_blocks, stratified_sample

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mol3(gpu):
    t0 = time.perf_counter()
    mol, pf = screening.molecule_setup(200, 202603)
    pp = screening.plan_sorted(mol, mol.screened_quartets(1e-12))
    batch = mol.pair_batch(pp)
    plan = runtime.EriPlan(batch)
    out = plan.run()
    print("200-atom molecule: %d quartets, %d integrals, FP64 output in %.1f s"
          % (batch.n_quartet, batch.out_size, time.perf_counter() - t0), flush=True)
    yield mol, pf, pp, batch, plan, out
    plan.close()


def test_full_molecule_fp64_vs_oracle(mol3):
    _, _, _, batch, _, out = mol3
    assert batch.n_quartet == 68_523_973
    idx, n_cls = stratified_sample(batch, 2000, seed=3)
    got = gather_blocks(out, batch, idx)
    ref, sub = oracle_blocks(batch, idx)
    n_bad, worst, emax = block_bar_check(got, ref, sub)
    print("FP64 vs oracle: %d quartets over %d classes, %d elements, max abs err %.3e, worst err/bar %.3f"
          % (len(idx), n_cls, len(ref), emax, worst), flush=True)
    assert n_cls == 21
    assert n_bad == 0


def test_full_molecule_learned_routing(mol3):
    import torch
    mol, pf, pp, batch, plan, out64 = mol3
    t0 = time.perf_counter()
    # the benchmark's router: 2 M training quartets from two molecules, the
    # margins calibrated on 1 M quartets from each of two held-out molecules
    # (a router trained on 0.6 M quartets and calibrated on one molecule
    # under-predicted a few rare one-centre (ds|ds)-type quartets by eight
    # decades and skipped blocks of up to 9e-9)
    router, stats = screening.build_router(n_train=3_000_000, train_seeds=(202611, 202612),
                                           val_seed=(202619, 202618), n_val=1_000_000, n_atoms=200, epochs=25,
                                           under_weight=150.0, tau_skip=5e-11)
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
    assert codes[2] > 0.2
    # with the primitive-pair screen (eps = 1e-11) on top
    plan.set_prim_screen(plan.prim_screen_cut(1e-11))
    try:
        plan.run_routed(route, got)
        err_ps = max_abs_diff(got, out64)
    finally:
        plan.set_prim_screen(0.0)
    print("routed + primitive screen: max abs err vs FP64 %.3e" % err_ps, flush=True)
    assert err_ps <= 1e-10
    del got
