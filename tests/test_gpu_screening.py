"""GPU: learned screening and routing end to end on a small molecule.

A router trained on two small molecules of their own seeds and calibrated on a
third routes the screened quartets of a fourth; the routed result stays within
the 1e-10 absolute error budget of the plain FP64 result, and the router
skips or demotes a share of the quartets that Schwarz screening alone keeps.
"""
import numpy as np
import pytest

from intception_amd import molecule, runtime, screening

pytestmark = pytest.mark.gpu


def test_learned_routing_small_molecule(gpu):
    import torch
    router, stats = screening.build_router(n_train=200_000, train_seeds=(11, 12), val_seed=19, n_val=60_000,
                                           n_atoms=40, epochs=6)
    print("router stats:", {k: v for k, v in stats.items() if k != "train_mse"}, flush=True)
    mol, pf = screening.molecule_setup(40, 23)
    pp = screening.plan_sorted(mol, mol.screened_quartets(1e-12))
    b = mol.pair_batch(pp)
    plan = runtime.EriPlan(b)
    ref = plan.run()
    route = torch.empty(b.n_quartet, dtype=torch.int8, device="cuda")
    d_pf = torch.from_numpy(pf).cuda()
    d_qp = torch.from_numpy(pp.astype(np.int32)).cuda()
    router.route(d_pf, d_qp, route)
    got = plan.run_routed(route)
    torch.cuda.synchronize()
    err = (got - ref).abs().max().item()
    codes = np.bincount(route.cpu().numpy().astype(np.int64), minlength=3) / b.n_quartet
    print("screened quartets %d; routed fp64/fp32/skip = %s; max abs err %.3e" % (b.n_quartet, codes, err),
          flush=True)
    assert err <= 1e-10
    assert codes[2] + codes[1] > 0.05
