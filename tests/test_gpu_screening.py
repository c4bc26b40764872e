"""GPU: learned screening and routing end to end on a small molecule.

A router trained on two small molecules of their own seeds and calibrated on a
third routes the screened quartets of a fourth; the routed result stays within
the 1e-10 absolute error budget of the plain FP64 result, and the router
skips more quartets than Schwarz screening raised to the same skip threshold
(the method without the network) would.
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
    # what the network gains over the method without it: Schwarz screening
    # raised to the same skip threshold (the Schwarz bound makes that safe);
    # the router skips on either test, so it never skips less, and the
    # network adds the quartets whose bound is loose
    qq = mol.Q[pp[:, 0]] * mol.Q[pp[:, 1]]
    schwarz = qq < 10 ** router.log_skip
    skipped = route.cpu().numpy() == 2
    # clear of the kernel's 1e-4-decade guard for its FP32 logarithms
    surely = np.log10(np.maximum(qq, 1e-300)) < router.log_skip - 2e-4
    print("skip fraction: router %.3f, Schwarz at the same threshold %.3f, by the network only %.3f"
          % (skipped.mean(), schwarz.mean(), (skipped & ~schwarz).mean()), flush=True)
    assert np.all(skipped[surely])
    assert (skipped & ~schwarz).mean() > 0.01


def test_three_center_molecule(gpu):
    """Config-4 style: (ab|P s) quartets of a small molecule with an auxiliary
    basis.  FP64 parity of a sample against the oracle (bar as for the
    three-centre classes: diffuse auxiliary shells far apart lose up to ~1e-11
    of the block maximum to rounding in the oracle's own scheme), then the
    learned router (trained on three-centre molecules of their own seeds) keeps
    the routed result within 1e-10 of FP64."""
    import torch
    from conftest import assert_parity  # noqa: F401
    from intception_amd import basis
    from oracle import oracle
    mol, pf = screening.molecule_setup(10, 41, three_center=True)
    pp = mol.screened_quartets(1e-12)
    assert len(pp) > 0
    samp = pp[np.random.default_rng(0).choice(len(pp), size=min(1500, len(pp)), replace=False)]
    b = mol.pair_batch(samp)
    got = runtime.compute(b)
    oracle.set_boys_mode(oracle.BOYS_EXACT)
    ref = oracle.eri_batch(b, True, basis.SPH)
    sz = b.sizes()
    for q in range(b.n_quartet):
        o, n = b.out_off[q], sz[q]
        g, r = got[o:o + n], ref[o:o + n]
        lim = 1e-14 + 1e-12 * np.abs(r) + 5e-11 * np.abs(r).max()
        assert np.all(np.abs(g - r) <= lim), (q, np.abs(g - r).max(), np.abs(r).max())
    router, stats = screening.build_router(n_train=150_000, train_seeds=(51, 52), val_seed=59, n_val=50_000,
                                           n_atoms=10, epochs=6, three_center=True)
    print("3c router stats:", {k: v for k, v in stats.items() if k != "train_mse"}, flush=True)
    pp = screening.plan_sorted(mol, pp)
    b = mol.pair_batch(pp)
    plan = runtime.EriPlan(b)
    ref = plan.run()
    route = torch.empty(b.n_quartet, dtype=torch.int8, device="cuda")
    router.route(torch.from_numpy(pf).cuda(), torch.from_numpy(pp.astype(np.int32)).cuda(), route)
    got = plan.run_routed(route)
    torch.cuda.synchronize()
    err = (got - ref).abs().max().item()
    codes = np.bincount(route.cpu().numpy().astype(np.int64), minlength=3) / b.n_quartet
    print("3c screened quartets %d; fp64/fp32/skip = %s; max abs err %.3e" % (b.n_quartet, codes, err), flush=True)
    assert err <= 1e-10
