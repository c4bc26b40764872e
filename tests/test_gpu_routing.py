"""GPU: FP32-recurrence kernels and device-side routing.

* every quartet routed to the FP32-recurrence kernels stays within a stated
  absolute error of the FP64 path.  The FP32 error follows the magnitude of
  the recurrence intermediates, not of the (possibly cancelled) result, so it
  is bounded absolutely and relative to the Schwarz-type scale
  sqrt(max|(ab|ab)| max|(cd|cd)|) of the quartet;
* a random route (FP64 / FP32 / skip per quartet) gives exactly the FP64 result
  on FP64 quartets, zeros on skipped ones and the FP32 result on the rest, and
  a captured graph replays it identically (route read at replay time).
"""
import numpy as np
import pytest

from conftest import assert_parity
from intception_amd import basis, runtime

pytestmark = pytest.mark.gpu


def _blocks(batch, arr):
    sz = batch.sizes()
    return [arr[o:o + n] for o, n in zip(batch.out_off, sz)]


def test_fp32_recurrences_error(gpu):
    import torch
    b = basis.config1(n_quartet=20000, seed=99)
    plan = runtime.EriPlan(b)
    ref = plan.run().cpu().numpy()
    route = torch.ones(b.n_quartet, dtype=torch.int8, device=plan.device)
    got = plan.run_routed(route).cpu().numpy()
    torch.cuda.synchronize()
    counts = plan.route_counts()
    assert counts[:, 0].sum() == 0 and counts[:, 1].sum() == b.n_quartet
    err = np.abs(got - ref)
    # Schwarz scale per quartet from the diagonal quartets (ab|ab), (cd|cd)
    q = b.quartets
    diag = basis.QuartetBatch(True, b.shell_l, b.shell_nprim, b.shell_off, b.shell_cen, b.exps, b.coefs,
                              np.concatenate([q[:, [0, 1, 0, 1]], q[:, [2, 3, 2, 3]]]).astype(np.int32))
    d = runtime.compute(diag)
    dmax = np.array([np.abs(x).max() for x in _blocks(diag, d)])
    scale = np.sqrt(dmax[:b.n_quartet] * dmax[b.n_quartet:])
    eq = np.array([e.max() for e in _blocks(b, err)])
    ratio = eq / np.maximum(scale, 1e-300)
    print("FP32 recurrences: max abs err %.3e; max err / Schwarz scale %.3e (p99 %.3e)"
          % (err.max(), ratio.max(), np.percentile(ratio, 99)), flush=True)
    assert err.max() < 1e-5
    assert ratio.max() < 1e-4


def test_random_route_and_graph(gpu):
    import torch
    b = basis.config1(n_quartet=8000, seed=5)
    plan = runtime.EriPlan(b)
    ref64 = plan.run().cpu().numpy()
    all32 = plan.run_routed(torch.ones(b.n_quartet, dtype=torch.int8, device=plan.device)).cpu().numpy()
    rng = np.random.default_rng(1)
    codes = rng.integers(0, 3, size=b.n_quartet).astype(np.int8)
    route = torch.from_numpy(codes).to(plan.device)
    out = torch.full((b.out_size,), 7.0, dtype=torch.float64, device=plan.device)
    got = plan.run_routed(route, out).cpu().numpy()
    for q, (g, r64, r32) in enumerate(zip(_blocks(b, got), _blocks(b, ref64), _blocks(b, all32))):
        if codes[q] == 0:
            assert np.array_equal(g, r64)
        elif codes[q] == 1:
            assert np.array_equal(g, r32)
        else:
            assert not g.any()
    out2 = torch.full((b.out_size,), 3.0, dtype=torch.float64, device=plan.device)
    plan.capture(out2, n_streams=4, route=route)
    for rep in range(3):
        codes2 = rng.integers(0, 3, size=b.n_quartet).astype(np.int8)
        route.copy_(torch.from_numpy(codes2).to(plan.device))
        torch.cuda.synchronize()
        plan.replay()
        torch.cuda.synchronize()
        g2 = out2.cpu().numpy()
        eager = plan.run_routed(route).cpu().numpy()
        assert np.array_equal(g2, eager), "graph replay %d differs from eager" % rep
