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
    """FP32 recurrences on the config-1 classes: relative error per class.
    Measured on MI355X (20 000 quartets, seed 99): median relative error
    1e-8..2e-5 and p99 up to 3e-4 for kets with lc+ld <= 3; kets with
    lc+ld = 4 reach p99 1e-2..0.4 -- the FP64 ket HRR amplifies the FP32
    error of its (e|f) inputs by ~|CD|^ld on distant, tight pairs -- which is
    why the router only sends quartets to this path below calibrated
    magnitude and Schwarz-product thresholds."""
    import torch
    from intception_amd.codegen import canonical
    b = basis.config1(n_quartet=20000, seed=99)
    plan = runtime.EriPlan(b)
    ref = plan.run().cpu().numpy()
    route = torch.ones(b.n_quartet, dtype=torch.int8, device=plan.device)
    got = plan.run_routed(route).cpu().numpy()
    torch.cuda.synchronize()
    counts = plan.route_counts()
    assert counts[:, 0].sum() == 0 and counts[:, 1].sum() == b.n_quartet
    assert np.all(np.isfinite(got))
    err = np.abs(got - ref)
    eq = np.maximum.reduceat(err, b.out_off)
    mq = np.maximum.reduceat(np.abs(ref), b.out_off)
    can = np.array([canonical(*c)[0] for c in map(tuple, b.classes())])
    print("FP32 recurrences: max abs err %.3e" % err.max(), flush=True)
    assert err.max() < 1e-5
    for c in sorted(set(map(tuple, can))):
        sel = np.all(can == np.array(c), axis=1) & (mq > 1e-12)
        rel = eq[sel] / mq[sel]
        print("  class %s: median rel %.1e p99 rel %.1e" % (c, np.median(rel), np.percentile(rel, 99)), flush=True)
        assert np.median(rel) < 1e-4
        if c[2] + c[3] <= 3:
            assert np.percentile(rel, 99) < 1e-3


@pytest.mark.parametrize("plan_sorted", [False, True])
def test_random_route_and_graph(gpu, plan_sorted):
    """Random route codes: exact FP64 / FP32 / zero blocks, queue lengths, and
    graph replays equal to eager runs.  plan_sorted lists the quartets in the
    plan's own order (identity permutation: route read and skipped blocks
    zeroed in place), otherwise the route is read through the permutation."""
    import torch
    b = basis.config1(n_quartet=8000, seed=5)
    if plan_sorted:
        p0 = runtime.EriPlan(b)
        order = p0.quartet_order()
        p0.close()
        assert sorted(order.tolist()) == list(range(b.n_quartet))
        b = b.subset(order)
    plan = runtime.EriPlan(b)
    assert bool(np.all(plan.quartet_order() == np.arange(b.n_quartet))) == plan_sorted
    ref64 = plan.run().cpu().numpy()
    all32 = plan.run_routed(torch.ones(b.n_quartet, dtype=torch.int8, device=plan.device)).cpu().numpy()
    rng = np.random.default_rng(1)
    codes = rng.integers(0, 3, size=b.n_quartet).astype(np.int8)
    route = torch.from_numpy(codes).to(plan.device)
    out = torch.full((b.out_size,), 7.0, dtype=torch.float64, device=plan.device)
    got = plan.run_routed(route, out).cpu().numpy()
    counts = plan.route_counts()
    assert np.array_equal(counts.sum(0), np.bincount(codes, minlength=3))
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


def test_fp32_error_check_redo(gpu):
    """FP32 error check with every quartet routed to FP32 (config-1 classes):
    a tolerance below every estimate sends every quartet of the FP32 register
    kernels to the FP64 redo, so the output equals the plain FP64 output bit
    for bit; tolerance 0 (off) gives the plain FP32 result; a tolerance of
    1e-11 redoes the quartets whose estimate exceeds it and keeps every
    element within 1e-11 of FP64, alone and in a captured graph."""
    import torch
    b = basis.config1(n_quartet=20000, seed=71)
    plan = runtime.EriPlan(b)
    ref = plan.run().cpu().numpy()
    route = torch.ones(b.n_quartet, dtype=torch.int8, device=plan.device)
    plain32 = plan.run_routed(route).cpu().numpy()
    plan.set_fp32_check(1e-300)
    got = plan.run_routed(route).cpu().numpy()
    redo = plan.redo_counts()
    counts = plan.route_counts()
    print("tolerance 1e-300: %d of %d quartets redone in FP64" % (redo.sum(), b.n_quartet), flush=True)
    # every quartet whose largest intermediate is a normal FP32 number is
    # redone; the others (intermediates below 1.2e-38, where the FP32 maximum
    # reads zero) keep their FP32 result, whose error is as small
    diff = got != ref
    print("  elements not equal to FP64: %d, largest |FP64| among them %.1e, largest difference %.1e, "
          "non-finite FP32-route elements without the check %d"
          % (diff.sum(), np.abs(ref[diff]).max() if diff.any() else 0.0,
             np.abs(got[diff] - ref[diff]).max() if diff.any() else 0.0, (~np.isfinite(plain32)).sum()), flush=True)
    assert not diff.any() or (np.abs(ref[diff]).max() < 1e-30 and np.abs(got[diff] - ref[diff]).max() < 1e-30)
    assert np.all(redo <= counts[:, 1]) and redo.sum() > 0.9 * b.n_quartet
    plan.set_fp32_check(0.0)
    got = plan.run_routed(route).cpu().numpy()
    assert np.array_equal(got, plain32) and plan.redo_counts().sum() == 0
    plan.set_fp32_check(1e-11)
    got = plan.run_routed(route).cpu().numpy()
    n_redo = int(plan.redo_counts().sum())
    err = np.abs(got - ref).max()
    print("tolerance 1e-11: %d quartets redone, max abs err %.3e (plain FP32: %.3e)"
          % (n_redo, err, np.abs(plain32 - ref).max()), flush=True)
    assert err <= 1e-11
    out = plan.alloc_out()
    plan.capture(out, route=route)
    for _ in range(3):
        out.zero_()
        plan.replay()
        torch.cuda.synchronize()
        assert np.array_equal(out.cpu().numpy(), got)
    plan.close()
