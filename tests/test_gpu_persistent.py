"""GPU: the persistent multi-class launch (intc_plan_set_persistent).

One launch runs every spherical register-class bin as tiles taken from a
device counter, each tile with its class's code and the bin's team size.  Its
output must equal that of the per-class launches (the same class code, so to
the FP64 parity bar 1e-14 + 1e-12 |ref|) for plain and routed runs, at tile
sizes that cut the bins at every point, for the three-centre classes with f
and g auxiliary shells, and in captured graphs replayed several times.  One
case is also checked against the oracle directly.
"""
import numpy as np
import pytest

from conftest import assert_parity
from intception_amd import basis, lib, runtime

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("tile_iters", [1, 7, 128])
def test_persistent_matches_per_class(gpu, tile_iters):
    b = basis.config1(n_quartet=6000, seed=11)
    print("persistent plain: config-1 classes, 6000 quartets, seed 11, tile_iters %d" % tile_iters, flush=True)
    plan = runtime.EriPlan(b)
    ref = plan.run().cpu().numpy()
    plan.set_persistent(True, tile_iters)
    got = plan.run().cpu().numpy()
    n_tiles, reg = plan.persistent_info()
    print("  %d tiles, %d LDS doubles per wave" % (n_tiles, reg), flush=True)
    assert n_tiles > 0 and reg > 0
    assert_parity(got, ref, what="persistent vs per-class (tile_iters %d)" % tile_iters)
    plan.close()


def test_persistent_matches_oracle(gpu):
    from oracle import oracle
    b = basis.config1(n_quartet=400, seed=31)
    print("persistent vs oracle: 400 quartets, seed 31", flush=True)
    plan = runtime.EriPlan(b)
    plan.set_persistent(True, 4)
    got = plan.run().cpu().numpy()
    plan.close()
    oracle.set_boys_mode(oracle.BOYS_EXACT)
    ref = oracle.eri_batch(b, True, basis.SPH)
    assert_parity(got, ref, what="persistent vs oracle")


def test_persistent_three_center(gpu):
    """(ab|P s) with auxiliary shells up to g: FP64 bodies of the f/g classes
    inside the multi-class kernel."""
    b = basis.random_three_center(7, 3000)
    print("persistent three-centre: 3000 quartets, seed 7", flush=True)
    plan = runtime.EriPlan(b)
    ref = plan.run().cpu().numpy()
    plan.set_persistent(True, 16)
    got = plan.run().cpu().numpy()
    plan.close()
    assert_parity(got, ref, what="persistent three-centre")


@pytest.mark.parametrize("use_fp32", [True, False])
def test_persistent_routed_and_graph(gpu, use_fp32):
    """Random routes (FP64 / FP32 / skip): persistent eager and graph replays
    equal the per-class routed run; FP32 queues of the f/g classes run the
    FP64 body as in the per-class path."""
    import torch
    b = basis.config1(n_quartet=8000, seed=5)
    print("persistent routed: 8000 quartets, seed 5, fp32 %s" % use_fp32, flush=True)
    plan = runtime.EriPlan(b)
    plan.set_route_fp32(use_fp32)
    rng = np.random.default_rng(3)
    codes = rng.integers(0, 3, size=b.n_quartet).astype(np.int8)
    route = torch.from_numpy(codes).to(plan.device)
    ref = plan.run_routed(route).cpu().numpy()
    ref_counts = plan.route_counts()
    plan.set_persistent(True, 8)
    got = plan.run_routed(route).cpu().numpy()
    assert np.array_equal(plan.route_counts(), ref_counts)
    assert_parity(got, ref, what="persistent routed eager")
    out = torch.full((plan.out_size,), float("nan"), dtype=torch.float64, device=plan.device)
    plan.capture(out, n_streams=4, route=route)
    for k in range(3):
        out.fill_(float("nan"))
        plan.replay()
        torch.cuda.synchronize(plan.device)
        assert_parity(out.cpu().numpy(), ref, what="persistent routed graph replay %d" % k)
    plan.close()


def test_persistent_plain_graph(gpu):
    import torch
    b = basis.config1(n_quartet=5000, seed=8)
    print("persistent plain graph: 5000 quartets, seed 8", flush=True)
    plan = runtime.EriPlan(b)
    ref = plan.run().cpu().numpy()
    plan.set_persistent(True, 32)
    out = torch.full((plan.out_size,), float("nan"), dtype=torch.float64, device=plan.device)
    plan.capture(out, n_streams=2)
    for k in range(3):
        out.fill_(float("nan"))
        plan.replay()
        torch.cuda.synchronize(plan.device)
        assert_parity(out.cpu().numpy(), ref, what="persistent plain graph replay %d" % k)
    plan.close()


def test_persistent_rejections(gpu):
    """Cartesian plans, caller-supplied matrices and the reference Boys
    scheme are refused before anything is launched."""
    b = basis.config0(n_per_class=50, seed=3)
    plan = runtime.EriPlan(b)
    with pytest.raises(lib.IntcError):
        plan.set_persistent(True)
    plan.close()
    b = basis.config1(n_quartet=100, seed=4)
    plan = runtime.EriPlan(b, sph=runtime.sph_blob())
    plan.set_persistent(True)
    with pytest.raises(lib.IntcError):
        plan.run()
    plan.close()
    plan = runtime.EriPlan(b)
    plan.set_persistent(True)
    runtime.set_boys_mode(1)
    try:
        with pytest.raises(lib.IntcError):
            plan.run()
    finally:
        runtime.set_boys_mode(0)
    plan.close()


def test_persistent_subset_of_bins(gpu):
    """intc_plan_set_persistent_bins: the persistent launch holds only the
    selected bins, the others run their own class kernels beside it; plain,
    routed and captured runs give the per-class output bit for bit."""
    import torch
    b = basis.config1(n_quartet=6000, seed=13)
    print("persistent subset: config-1 classes, 6000 quartets, seed 13", flush=True)
    plan = runtime.EriPlan(b)
    ref = plan.run().cpu().numpy()
    nb = len(plan.bins())
    rng = np.random.default_rng(3)
    cover = rng.random(nb) < 0.5
    plan.set_persistent(True, 8)
    plan.set_persistent_bins(cover)
    got = plan.run().cpu().numpy()
    assert np.array_equal(got, ref)
    route = torch.from_numpy(rng.integers(0, 3, size=b.n_quartet).astype(np.int8)).to(plan.device)
    plan.set_persistent(False)
    rref = plan.run_routed(route).cpu().numpy()
    plan.set_persistent(True, 8)
    rgot = plan.run_routed(route).cpu().numpy()
    assert np.array_equal(rgot, rref)
    out = plan.alloc_out()
    plan.capture(out, n_streams=4, route=route)
    for rep in range(3):
        out.fill_(float("nan"))
        plan.replay()
        torch.cuda.synchronize()
        assert np.array_equal(out.cpu().numpy(), rref), "graph replay %d" % rep
    plan.set_persistent_bins(None)
    assert np.array_equal(plan.run().cpu().numpy(), ref)
    with pytest.raises(Exception):
        plan.set_persistent_bins(np.ones(nb + 1, dtype=bool))
    plan.close()
