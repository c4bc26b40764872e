"""GPU: the primitive-pair screen (intc_plan_set_prim_screen).

After the pair build, a device pass keeps the primitive pairs of each shell
pair whose s-type Schwarz factor w = |kappa| (2 zeta)^(-1/4) reaches the cut.
The class kernels then contract only the kept pairs.  Checks, on a
molecule-like batch (shared shell pairs, s/p/d shells with K up to 4):

* the kept counts equal the host's count of w >= cut;
* the screened integrals equal the oracle's contraction over exactly the kept
  primitive quartets (the full oracle value minus the oracle values of the
  dropped primitive quartets, evaluated as single-primitive shells), to the
  FP64 parity bar 1e-14 + 1e-12 |ref|;
* the change against the unscreened integrals stays within
  16 (D_bra W_ket + W_bra D_ket), where W is a pair's sum of w and D the sum of
  its dropped w;
* cut 0 after a screened build restores the unscreened output bit for bit;
* captured graphs and the persistent launch give the eager screened output;
* a cut above every w drops every primitive and the output is exactly zero.
"""
import numpy as np
import pytest

from conftest import assert_parity
from intception_amd import basis, molecule, runtime

pytestmark = pytest.mark.gpu


def _prim_w(m, a, b):
    ea = m.exps[m.shell_off[a]:m.shell_off[a] + m.shell_nprim[a]]
    ca = m.coefs[m.shell_off[a]:m.shell_off[a] + m.shell_nprim[a]]
    eb = m.exps[m.shell_off[b]:m.shell_off[b] + m.shell_nprim[b]]
    cb = m.coefs[m.shell_off[b]:m.shell_off[b] + m.shell_nprim[b]]
    r2 = float(np.sum((m.shell_cen[a] - m.shell_cen[b]) ** 2))
    z = ea[:, None] + eb[None, :]
    kap = 5.914967172795612878 * np.exp(-ea[:, None] * eb[None, :] / z * r2) * ca[:, None] * cb[None, :] / z
    return np.abs(kap) * (2.0 * z) ** -0.25


def _molecule_batch(n_quartet, seed):
    m = molecule.Molecule(n_atoms=24, seed=seed)
    m.pairs = m.prescreen_pairs()
    rng = np.random.default_rng(seed)
    pp = rng.integers(0, len(m.pairs), size=(n_quartet, 2))
    return m, m.pair_batch(pp)


def _cut_with_gap(ws, cut):
    """Move the cut off any w within 1e-9 (relative) so that the device and
    the host decide every pair alike."""
    allw = np.concatenate([w.ravel() for w in ws])
    while np.any(np.abs(allw - cut) <= 1e-9 * cut):
        cut *= 1.0 + 1e-6
    return cut


@pytest.fixture(scope="module")
def case(gpu):
    m, b = _molecule_batch(3000, 202641)
    ws = [_prim_w(m, a, c) for a, c in m.pairs]
    plan = runtime.EriPlan(b)
    full = plan.run().cpu().numpy()
    yield m, b, ws, plan, full
    plan.close()


@pytest.mark.parametrize("eps", [1e-12, 1e-8])
def test_screen_counts_and_bound(case, eps):
    m, b, ws, plan, full = case
    cut = _cut_with_gap(ws, plan.prim_screen_cut(eps))
    print("prim screen: 24-atom molecule, 3000 quartets, eps %g, cut %.3e" % (eps, cut), flush=True)
    plan.set_prim_screen(cut)
    got = plan.run().cpu().numpy()
    kept = plan.pair_prims(screened=True)
    allp = plan.pair_prims(screened=False)
    host = np.array([(w >= cut).sum() for w in ws], dtype=np.int32)
    print("  kept %d of %d primitive pairs" % (kept.sum(), allp.sum()), flush=True)
    assert np.array_equal(allp, np.array([w.size for w in ws], dtype=np.int32))
    assert np.array_equal(kept, host)
    assert kept.sum() < allp.sum()
    W = np.array([w.sum() for w in ws])
    D = np.array([w[w < cut].sum() for w in ws])
    qp = b.quartet_pairs
    bound = 16.0 * (D[qp[:, 0]] * W[qp[:, 1]] + W[qp[:, 0]] * D[qp[:, 1]])
    sz = b.sizes()
    lim = np.repeat(bound, sz) + 1e-14 + 1e-12 * np.abs(full)
    d = np.abs(got - full)
    print("  max |screened - full| %.3e, max ratio to bound %.3g" % (d.max(), (d / lim).max()), flush=True)
    assert np.all(d <= lim)
    plan.set_prim_screen(0.0)


def test_screen_matches_oracle_over_kept_primitives(case):
    """Parity with the oracle on the same kept primitive quartets."""
    from oracle import oracle
    m, b, ws, plan, full = case
    cut = _cut_with_gap(ws, plan.prim_screen_cut(1e-8))
    plan.set_prim_screen(cut)
    got = plan.run().cpu().numpy()
    plan.set_prim_screen(0.0)
    sel = np.arange(0, b.n_quartet, 10)
    sub = m.pair_batch(b.quartet_pairs[sel])
    oracle.set_boys_mode(oracle.BOYS_EXACT)
    ref = oracle.eri_batch(sub, True, basis.SPH)
    # single-primitive shells: one per (shell, primitive)
    sid, L, NP, OFF, CEN, EX, CO = {}, [], [], [], [], [], []
    for s in range(m.n_shell):
        for i in range(m.shell_nprim[s]):
            sid[(s, i)] = len(L)
            L.append(m.shell_l[s]); NP.append(1); OFF.append(len(EX)); CEN.append(m.shell_cen[s])
            EX.append(m.exps[m.shell_off[s] + i]); CO.append(m.coefs[m.shell_off[s] + i])
    items, owner = [], []
    for k, (P, R) in enumerate(sub.quartet_pairs):
        (a, bb), (c, d) = m.pairs[P], m.pairs[R]
        kb, kk = ws[P] >= cut, ws[R] >= cut
        for i, j in np.ndindex(kb.shape):
            for u, v in np.ndindex(kk.shape):
                if not (kb[i, j] and kk[u, v]):
                    items.append((sid[(a, i)], sid[(bb, j)], sid[(c, u)], sid[(d, v)]))
                    owner.append(k)
    print("screen vs oracle: %d quartets, %d dropped primitive quartets" % (len(sel), len(items)), flush=True)
    assert items
    tb = basis.QuartetBatch(True, np.array(L, np.int32), np.array(NP, np.int32), np.array(OFF, np.int32),
                            np.array(CEN), np.array(EX), np.array(CO), np.array(items, np.int32))
    dropped = oracle.eri_batch(tb, True, basis.SPH)
    sz, tsz = sub.sizes(), tb.sizes()
    for t, k in enumerate(owner):
        ref[sub.out_off[k]:sub.out_off[k] + sz[k]] -= dropped[tb.out_off[t]:tb.out_off[t] + tsz[t]]
    g = np.concatenate([got[b.out_off[q]:b.out_off[q] + sz[k]] for k, q in enumerate(sel)])
    assert_parity(g, ref, what="screened vs oracle over the kept primitives")


def test_screen_off_restores_and_paths_agree(case):
    m, b, ws, plan, full = case
    cut = _cut_with_gap(ws, plan.prim_screen_cut(1e-8))
    plan.set_prim_screen(cut)
    eager = plan.run().cpu().numpy()
    print("screen paths: graph replays and persistent launch vs eager", flush=True)
    out = plan.alloc_out()
    plan.capture(out, n_streams=4)
    import torch
    for rep in range(3):
        out.fill_(np.nan)
        plan.replay()
        torch.cuda.synchronize()
        assert np.array_equal(out.cpu().numpy(), eager), "graph replay %d" % rep
    plan.set_persistent(True, 16)
    pers = plan.run().cpu().numpy()
    plan.set_persistent(False)
    assert np.array_equal(pers, eager)
    plan.set_prim_screen(0.0)
    again = plan.run().cpu().numpy()
    assert np.array_equal(again, full)


def test_screen_everything_gives_zeros(gpu):
    """A cut above every w leaves every shell pair without primitives: every
    kernel (register and workgroup classes, f shells included) writes zeros."""
    b = basis.config2(n_quartet=200, seed=5)
    print("screen everything: config-2 classes, 200 quartets", flush=True)
    plan = runtime.EriPlan(b)
    plan.set_prim_screen(1e300)
    got = plan.run().cpu().numpy()
    assert int(plan.pair_prims(screened=True).sum()) == 0
    plan.close()
    assert np.all(got == 0.0)
