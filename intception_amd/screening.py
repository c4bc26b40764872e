"""Learned screening and FP32/FP64 routing for molecule-like workloads.

Pipeline (every numeric step on the device through libintc):

1. Schwarz factors Q_P per shell pair (FP64 class kernels on (ab|ab)).
2. Training data from molecules with their own seeds: quartets sampled
   uniformly from each molecule's Schwarz-screened set, target
   log10 max|integral| from the FP64 path, 12 pair-derived features.
3. ``MagnitudeNet`` trained on the device (BF16 weights, FP32 accumulation).
4. Calibration on a held-out molecule:
   * skip margin: the smallest margin under which every held-out quartet
     whose true magnitude reaches tau_skip keeps predicted + margin >=
     log10(tau_skip), plus a pad; a quartet is skipped only when
     predicted + margin < log10(tau_skip);
   * FP32 guard: the largest Schwarz product log10(Q_P Q_P') under which every
     held-out quartet's FP32-recurrence error stays below the FP32 budget.
5. Route codes per quartet on the device (0 FP64, 1 FP32, 2 skip), then the
   routed plan execution.
"""
import numpy as np

from . import basis, mlp, molecule, runtime


def molecule_setup(n_atoms, seed, tau=1e-12, three_center=False):
    """Molecule (config 3) or three-centre molecule (config 4) with Schwarz
    factors and the per-pair feature table."""
    mol = (molecule.ThreeCenterMolecule if three_center else molecule.Molecule)(n_atoms, seed=seed)
    mol.schwarz()
    pf = mlp.pair_features(mol.shell_l, mol.shell_nprim, mol.shell_off, mol.shell_cen, mol.exps, mol.pairs, mol.Q)
    return mol, pf


def plan_sorted(mol, pp):
    """The quartet list (pair-index pairs) reordered into the execution plan's
    order (class, then primitive count descending).  A batch listed in this
    order makes the plan order the identity, so route codes are read and
    skipped output blocks zeroed sequentially."""
    plan = runtime.EriPlan(mol.pair_batch(pp))
    order = plan.quartet_order()
    plan.close()
    return np.ascontiguousarray(pp[order])


def _device(a, torch, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a))
    return t.to("cuda") if dtype is None else t.to("cuda", dtype=dtype)


def sample_targets(mol, n, seed, tau=1e-12, fp32=False):
    """Uniform sample of the screened quartets with FP64 targets (and the
    FP32-recurrence error per quartet when fp32=True)."""
    import torch
    pp, total = mol.sample_screened(tau, n, seed=seed)
    b = mol.pair_batch(pp)
    plan = runtime.EriPlan(b)
    out = plan.run()
    off = torch.from_numpy(b.out_off).to(out.device)
    host = out.cpu().numpy()
    target = mlp.block_log_max(b, host)
    e32 = None
    if fp32:
        route = torch.ones(b.n_quartet, dtype=torch.int8, device=out.device)
        o32 = plan.run_routed(route).cpu().numpy()
        e32 = np.maximum.reduceat(np.abs(o32 - host), b.out_off.astype(np.int64))
    plan.close()
    del off
    return pp, target, e32, total


def training_set(n_samples, seeds, n_atoms=200, tau=1e-12, log=None, three_center=False):
    """Features and targets log10(max|integral| / (Q_P Q_P')) (clamped to
    [-25, 0]) for quartets of molecules with the given seeds."""
    feats, targets = [], []
    for s in seeds:
        mol, pf = molecule_setup(n_atoms, s, tau, three_center)
        pp, t, _, total = sample_targets(mol, n_samples // len(seeds), seed=s + 7, tau=tau)
        F = mlp.quartet_features_np(pf, pp).astype(np.float32)
        feats.append(F)
        targets.append(np.clip(t - mlp.log_schwarz_product(F), -25.0, 0.0).astype(np.float32))
        if log:
            log("training molecule seed %d: %d screened quartets, sampled %d" % (s, total, len(pp)))
    return np.concatenate(feats), np.concatenate(targets)


class LearnedRouter:
    """Trained network + thresholds; writes route codes for a quartet list."""

    def __init__(self, net, log_skip=-10.0, fp32_budget=1e-11):
        self.net = net
        self.log_skip = log_skip          # skip when predicted + margin < log10(tau_skip)
        self.fp32_budget = fp32_budget
        self.log_32 = log_skip + 2.0      # FP32 when predicted + margin below this ...
        self.log_qq32 = -99.0             # ... and log10(Q_P Q_P') below this (set by calibrate_fp32)

    def calibrate_fp32(self, pred, logqq, e32):
        """Thresholds (predicted magnitude, Schwarz product) for the FP32 route:
        on held-out quartets that are not skipped, the pair of thresholds
        (0.5-decade grid) that sends the most quartets to FP32 while every one
        of them keeps its FP32-recurrence error within the FP32 budget, and
        the budget still holds one grid step further in both thresholds.
        The FP32 error has a heavy tail (the HRR and the contraction amplify
        the rounding of the recurrence values by up to ~1e4 of the block
        maximum), so a threshold pair picked at the held-out set's edge can
        admit a far larger error on another molecule; the step of margin
        keeps the chosen pair half a decade inside the region that held."""
        m = self.net.margin_32
        keep = pred + self.net.margin_skip >= self.log_skip
        keep &= logqq >= self.log_skip - 1e-4        # the route kernel also skips on the Schwarz bound
        best = (0, -99.0, -99.0)
        for tp in np.arange(-14.0, 2.01, 0.5):
            sel_p = keep & (pred + m < tp)
            if not sel_p.any():
                continue
            wide_p = keep & (pred + m < tp + 0.5)
            for tq in np.arange(-16.0, 4.01, 0.5):
                sel = sel_p & (logqq < tq)
                n = int(sel.sum())
                wide = wide_p & (logqq < tq + 0.5)
                if n > best[0] and e32[wide].max() <= self.fp32_budget:
                    best = (n, tp, tq)
        _, self.log_32, self.log_qq32 = best
        return best

    def route(self, d_pf, d_qp, route_out, pred_out=None):
        self.net.route(d_pf, d_qp, route_out, self.log_skip, self.log_32, self.log_qq32, pred_out)

    def summary(self):
        return {"log10_tau_skip": self.log_skip, "margin": self.net.margin_skip, "log10_tau_fp32": self.log_32,
                "log10_schwarz_fp32_guard": self.log_qq32, "fp32_budget": self.fp32_budget}


def build_router(n_train=2_000_000, train_seeds=(202611, 202612), val_seed=202619, n_val=400_000, n_atoms=200,
                 tau=1e-12, tau_skip=5e-11, fp32_budget=1e-11, epochs=10, under_weight=10.0, log=None,
                 three_center=False):
    """Train and calibrate a router on molecules of their own seeds
    (three-centre molecules for the (ab|P s) workload)."""
    import time
    import torch
    t0 = time.perf_counter()
    X, T = training_set(n_train, train_seeds, n_atoms, tau, log, three_center)
    t_data = time.perf_counter() - t0
    net = mlp.MagnitudeNet()
    mean, scale = X.mean(0), X.std(0) + 1e-6
    net.set_norm(mean, scale)
    Xs = np.zeros((len(X), mlp.KF), dtype=np.float32)
    Xs[:, :mlp.NF] = (X - mean) / scale
    t0 = time.perf_counter()
    hist = net.fit(_device(Xs, torch), _device(T, torch), epochs=epochs, batch=4096, lr=2e-3, log=log,
                   under_weight=under_weight)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    # calibration on held-out molecules (val_seed: one seed or several; their
    # samples are pooled, n_val quartets from each)
    seeds = tuple(val_seed) if isinstance(val_seed, (tuple, list)) else (val_seed,)
    preds, tvs, e32s, logqqs = [], [], [], []
    for vs in seeds:
        mol, pf = molecule_setup(n_atoms, vs, tau, three_center)
        pp, tv, e32, total = sample_targets(mol, n_val, seed=vs + 7, tau=tau, fp32=True)
        d_pf, d_qp = _device(pf, torch), _device(pp.astype(np.int32), torch)
        preds.append(net.predict(d_pf, d_qp).cpu().numpy())
        tvs.append(tv)
        e32s.append(e32)
        logqqs.append(np.log10(np.maximum(mol.Q[pp[:, 0]] * mol.Q[pp[:, 1]], 1e-300)))
    pred, tv, e32, logqq = (np.concatenate(x) for x in (preds, tvs, e32s, logqqs))
    margin = net.calibrate(pred, tv, np.log10(tau_skip))
    router = LearnedRouter(net, np.log10(tau_skip), fp32_budget)
    n32, _, _ = router.calibrate_fp32(pred, logqq, e32)
    stats = {"train_samples": int(len(X)), "train_seconds": t_train, "data_seconds": t_data,
             "train_mse": hist, "val_rmse": float(np.sqrt(np.mean((pred - tv) ** 2))),
             "val_max_underprediction": float(np.max(tv - pred)), "margin": margin,
             "val_fp32_max_err": float(e32.max()), "val_fp32_fraction": n32 / len(pred),
             "val_samples": int(len(pred)),
             "val_skip_fraction": float(np.mean((pred + margin < np.log10(tau_skip)) |
                                                (logqq < np.log10(tau_skip) - 1e-4))),
             # the method without the network: Schwarz screening raised to tau_skip
             "val_skip_schwarz_only": float(np.mean(logqq < np.log10(tau_skip))),
             # what a perfect predictor could skip: true magnitude below tau_skip
             "val_skip_ideal": float(np.mean(tv < np.log10(tau_skip))), **router.summary()}
    if log:
        log("router: %s" % {k: v for k, v in stats.items() if k != "train_mse"})
    return router, stats
