"""Quartet-magnitude network (12 -> 64 -> 64 -> 1, ReLU, BF16 weights, FP32
accumulation) that tightens Schwarz screening and routes quartets between the
FP64 and FP32-recurrence kernels.

Host side only: the forward, backward, Adam and routing arithmetic all run in
libintc's HIP kernels (csrc/intc_mlp.hip); torch provides device memory.

* ``pair_features(mol)``: per shell pair log10 Q, log10 zeta_min, |AB|, L, the
  centre of the most diffuse primitive pair and the pair's lowest multipole
  order (8 floats).
* ``training_set(...)``: quartets of separate molecules (their own seeds),
  targets log10(max|integral| / (Q_bra Q_ket)) from the device FP64 path (the
  decay below the Schwarz bound; the network's output is clamped at 0, so its
  bound is never looser than Schwarz).
* ``MagnitudeNet.fit`` trains on the device; ``calibrate`` sets the margins
  from held-out quartets so that skipped quartets stay below the error budget;
  ``route`` writes the per-quartet route codes on the device.
"""
import ctypes
import json
import os

import numpy as np

from . import lib as _lib

NF, KF, NH = 12, 16, 64
OW1, OB1 = 0, NH * KF
OW2 = OB1 + NH
OB2 = OW2 + NH * NH
OW3 = OB2 + NH
OB3 = OW3 + NH
NPARAM = OB3 + 1
BW1, BW2, BW3 = 0, NH * KF, NH * KF + NH * NH
NB16 = BW3 + NH
LOG_FLOOR = -20.0

_c = ctypes.c_void_p


def _L():
    L = _lib.load()
    if not getattr(L, "_mlp_typed", False):
        f, vp, ll, i = ctypes.c_float, ctypes.c_void_p, ctypes.c_longlong, ctypes.c_int
        L.intc_mlp_features.argtypes = [ll, vp, vp, vp, vp, vp]
        L.intc_mlp_route.argtypes = [ll, vp, vp, vp, vp, vp, f, f, f, f, f, vp, vp, vp]
        L.intc_mlp_train_step.argtypes = [i, vp, vp, vp, vp, vp, vp, vp, f, vp]
        L.intc_mlp_adam.argtypes = [i, vp, vp, vp, vp, vp, f, f, f, f, i, vp, vp]
        for n in ("intc_mlp_features", "intc_mlp_route", "intc_mlp_train_step", "intc_mlp_adam",
                  "intc_mlp_num_params", "intc_mlp_num_bf16", "intc_mlp_train_block"):
            getattr(L, n).restype = ctypes.c_int
        assert L.intc_mlp_num_params() == NPARAM and L.intc_mlp_num_bf16() == NB16
        L._mlp_typed = True
    return L


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p()


# ---------------------------------------------------------------------------
# features
# ---------------------------------------------------------------------------
def pair_features(shell_l, shell_nprim, shell_off, shell_cen, exps, pairs, Q):
    """(n_pair, 8) float32: log10 Q, log10 zeta_min, |AB|, la+lb, P (3), and the
    pair's lowest multipole order: |la - lb| for a one-centre pair (its charge
    distribution has no moment below that order, e.g. a d*s pair on one atom
    starts at the quadrupole), 0 for a two-centre pair."""
    a, b = pairs[:, 0], pairs[:, 1]
    emin = np.array([exps[o:o + n].min() for o, n in zip(shell_off, shell_nprim)])
    za, zb = emin[a], emin[b]
    z = za + zb
    A, B = shell_cen[a], shell_cen[b]
    P = (za[:, None] * A + zb[:, None] * B) / z[:, None]
    ab = np.linalg.norm(A - B, axis=1)
    out = np.empty((len(pairs), 8), dtype=np.float32)
    out[:, 0] = np.log10(np.maximum(Q, 1e-300))
    out[:, 1] = np.log10(z)
    out[:, 2] = ab
    out[:, 3] = shell_l[a] + shell_l[b]
    out[:, 4:7] = P
    out[:, 7] = np.where(ab < 1e-8, np.abs(shell_l[a] - shell_l[b]), 0)
    return out


def quartet_features_np(pf, qp):
    """Numpy twin of the device feature function (unstandardised)."""
    a, c = pf[qp[:, 0]].astype(np.float64), pf[qp[:, 1]].astype(np.float64)
    r2 = np.sum((a[:, 4:7] - c[:, 4:7]) ** 2, axis=1)
    za, zc = 10.0 ** a[:, 1], 10.0 ** c[:, 1]
    rho = za * zc / (za + zc)
    T = rho * r2
    sT = np.sqrt(np.maximum(T, 1e-300))
    from scipy.special import erf
    f0 = np.where(T < 1e-6, 1.0, 0.886226925452758 * erf(sT) / sT)
    decay = np.log10(f0) + 0.25 * np.log10(4 * za * zc) - 0.5 * np.log10(za + zc)
    R = np.sqrt(r2)
    mdecay = -(a[:, 7] + c[:, 7]) * np.log10(1.0 + R * np.sqrt(rho))
    return np.stack([a[:, 0], c[:, 0], a[:, 1], c[:, 1], a[:, 2], c[:, 2], a[:, 3], c[:, 3], R,
                     np.log10(rho), decay, mdecay], axis=1)


def log_schwarz_product(F):
    """log10(Q_bra Q_ket) from unstandardised quartet features."""
    return F[:, 0] + F[:, 1]


def block_log_max(batch, out):
    """log10 max|integral| per quartet block (floored)."""
    mx = np.maximum.reduceat(np.abs(out), batch.out_off.astype(np.int64))
    return np.log10(np.maximum(mx, 10.0 ** LOG_FLOOR))


# ---------------------------------------------------------------------------
# bf16 helpers and the numpy reference of the device network
# ---------------------------------------------------------------------------
def bf16_round(x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def to_bf16_bits(x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def unpack(master):
    W1 = master[OW1:OB1].reshape(NH, KF)
    return W1, master[OB1:OW2], master[OW2:OB2].reshape(NH, NH), master[OB2:OW3], master[OW3:OB3], master[OB3]


def reference_forward_backward(master, X, T, rounding=True, under_weight=1.0):
    """FP64 numpy restatement of intc_mlp_train_k + the Adam gradient sum:
    forward with BF16-rounded weights and activations, MSE loss, gradient
    w.r.t. every master parameter (straight-through for the rounding).
    rounding=False gives the exact FP64 network (for finite differences)."""
    rnd = (lambda v: bf16_round(v).astype(np.float64)) if rounding else (lambda v: np.asarray(v, np.float64))
    W1, b1, W2, b2, w3, b3 = (np.asarray(v, dtype=np.float64) for v in unpack(np.asarray(master, np.float64)))
    W1b, W2b, w3b, Xb = rnd(W1), rnd(W2), rnd(w3), rnd(X)
    n = len(X)
    z1 = Xb @ W1b.T + b1
    h1 = np.maximum(z1, 0)
    h1b = rnd(h1)
    z2 = h1b @ W2b.T + b2
    h2 = np.maximum(z2, 0)
    h2b = rnd(h2)
    y = h2b @ w3b + b3
    r = y - T
    w = np.where(r < 0, under_weight, 1.0)
    loss = np.mean(w * r * r)
    dy = 2 * w * r / n
    d2 = (dy[:, None] * w3b[None, :]) * (h2 > 0)
    d1 = (d2 @ W2b) * (h1 > 0)
    g = np.zeros(NPARAM)
    gW1 = d1.T @ Xb
    gW1[:, NF:] = 0
    g[OW1:OB1] = gW1.reshape(-1)
    g[OB1:OW2] = d1.sum(0)
    g[OW2:OB2] = (d2.T @ h1b).reshape(-1)
    g[OB2:OW3] = d2.sum(0)
    g[OW3:OB3] = dy @ h2b
    g[OB3] = dy.sum()
    return y, loss, g


# ---------------------------------------------------------------------------
# the network on the device
# ---------------------------------------------------------------------------
class MagnitudeNet:
    def __init__(self, device="cuda", seed=0):
        import torch
        self.torch = torch
        from .runtime import _cuda_device
        self.device = _cuda_device(torch, device)
        rng = np.random.default_rng(seed)
        m = np.zeros(NPARAM, dtype=np.float32)
        m[OW1:OB1] = (rng.normal(0, np.sqrt(2.0 / NF), size=(NH, KF)) * (np.arange(KF) < NF)).reshape(-1)
        m[OW2:OB2] = rng.normal(0, np.sqrt(2.0 / NH), size=NH * NH)
        m[OW3:OB3] = rng.normal(0, np.sqrt(1.0 / NH), size=NH)
        m[OB3] = -8.0
        self.norm = np.concatenate([np.zeros(NF), np.ones(NF)]).astype(np.float32)
        self._upload(m)
        self.margin_skip = 0.0
        self.margin_32 = 0.0   # the FP32 thresholds are calibrated on the predictions themselves
        self.step = 0

    def _upload(self, master):
        t = self.torch
        self.master = t.from_numpy(np.ascontiguousarray(master, dtype=np.float32)).to(self.device)
        self.m1 = t.zeros(NPARAM, dtype=t.float32, device=self.device)
        self.m2 = t.zeros(NPARAM, dtype=t.float32, device=self.device)
        self.wb = t.zeros(NB16, dtype=t.int16, device=self.device)
        self.d_norm = t.from_numpy(self.norm).to(self.device)
        self._refresh_bf16()

    def _refresh_bf16(self):
        # lr = 0 Adam call: copies the master weights into the BF16 buffer
        t = self.torch
        part = t.zeros(NPARAM + 1, dtype=t.float32, device=self.device)
        _lib.check(_L().intc_mlp_adam(1, _p(part), _p(self.master), _p(self.m1), _p(self.m2), _p(self.wb), 0.0, 0.9,
                                      0.999, 1e-8, 1, None, self._stream()))

    def _stream(self):
        return ctypes.c_void_p(self.torch.cuda.current_stream(self.device).cuda_stream)

    def master_numpy(self):
        return self.master.cpu().numpy()

    def set_norm(self, mean, scale):
        self.norm = np.concatenate([mean, 1.0 / scale]).astype(np.float32)
        self.d_norm = self.torch.from_numpy(self.norm).to(self.device)

    def features(self, d_pair_feat, d_qp):
        t = self.torch
        n = d_qp.shape[0]
        X = t.empty((n, KF), dtype=t.float32, device=self.device)
        _lib.check(_L().intc_mlp_features(n, _p(d_pair_feat), _p(d_qp), _p(self.d_norm), _p(X), self._stream()))
        return X

    def grad(self, X, T, idx, under_weight=1.0):
        """Loss and summed gradient of one minibatch (device kernels)."""
        t = self.torch
        nb = idx.shape[0]
        nblk = -(-nb // _L().intc_mlp_train_block())
        part = t.empty(nblk * (NPARAM + 1), dtype=t.float32, device=self.device)
        loss = t.zeros(1, dtype=t.float32, device=self.device)
        _lib.check(_L().intc_mlp_train_step(nb, _p(X), _p(T), _p(idx), _p(self.wb), _p(self.master), _p(part),
                                            _p(loss), float(under_weight), self._stream()))
        g = t.empty(NPARAM, dtype=t.float32, device=self.device)
        _lib.check(_L().intc_mlp_adam(nblk, _p(part), _p(self.master), _p(self.m1), _p(self.m2), _p(self.wb), 0.0,
                                      0.9, 0.999, 1e-8, 1, _p(g), self._stream()))
        return loss, g

    def fit(self, X, T, epochs=8, batch=4096, lr=2e-3, log=None, seed=0, under_weight=1.0):
        """Adam on the device over shuffled minibatches of (X, T); squared
        errors of under-predictions weigh under_weight."""
        t = self.torch
        n = X.shape[0]
        gen = t.Generator(device=self.device)
        gen.manual_seed(seed)
        tb = _L().intc_mlp_train_block()
        part = t.empty((-(-batch // tb)) * (NPARAM + 1), dtype=t.float32, device=self.device)
        loss = t.zeros(1, dtype=t.float32, device=self.device)
        history = []
        for ep in range(epochs):
            perm = t.randperm(n, generator=gen, device=self.device).to(t.int32)
            loss.zero_()
            steps = 0
            lr_ep = lr * (0.5 ** max(0, ep - epochs // 2))
            for s in range(0, n, batch):
                idx = perm[s:s + batch]
                nb = idx.shape[0]
                nblk = -(-nb // tb)
                self.step += 1
                _lib.check(_L().intc_mlp_train_step(nb, _p(X), _p(T), _p(idx), _p(self.wb), _p(self.master),
                                                    _p(part), _p(loss), float(under_weight), self._stream()))
                _lib.check(_L().intc_mlp_adam(nblk, _p(part), _p(self.master), _p(self.m1), _p(self.m2),
                                              _p(self.wb), float(lr_ep), 0.9, 0.999, 1e-8, self.step, None,
                                              self._stream()))
                steps += 1
            mse = float(loss.item()) / steps
            history.append(mse)
            if log:
                log("mlp epoch %d: train mse %.4f (lr %.1e)" % (ep, mse, lr_ep))
        return history

    def predict(self, d_pair_feat, d_qp, route=None, m_skip=0.0, log_skip=-99.0, m_32=0.0, log_32=-99.0,
                log_qq32=-99.0):
        """Predicted log10 max|integral| = log10(Q_bra Q_ket) + min(0, network)."""
        t = self.torch
        n = d_qp.shape[0]
        pred = t.empty(n, dtype=t.float32, device=self.device)
        _lib.check(_L().intc_mlp_route(n, _p(d_pair_feat), _p(d_qp), _p(self.d_norm), _p(self.wb),
                                       _p(self.master), float(m_skip), float(log_skip), float(m_32), float(log_32),
                                       float(log_qq32), _p(pred), _p(route), self._stream()))
        return pred

    def route(self, d_pair_feat, d_qp, route_out, log_skip, log_32, log_qq32, pred_out=None):
        """Route codes on the device (0 FP64, 1 FP32, 2 skip) with the
        calibrated margins."""
        n = d_qp.shape[0]
        _lib.check(_L().intc_mlp_route(n, _p(d_pair_feat), _p(d_qp), _p(self.d_norm), _p(self.wb),
                                       _p(self.master), float(self.margin_skip), float(log_skip),
                                       float(self.margin_32), float(log_32), float(log_qq32), _p(pred_out),
                                       _p(route_out), self._stream()))

    def calibrate(self, pred, true, log_thresh, pad=0.25):
        """Skip margin from held-out quartets: the smallest margin under which
        no held-out quartet whose true magnitude reaches 10**log_thresh is
        predicted below it (pred + margin >= log_thresh for every such
        quartet), plus a pad.  Large under-predictions of big quartets that
        stay far above the threshold do not widen it."""
        pred = np.asarray(pred, dtype=np.float64)
        need = np.asarray(true, dtype=np.float64) >= log_thresh
        gap = (log_thresh - pred[need]).max() if need.any() else 0.0
        self.margin_skip = float(max(0.0, gap) + pad)
        return self.margin_skip

    def save(self, path):
        np.savez(path, master=self.master_numpy(), norm=self.norm,
                 margins=np.array([self.margin_skip, self.margin_32]))

    def load(self, path):
        z = np.load(path, allow_pickle=False)
        self.norm = z["norm"].astype(np.float32)
        self._upload(z["master"])
        self.margin_skip, self.margin_32 = (float(v) for v in z["margins"])
        return self
