"""The quartet-magnitude network.

CPU: the FP64 numpy restatement of the training kernel's backward pass agrees
with central finite differences of its own loss (a dropped term or transposed
operand in the backward formulas fails this).
GPU: the MFMA inference kernel and the training kernel agree with that
restatement (same BF16 rounding of weights and activations, FP32
accumulation); the feature kernel agrees with numpy; training lowers the loss.
"""
import numpy as np
import pytest

from intception_amd import mlp


def _rand_master(seed=0):
    rng = np.random.default_rng(seed)
    m = np.zeros(mlp.NPARAM)
    m[mlp.OW1:mlp.OB1] = (rng.normal(0, 0.4, size=(mlp.NH, mlp.KF)) * (np.arange(mlp.KF) < mlp.NF)).reshape(-1)
    m[mlp.OB1:mlp.OW2] = rng.normal(0, 0.1, size=mlp.NH)
    m[mlp.OW2:mlp.OB2] = rng.normal(0, 0.18, size=mlp.NH * mlp.NH)
    m[mlp.OB2:mlp.OW3] = rng.normal(0, 0.1, size=mlp.NH)
    m[mlp.OW3:mlp.OB3] = rng.normal(0, 0.12, size=mlp.NH)
    m[mlp.OB3] = -3.0
    return m


def test_reference_backward_finite_differences():
    rng = np.random.default_rng(1)
    m = _rand_master(2)
    X = np.zeros((40, mlp.KF))
    X[:, :mlp.NF] = rng.normal(size=(40, mlp.NF))
    T = rng.normal(-5, 2, size=40)
    _, loss, g = mlp.reference_forward_backward(m, X, T, rounding=False, under_weight=3.0)
    idx = list(range(mlp.OW1, mlp.OW1 + 24)) + list(range(mlp.OB1, mlp.OB1 + 5)) + \
        list(range(mlp.OW2, mlp.OW2 + 40)) + list(range(mlp.OB2, mlp.OB2 + 5)) + \
        list(range(mlp.OW3, mlp.OB3)) + [mlp.OB3]
    h = 1e-6
    for i in idx:
        mp, mm = m.copy(), m.copy()
        mp[i] += h
        mm[i] -= h
        fd = (mlp.reference_forward_backward(mp, X, T, False, 3.0)[1] -
              mlp.reference_forward_backward(mm, X, T, False, 3.0)[1]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_bf16_rounding():
    x = np.array([1.0, 1.00390625, 1.005859375, -3.14159, 1e-30], dtype=np.float32)
    r = mlp.bf16_round(x)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0078125   # ties to even
    assert abs(r[3] + 3.140625) < 1e-7


@pytest.mark.gpu
def test_mlp_kernels_match_reference(gpu):
    import torch
    rng = np.random.default_rng(3)
    n_pair, n = 500, 4000
    pf = np.zeros((n_pair, 8), dtype=np.float32)
    pf[:, 0] = rng.uniform(-8, 1, n_pair)
    pf[:, 1] = rng.uniform(-1, 3, n_pair)
    pf[:, 2] = rng.uniform(0, 6, n_pair)
    pf[:, 3] = rng.integers(0, 5, n_pair)
    pf[:, 4:7] = rng.uniform(-10, 10, (n_pair, 3))
    pf[:, 7] = rng.integers(0, 3, n_pair) * (rng.uniform(0, 1, n_pair) < 0.3)
    qp = rng.integers(0, n_pair, (n, 2)).astype(np.int32)
    F = mlp.quartet_features_np(pf, qp)
    net = mlp.MagnitudeNet(seed=4)
    net.set_norm(F.mean(0), F.std(0) + 1e-6)
    master = _rand_master(5).astype(np.float32)
    net._upload(master)
    d_pf = torch.from_numpy(pf).cuda()
    d_qp = torch.from_numpy(qp).cuda()
    X = net.features(d_pf, d_qp).cpu().numpy()
    Xref = np.zeros((n, mlp.KF), dtype=np.float32)
    Xref[:, :mlp.NF] = (F - F.mean(0)) / (F.std(0) + 1e-6)
    assert np.abs(X - Xref).max() < 1e-4
    pred = net.predict(d_pf, d_qp).cpu().numpy()
    T = rng.normal(-6, 2, size=n).astype(np.float32)
    yref, lref, gref = mlp.reference_forward_backward(master, X, T, under_weight=10.0)
    # the routing kernel reports log10 Q_bra Q_ket + min(0, network output)
    yabs = mlp.log_schwarz_product(F) + np.minimum(yref, 0.0)
    err = np.abs(pred - yabs)
    print("mfma forward: max |err| %.2e (|y| ~ %.1f)" % (err.max(), np.abs(yref).mean()), flush=True)
    assert err.max() < 2e-3 * (1 + np.abs(yabs).max())
    d_X = torch.from_numpy(X).cuda()
    d_T = torch.from_numpy(T).cuda()
    idx = torch.arange(n, dtype=torch.int32, device="cuda")
    loss, g = net.grad(d_X, d_T, idx, under_weight=10.0)
    g = g.cpu().numpy()
    print("train step: loss %.5f ref %.5f; grad max rel err %.2e" % (loss.item(), lref,
          np.abs(g - gref).max() / np.abs(gref).max()), flush=True)
    assert abs(loss.item() - lref) <= 1e-3 * lref
    assert np.abs(g - gref).max() <= 2e-3 * np.abs(gref).max()


@pytest.mark.gpu
def test_mlp_training_lowers_loss(gpu):
    import torch
    rng = np.random.default_rng(7)
    n = 20000
    X = np.zeros((n, mlp.KF), dtype=np.float32)
    X[:, :mlp.NF] = rng.normal(size=(n, mlp.NF))
    T = (X[:, 0] * 2 - X[:, 1] ** 2 + 0.5 * X[:, 2] * X[:, 3] - 6).astype(np.float32)
    net = mlp.MagnitudeNet(seed=1)
    hist = net.fit(torch.from_numpy(X).cuda(), torch.from_numpy(T).cuda(), epochs=6, batch=1024, lr=3e-3)
    print("training mse by epoch", ["%.4f" % h for h in hist], flush=True)
    assert hist[-1] < 0.1 * hist[0]
