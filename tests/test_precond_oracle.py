"""Pins the numpy oracle of the learned block-Jacobi correction: its hand-derived
gradient against torch autograd (float64, CPU), its preconditioner blocks against the
closed form at zero output (block-Jacobi: M^-1 = D^-1) and its features on a small
hand-checkable matrix."""
import numpy as np
import pytest
import torch

import precond_oracle as po


def _spd_blocks(rng, n):
    Ds = []
    for _ in range(n):
        A = rng.normal(size=(6, 6))
        Ds.append(A @ A.T + 6 * np.eye(6) * np.exp(rng.normal(size=6)))
    return np.array(Ds)


def _torch_loss(F, Ds, u, w, weights):
    wt = torch.tensor(weights, dtype=torch.float64, requires_grad=True)
    v, o = {}, 0
    for name, shape in (("W1", (128, 32)), ("b1", (128,)), ("W2", (128, 128)), ("b2", (128,)),
                        ("W3", (21, 128)), ("b3", (21,))):
        n = int(np.prod(shape))
        v[name] = wt[o:o + n].reshape(shape)
        o += n
    X = torch.tensor(F, dtype=torch.float64)
    H1 = torch.tanh(X @ v["W1"].T + v["b1"])
    H2 = torch.tanh(H1 @ v["W2"].T + v["b2"])
    O = H2 @ v["W3"].T + v["b3"]
    total = torch.zeros((), dtype=torch.float64)
    tril = [(i, j) for i in range(6) for j in range(i + 1)]
    for b in range(F.shape[0]):
        B = torch.tensor(po.whiten(Ds[b]), dtype=torch.float64)
        G = torch.zeros(6, 6, dtype=torch.float64)
        for k, (i, j) in enumerate(tril):
            G = G.index_put((torch.tensor(i), torch.tensor(j)), torch.exp(O[b, k]) if i == j else O[b, k])
        ut = B.T @ torch.tensor(u[b], dtype=torch.float64)
        wb = torch.tensor(w[b], dtype=torch.float64)
        f = G @ (G.T @ ut) - wb
        total = total + (f @ f) / (wb @ wb)
    total.backward()
    return float(total.detach()), wt.grad.numpy()


def test_oracle_gradient_matches_torch_autograd():
    rng = np.random.default_rng(0)
    n = 7
    F = rng.normal(size=(n, 32)) * 0.5
    Ds = _spd_blocks(rng, n)
    u = rng.normal(size=(n, 6))
    w = rng.normal(size=(n, 6))
    weights = rng.normal(size=po.NW) * 0.1
    l_o, g_o = po.loss_grad(F, Ds, u, w, weights)
    l_t, g_t = _torch_loss(F, Ds, u, w, weights)
    assert abs(l_o - l_t) <= 1e-12 * abs(l_t)
    assert np.max(np.abs(g_o - g_t)) <= 1e-10 * np.max(np.abs(g_t))


def test_zero_output_is_block_jacobi():
    rng = np.random.default_rng(1)
    for D in _spd_blocks(rng, 5):
        M = po.minv(D, np.zeros(21))
        assert np.allclose(M, np.linalg.inv(D), rtol=1e-12, atol=1e-14 * np.abs(np.linalg.inv(D)).max())


def test_minv_is_spd_for_any_output():
    rng = np.random.default_rng(2)
    for D in _spd_blocks(rng, 5):
        M = po.minv(D, rng.normal(size=21))
        assert np.allclose(M, M.T, rtol=1e-12, atol=0)
        assert np.linalg.eigvalsh(M).min() > 0


def test_features_on_a_two_block_matrix():
    """Two blocks coupled by S_10 = c I: correlations 0 for diagonal D, log-scales of
    diag(D), one neighbour, coupling norm |c| sqrt(sum_i 1/(d0_i d1_i))."""
    d0 = np.array([1.0, 2.0, 3.0, 4.0, 5.0, 6.0])
    d1 = np.array([2.0, 2.0, 2.0, 8.0, 8.0, 8.0])
    A = np.zeros((12, 12))
    A[:6, :6] = np.diag(d0)
    A[6:, 6:] = np.diag(d1)
    A[6:, :6] = 0.5 * np.eye(6)
    A[:6, 6:] = 0.5 * np.eye(6)
    F = po.features(A)
    nrm = 0.5 * np.sqrt(np.sum(1.0 / (d0 * d1)))
    assert np.allclose(F[0, :15], 0.0)
    assert np.allclose(F[0, 15:21], np.log(d0) - np.log(d0).mean())
    assert np.isclose(F[0, 21], np.log1p(1) / 4)
    assert np.allclose(F[0, 22:25], [np.log1p(nrm), np.log1p(nrm), np.log1p(nrm ** 2)])
    assert np.allclose(F[0, 25:], 0.0)
