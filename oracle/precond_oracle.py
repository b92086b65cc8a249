"""ORACLE — test infrastructure only.

numpy float64 restatement of the learned block-Jacobi correction of the reduced camera
system: the block features, the 3-layer MLP (forward and hand-derived backward), the
preconditioner blocks M_b^-1 = B G G^T B^T and the training loss.  The model has no
counterpart in the reference (it is the build's addition on top of the Schur solver);
its gradients are pinned by torch autograd in the CPU tests.
"""
from __future__ import annotations

import numpy as np

FEAT, HID, OUT = 32, 128, 21
NW = HID * FEAT + HID + HID * HID + HID + OUT * HID + OUT


def views(w):
    w = np.asarray(w, dtype=np.float64).reshape(-1)
    out, o = {}, 0
    for name, shape in (("W1", (HID, FEAT)), ("b1", (HID,)), ("W2", (HID, HID)), ("b2", (HID,)),
                        ("W3", (OUT, HID)), ("b3", (OUT,))):
        n = int(np.prod(shape))
        out[name] = w[o:o + n].reshape(shape)
        o += n
    return out


def flatten(d) -> np.ndarray:
    return np.concatenate([d[k].reshape(-1) for k in ("W1", "b1", "W2", "b2", "W3", "b3")])


def features(A: np.ndarray) -> np.ndarray:
    """Per 6x6 block row of a dense symmetric matrix: J D J correlations above the
    diagonal (15), log d_i - mean (6), log1p(#off-diagonal blocks)/4, log1p of the sum,
    max and sum of squares of |J_b S_bc J_c|_F over structurally nonzero blocks."""
    nb = A.shape[0] // 6
    D = [A[6 * b:6 * b + 6, 6 * b:6 * b + 6] for b in range(nb)]
    J = [1.0 / np.sqrt(np.diag(Db)) for Db in D]
    F = np.zeros((nb, FEAT))
    for b in range(nb):
        Db, Jb = D[b], J[b]
        f = [Db[j, i] * Jb[i] * Jb[j] for i in range(6) for j in range(i + 1, 6)]
        ld = np.log(np.diag(Db))
        f += list(ld - ld.mean())
        norms = []
        for c in range(nb):
            if c == b:
                continue
            blk = A[6 * b:6 * b + 6, 6 * c:6 * c + 6]
            if np.any(blk != 0):
                norms.append(np.linalg.norm(Jb[:, None] * blk * J[c][None, :]))
        norms = np.array(norms) if norms else np.zeros(0)
        f.append(np.log1p(len(norms)) / 4.0)
        f.append(np.log1p(norms.sum()))
        f.append(np.log1p(norms.max() if len(norms) else 0.0))
        f.append(np.log1p((norms ** 2).sum()))
        F[b, :len(f)] = f
    return F


def forward(F, w):
    v = views(w)
    X = np.asarray(F, dtype=np.float64)
    H1 = np.tanh(X @ v["W1"].T + v["b1"])
    H2 = np.tanh(H1 @ v["W2"].T + v["b2"])
    O = H2 @ v["W3"].T + v["b3"]
    return O, (X, H1, H2)


def g_matrix(o):
    G = np.zeros((6, 6))
    idx = 0
    for i in range(6):
        for j in range(i + 1):
            G[i, j] = np.exp(o[idx]) if i == j else o[idx]
            idx += 1
    return G


def whiten(D):
    """B = J L^-T with J = diag(D)^-1/2 and J D J = L L^T."""
    J = 1.0 / np.sqrt(np.diag(D))
    L = np.linalg.cholesky(J[:, None] * D * J[None, :])
    return J[:, None] * np.linalg.inv(L).T


def minv(D, o):
    B = whiten(D)
    G = g_matrix(o)
    return B @ G @ G.T @ B.T


def loss_grad(F, Ds, u, w, weights):
    """sum_b |G G^T B^T u_b - w_b|^2 / |w_b|^2 and its gradient w.r.t. the weights."""
    O, (X, H1, H2) = forward(F, weights)
    n = O.shape[0]
    dO = np.zeros_like(O)
    total = 0.0
    for b in range(n):
        B = whiten(Ds[b])
        ut = B.T @ u[b]
        G = g_matrix(O[b])
        t = G.T @ ut
        f = G @ t - w[b]
        s = G.T @ f
        ww = w[b] @ w[b]
        total += f @ f / ww
        dG = 2.0 / ww * (np.outer(f, t) + np.outer(ut, s))
        idx = 0
        for i in range(6):
            for j in range(i + 1):
                dO[b, idx] = dG[i, j] * (G[i, i] if i == j else 1.0)
                idx += 1
    v = views(weights)
    g = {"W3": dO.T @ H2, "b3": dO.sum(0)}
    dA2 = (dO @ v["W3"]) * (1 - H2 ** 2)
    g["W2"], g["b2"] = dA2.T @ H1, dA2.sum(0)
    dA1 = (dA2 @ v["W2"]) * (1 - H1 ** 2)
    g["W1"], g["b1"] = dA1.T @ X, dA1.sum(0)
    return total, flatten(g)
