"""Writes the linear-algebra and Lie fixtures of the reference's own tests as data.

Every matrix here is restated from the construction in the reference test it names;
the output, `reference_fixtures.json`, holds only inputs and the expected properties the
reference asserts (tolerances included).  Run:  python tests/golden/make_fixtures.py
"""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent


def schur_spd(n_free: int, n_marg: int, marg_block_dim: int) -> np.ndarray:
    """crates/sophus_solver/src/ldlt/schur_ldlt.rs:1045-1078 (build_schur_spd_matrix)."""
    n = n_free + n_marg
    h = np.zeros((n, n))
    for i in range(n_free):
        h[i, i] = 10.0 + i
        if i + 1 < n_free:
            h[i, i + 1] = h[i + 1, i] = 1.0
    for b in range(n_marg // marg_block_dim):
        off = n_free + b * marg_block_dim
        for i in range(marg_block_dim):
            h[off + i, off + i] = 10.0 + off + i
            if i + 1 < marg_block_dim:
                h[off + i, off + i + 1] = h[off + i + 1, off + i] = 0.5
    for i in range(min(n_free, n_marg)):
        h[i, n_free + i] = h[n_free + i, i] = 0.3
    return h


def small_system():
    """crates/sophus_solver/src/test_examples/positive_semidefinite/small.rs:16-75."""
    a = np.array([[3.3, 1.0, 0.0, 0.0, 0.0, 1.0],
                  [1.0, 3.2, 1.0, 0.0, 0.0, 0.0],
                  [0.0, 1.0, 3.1, 0.0, 0.0, 0.0],
                  [0.0, 0.0, 0.0, 2.3, 0.0, 0.0],
                  [0.0, 0.0, 0.0, 0.0, 2.2, 0.0],
                  [1.0, 0.0, 0.0, 0.0, 0.0, 2.1]])
    return a, [3, 3], np.arange(1.0, 7.0)


def single_partition_system():
    """crates/sophus_solver/src/test_examples/positive_semidefinite/single_partition.rs:17-74."""
    nb, d = 4, 2
    n = nb * d
    L = np.zeros((n, n))
    for j in range(nb):
        for k in range(d):
            L[j * d + k, j * d + k] = 1.0
        if j > 0:
            for c in range(d):
                for r in range(d):
                    L[j * d + r, (j - 1) * d + c] = 0.05 * (1.0 + r + c)
    return L @ L.T, [d] * nb, 1.0 + 0.1 * np.arange(n)


def scalar_pattern_system():
    """crates/sophus_solver/src/test_examples/positive_semidefinite/scalar_pattern.rs:15-72."""
    n = 8
    L = np.eye(n)
    for i in range(1, n):
        L[i, i - 1] = 0.2
    for i in range(2, n):
        L[i, i - 2] = 0.05
    return L @ L.T, [1] * n, np.array([1.0, -1.0, 2.0, -2.0, 0.5, -0.5, 3.0, -3.0])


def medium_system():
    """crates/sophus_solver/src/test_examples/positive_semidefinite/medium.rs:22-160:
    a 5x10 grid of 4-dim blocks and a 60-chain of 3-dim blocks, coupled at (1:0, 0:0)."""
    i4 = np.array([[1.0, 0.01, 0.02, 0.01], [0.01, 1.1, 0.02, 0.0],
                   [0.02, 0.02, 1.2, 0.0], [0.01, 0.0, 0.0, 1.3]])
    four_i4 = np.array([[4.0, 0.001, 0.1, 0.002], [0.001, 4.2, 0.0009, 0.0],
                        [0.1, 0.0009, 4.3, 0.0], [0.002, 0.0, 0.0, 4.0]])
    i3 = np.array([[1.0, 0.02, 0.0], [0.02, 1.1, 0.01], [0.0, 0.01, 1.2]])
    three_i3 = np.array([[3.0, 0.001, 0.0], [0.001, 3.1, 0.0007], [0.0, 0.0007, 3.2]])
    c10 = np.array([[0.1, 0.0, 0.0, 0.0], [0.0, 0.08, 0.01, 0.0], [0.0, 0.01, 0.07, 0.0]])
    R0, C0, NB1 = 5, 10, 60
    NB0 = R0 * C0
    n0 = 4 * NB0
    n = n0 + 3 * NB1
    A = np.zeros((n, n))

    def blk0(i):
        return slice(4 * i, 4 * i + 4)

    def blk1(i):
        return slice(n0 + 3 * i, n0 + 3 * i + 3)

    def add_lower(rs, cs, m):
        A[rs, cs] += m
        if rs != cs:
            A[cs, rs] += m.T

    for i in range(NB0):
        add_lower(blk0(i), blk0(i), four_i4)

    def edge(blk, a, b, c):
        i, j = max(a, b), min(a, b)
        add_lower(blk(i), blk(j), -c)
        add_lower(blk(i), blk(i), c)
        add_lower(blk(j), blk(j), c)

    for r in range(R0):
        for c in range(C0):
            i = r * C0 + c
            if r + 1 < R0:
                edge(blk0, i, (r + 1) * C0 + c, i4)
            if c + 1 < C0:
                edge(blk0, i, r * C0 + c + 1, i4)
    for i in range(NB1):
        add_lower(blk1(i), blk1(i), three_i3)
    for i in range(NB1 - 1):
        edge(blk1, i, i + 1, i3)
    add_lower(blk1(0), blk0(0), -c10)
    add_lower(blk1(0), blk1(0), i3)
    add_lower(blk0(0), blk0(0), i4)
    return A, [4] * NB0 + [3] * NB1, np.ones(n)


def lower_nonzeros(a: np.ndarray) -> list:
    """[row, col, value] of the nonzeros on and below the diagonal."""
    r, c = np.nonzero(np.tril(a))
    return [[int(i), int(j), float(a[i, j])] for i, j in zip(r, c)]


def main() -> None:
    out = {}
    systems = {}
    for name, fn, src in (
            ("small", small_system, "test_examples/positive_semidefinite/small.rs:16"),
            ("single_partition", single_partition_system,
             "test_examples/positive_semidefinite/single_partition.rs:17"),
            ("scalar_pattern", scalar_pattern_system,
             "test_examples/positive_semidefinite/scalar_pattern.rs:15"),
            ("medium", medium_system, "test_examples/positive_semidefinite/medium.rs:22")):
        a, dims, b = fn()
        systems[name] = {"source": "crates/sophus_solver/src/" + src, "n": a.shape[0],
                         "A_lower_nonzeros": lower_nonzeros(a),
                         "block_dims": dims, "b": b.tolist(),
                         "asserts": "|A x - b|_inf <= 1e-6 (positive_semidefinite.rs:63,135)"}
    out["psd_systems"] = systems
    # the scalar_pattern test pins its dense matrix explicitly (scalar_pattern.rs:55-67)
    out["scalar_pattern_dense"] = [
        [1.000, 0.2000, 0.0500, 0.0, 0.0, 0.0, 0.0, 0.0],
        [0.200, 1.0400, 0.2100, 0.0500, 0.0, 0.0, 0.0, 0.0],
        [0.050, 0.2100, 1.0425, 0.2100, 0.0500, 0.0, 0.0, 0.0],
        [0.0, 0.0500, 0.2100, 1.0425, 0.2100, 0.0500, 0.0, 0.0],
        [0.0, 0.0, 0.0500, 0.2100, 1.0425, 0.2100, 0.0500, 0.0],
        [0.0, 0.0, 0.0, 0.0500, 0.2100, 1.0425, 0.2100, 0.0500],
        [0.0, 0.0, 0.0, 0.0, 0.0500, 0.2100, 1.0425, 0.2100],
        [0.0, 0.0, 0.0, 0.0, 0.0, 0.0500, 0.2100, 1.0425]]
    h = schur_spd(4, 6, 2)
    out["schur_vs_dense"] = {
        "source": "crates/sophus_solver/src/ldlt/schur_ldlt.rs:1045-1115",
        "H": h.tolist(), "block_dims": [2, 2, 2, 2, 2], "num_free_blocks": 2,
        "g": ((np.arange(10) + 1.0) * 0.3).tolist(),
        "asserts": "dx == H^-1 g within 1e-9 (abs and rel)"}
    m = np.array([[1.0, 2.0, 3.0], [0.0, 1.0, 4.0], [5.0, 6.0, 0.0]])
    mm = np.array([[1.0, 0.0, 1.0], [0.0, 1.0, 1.0]])
    out["dense_ldlt"] = {
        "source": "crates/sophus_solver/src/ldlt/dense_ldlt.rs:342-420",
        "spd": {"A": (m.T @ m + 0.1 * np.eye(3)).tolist(), "rank": 3, "definiteness": "positive_definite",
                "min_pivot_condition": 0.01, "reconstruction_tol": 1e-9},
        "psd": {"A": (mm.T @ mm).tolist(), "rank": 2, "definiteness": "positive_semidefinite",
                "zero_pivots": 1, "reconstruction_tol": 1e-9},
        "indefinite": {"A": [[1.0, 2.0], [2.0, 1.0]], "rank": 2, "definiteness": "indefinite",
                       "min_pivot_condition": 0.1, "b": [1.0, 2.0], "solve_tol": 1e-10}}
    out["rotation3_tangent_examples"] = {
        "source": "crates/sophus_lie/src/groups/rotation3.rs:325-337",
        "tangents": [[0.0, 0.0, 0.0], [0.0001, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0],
                     [0.0, 0.0, 1.0], [0.5, 0.5, 0.1], [-0.1, -0.5, -0.5], [2.5, 0.5, 0.1],
                     [0.5, 2.5, 0.1], [0.5, 0.1, -2.5]]}
    (HERE / "reference_fixtures.json").write_text(json.dumps(out, indent=1))
    print("wrote", HERE / "reference_fixtures.json")


if __name__ == "__main__":
    main()
