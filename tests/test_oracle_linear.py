"""Pins the oracle's linear algebra (RCS LDLt, Schur solve, dense LDLt, 3x3 inverse)
against the reference's own fixtures and against numpy.

Fixtures: tests/golden/reference_fixtures.json, restated from
crates/sophus_solver/src/test_examples/positive_semidefinite/*.rs,
crates/sophus_solver/src/ldlt/schur_ldlt.rs:1045-1115 and
crates/sophus_solver/src/ldlt/dense_ldlt.rs:342-420 by tests/golden/make_fixtures.py.
"""
import json
from pathlib import Path

import numpy as np
import pytest

FIX = json.loads((Path(__file__).parent / "golden" / "reference_fixtures.json").read_text())


def _dense(sysd):
    n = sysd["n"]
    A = np.zeros((n, n))
    for i, j, v in sysd["A_lower_nonzeros"]:
        A[i, j] = v
        A[j, i] = v
    return A


@pytest.mark.parametrize("name", ["small", "single_partition", "scalar_pattern", "medium"])
def test_block_ldlt_on_reference_psd_systems(orc, name):
    """positive_semidefinite.rs:36-66,118-141: A x = b within 1e-6 for every solver."""
    s = FIX["psd_systems"][name]
    A = _dense(s)
    b = np.array(s["b"])
    x = orc.block_ldlt_solve_dense(A, s["block_dims"], b)
    assert x is not None
    assert np.max(np.abs(A @ x - b)) <= 1e-6
    # tighter: these systems are well conditioned, so the LDLt solve is exact to rounding
    assert np.max(np.abs(x - np.linalg.solve(A, b))) <= 1e-12 * max(1.0, np.max(np.abs(x)))


def test_scalar_pattern_fixture_matches_reference_dense():
    """scalar_pattern.rs:55-67 pins the dense matrix of L L^T."""
    A = _dense(FIX["psd_systems"]["scalar_pattern"])
    assert np.max(np.abs(A - np.array(FIX["scalar_pattern_dense"]))) < 1e-12


def test_schur_vs_dense_equivalence(orc):
    """schur_ldlt.rs:1084-1115: Schur solve == dense H^-1 g (1e-9)."""
    f = FIX["schur_vs_dense"]
    H, g = np.array(f["H"]), np.array(f["g"])
    dx = orc.schur_solve_dense(H, f["block_dims"], f["num_free_blocks"], g)
    ref = np.linalg.inv(H) @ g
    assert np.allclose(dx, ref, atol=1e-9, rtol=1e-9)


def test_schur_solve_on_random_ba_shaped_systems(orc):
    """Larger Schur systems (6-dim free blocks, 3-dim marg blocks coupling only to free
    blocks) against numpy's dense solve."""
    rng = np.random.default_rng(0)
    nf, nm = 5, 40
    n = 6 * nf + 3 * nm
    J = np.zeros((4 * nm, n))
    for b in range(nm):
        cams = rng.choice(nf, 2, replace=False)
        for k, c in enumerate(cams):
            J[4 * b + 2 * k:4 * b + 2 * k + 2, 6 * c:6 * c + 6] = rng.normal(size=(2, 6))
            J[4 * b + 2 * k:4 * b + 2 * k + 2, 6 * nf + 3 * b:6 * nf + 3 * b + 3] = rng.normal(size=(2, 3))
    H = J.T @ J + np.eye(n)
    g = rng.normal(size=n)
    dx = orc.schur_solve_dense(H, [6] * nf + [3] * nm, nf, g)
    assert np.max(np.abs(dx - np.linalg.solve(H, g))) < 1e-10 * np.max(np.abs(dx))


def _reconstruct(L, d):
    n = L.shape[0]
    Lu = np.tril(L, -1) + np.eye(n)
    return Lu @ np.diag(d) @ Lu.T


def test_dense_ldlt_reference_cases(orc):
    """dense_ldlt.rs:342-420: SPD reconstructs with positive pivots; rank-2 PSD has one
    zero pivot; the indefinite 2x2 solves like LU with one negative pivot."""
    f = FIX["dense_ldlt"]
    A = np.array(f["spd"]["A"])
    L, d, rank, df, pc = orc.dense_ldlt(A)
    assert rank == 3 and df == 0 and pc > f["spd"]["min_pivot_condition"]
    assert np.linalg.norm(_reconstruct(L, d) - A) <= 1e-9 and np.all(d > 0)

    A = np.array(f["psd"]["A"])
    L, d, rank, df, pc = orc.dense_ldlt(A)
    assert rank == 2 and df == 1
    assert np.linalg.norm(_reconstruct(L, d) - A) <= 1e-9
    assert int(np.sum(d == 0.0)) == 1

    A = np.array(f["indefinite"]["A"])
    L, d, rank, df, pc = orc.dense_ldlt(A)
    assert rank == 2 and df == 2 and pc > 0.1 and d[0] > 0 and d[1] < 0
    x = orc.dense_ldlt_solve(L, d, np.array(f["indefinite"]["b"]))
    assert np.allclose(x, np.linalg.solve(A, f["indefinite"]["b"]), atol=1e-10)


def test_block_ldlt_zero_pivot_directions_solve_to_zero(orc):
    """A PSD block system with a rank-deficient diagonal block: the solve returns the
    minimum-norm-like solution with zero in the null directions (dense_ldlt.rs:164-256)."""
    rng = np.random.default_rng(1)
    B = rng.normal(size=(6, 4))
    A = np.zeros((12, 12))
    A[:6, :6] = B @ B.T
    A[6:, 6:] = np.eye(6) * 3.0
    b = A @ rng.normal(size=12)
    x = orc.block_ldlt_solve_dense(A, [6, 6], b)
    assert np.allclose(A @ x, b, atol=1e-9)


def test_inverse3_matches_numpy_and_flags_singular(orc):
    rng = np.random.default_rng(2)
    for _ in range(20):
        M = rng.normal(size=(3, 3))
        M = M @ M.T + 0.1 * np.eye(3)
        assert np.allclose(orc.inverse3(M), np.linalg.inv(M), rtol=1e-12, atol=1e-12)
    assert orc.inverse3(np.zeros((3, 3))) is None
    # rank 2: the adjugate formula's determinant is exactly zero here
    assert orc.inverse3(np.array([[1.0, 2.0, 3.0], [2.0, 4.0, 6.0], [0.0, 1.0, 1.0]])) is None
