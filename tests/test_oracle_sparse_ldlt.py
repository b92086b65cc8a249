"""The oracle's SparseLdlt (the RCS solver of SchurSparseLdlt) on the CPU.

The oracle restates sparse_ldlt.rs (crates/sophus_solver/src/ldlt/sparse_ldlt.rs:121-445,
571-650): the scalar left-looking LDLt in the block order with a per-pivot rule,
|d_jj| <= max(|d_jj|, 1) 1e-12 -> 0 (an empty column of L), and a solve that divides by
every pivot.  It is pinned by:
- the reference's own test (sparse_ldlt.rs:690-760): [[eps, 1], [1, -eps]] with eps =
  1e-12 has both pivots at the threshold, so both are zero and the solve divides by zero;
  the reference asserts an error above 1e-6 against LU, here x = (inf, inf) exactly;
- SPD block matrices against numpy's solve (the factorization is exact in exact
  arithmetic, any order);
- constructed L D L^T blocks whose pivots are known: a pivot of 3e-7 after one of 1e6 is
  kept by the scalar rule (the solve matches numpy's) and zeroed by the block rule of
  BlockSparseLdlt (max(1, max |d|) 1e-12 = 1e-6); a pivot of 1e-13 is zero under both, so
  the scalar solve is not finite while the block one is;
- the LM step on the ring: with no pivot near either threshold both solvers give the
  same trajectory.
"""
import numpy as np
import pytest


def _spd_blocks(nb, band, seed):
    rng = np.random.default_rng(seed)
    n = 6 * nb
    A = np.zeros((n, n))
    for i in range(nb):
        for j in range(max(0, i - band), i + 1):
            B = rng.normal(size=(6, 6))
            A[6 * i:6 * i + 6, 6 * j:6 * j + 6] = B
            A[6 * j:6 * j + 6, 6 * i:6 * i + 6] = B.T
    A = 0.5 * (A + A.T) + np.eye(n) * (np.abs(A).sum(axis=1).max() + 1.0)
    return A


def _ldl_block(d, seed):
    rng = np.random.default_rng(seed)
    L = np.tril(rng.normal(0, 0.1, (6, 6)), -1) + np.eye(6)
    A = L @ np.diag(d) @ L.T
    return 0.5 * (A + A.T)


def test_reference_ill_conditioned_kkt(orc):
    """sparse_ldlt.rs:690-760: both pivots sit at the threshold (|eps| <= 1e-12) and are zero;
    the solve divides y = b by them."""
    eps = 1e-12
    A = np.array([[eps, 1.0], [1.0, -eps]])
    b = np.array([1.0, 2.0])
    x = orc.sparse_ldlt_solve_dense(A, [2], b)
    assert np.all(np.isinf(x)) and np.all(x > 0)
    x_ref = np.linalg.solve(A, b)
    with np.errstate(invalid="ignore"):
        assert not np.linalg.norm(x - x_ref) <= 1e-6  # the reference's assertion: error > 1e-6


@pytest.mark.parametrize("nb,band", [(1, 0), (9, 2), (40, 6)])
def test_spd_matches_numpy(orc, nb, band):
    A = _spd_blocks(nb, band, nb)
    b = np.random.default_rng(nb + 1).normal(size=6 * nb)
    x = orc.sparse_ldlt_solve_dense(A, [6] * nb, b)
    ref = np.linalg.solve(A, b)
    assert np.max(np.abs(x - ref)) / np.max(np.abs(ref)) < 1e-12


def test_per_pivot_rule_keeps_what_the_block_rule_drops(orc):
    d = np.array([1e6, 3e-7, 2.0, 5.0, 8.0, 7.0])
    A = _ldl_block(d, 3)
    b = np.random.default_rng(4).normal(size=6)
    xs = orc.sparse_ldlt_solve_dense(A, [6], b)
    xb = orc.block_ldlt_solve_dense(A, [6], b)
    ref = np.linalg.solve(A, b)
    # kept: the exact solve (condition ~3e12, so the bar is relative to that)
    assert np.all(np.isfinite(xs))
    assert np.max(np.abs(xs - ref)) / np.max(np.abs(ref)) < 1e-3
    # the block rule zeroes the 3e-7 pivot (<= 1e6 * 1e-12): its solution is bounded
    assert np.max(np.abs(xb)) < 1e-2 * np.max(np.abs(ref))
    # both split that pivot's direction off: A x = b holds outside it
    _, dd = orc.dense_ldlt(A)[:2]
    assert dd[1] == 0.0


def test_tiny_pivot_is_zero_under_both_rules_and_the_scalar_solve_divides_by_it(orc):
    d = np.array([4.0, 1e-13, 2.0, 5.0, 8.0, 7.0])
    A = _ldl_block(d, 5)
    b = np.random.default_rng(6).normal(size=6)
    xs = orc.sparse_ldlt_solve_dense(A, [6], b)
    xb = orc.block_ldlt_solve_dense(A, [6], b)
    assert not np.all(np.isfinite(xs))
    assert np.all(np.isfinite(xb))


def test_lm_on_the_ring_matches_the_block_solver(sg, orc):
    """configs-style ring (10 cameras, 200 points, 20 iterations): the same accept sequence
    and merits to 1e-10 with either RCS solver."""
    p = sg.generate_ba(sg.config_spec(0))
    runs = []
    for solver in (0, 3):
        sol, mem, tr = orc.optimize_nlls(orc.ba_families(p), orc.ba_costs(p),
                                         orc.default_params(num_iterations=20, initial_lm_damping=1.0,
                                                            solver=solver))
        runs.append((sol, mem, tr))
    (s0, m0, t0), (s3, m3, t3) = runs
    assert [t["accepted"] for t in t0] == [t["accepted"] for t in t3]
    for a, b in zip(t0, t3):
        assert abs(a["merit"] - b["merit"]) <= 1e-10 * a["merit"]
    assert np.max(np.abs(m0["poses"] - m3["poses"])) < 1e-9
