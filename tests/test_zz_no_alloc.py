"""The allocation-free block LDLt solve (sgpu_block_ldlt_solve_noalloc) against the
allocating one (sgpu_block_ldlt_solve), through the C ABI: two calls in a row, the second
on a smaller system, so the second runs on scratch that still holds the first call's
values; and the size limit, which is checked on the host before any device work."""
import numpy as np
import pytest


def _banded(nb, band, extra, seed):
    rng = np.random.default_rng(seed)
    n = 6 * nb
    A = np.zeros((n, n))
    for i in range(nb):
        for j in range(max(0, i - band), i + 1):
            B = rng.normal(0, 1, (6, 6))
            A[6 * i:6 * i + 6, 6 * j:6 * j + 6] += B
            if i != j:
                A[6 * j:6 * j + 6, 6 * i:6 * i + 6] += B.T
    for _ in range(extra):
        i, j = sorted(rng.integers(0, nb, 2))
        if i != j:
            B = rng.normal(0, 1, (6, 6))
            A[6 * j:6 * j + 6, 6 * i:6 * i + 6] += B
            A[6 * i:6 * i + 6, 6 * j:6 * j + 6] += B.T
    A = 0.5 * (A + A.T)
    A += np.eye(n) * (np.abs(A).sum(axis=1).max() + 1.0)
    return A


@pytest.mark.gpu
def test_noalloc_block_ldlt_matches_allocating_solve(sg):
    # the larger system first, then a smaller one on the same (now dirty) scratch, then the
    # larger one again
    cases = [(300, 12, 50, 1), (60, 5, 20, 2), (300, 12, 50, 3)]
    for nb, band, extra, seed in cases:
        A = _banded(nb, band, extra, seed)
        b = np.random.default_rng(seed + 10).normal(size=6 * nb)
        x_ref = sg.block_ldlt_solve(A, b)
        x = sg.block_ldlt_solve(A, b, noalloc=True)
        rel = float(np.max(np.abs(x - x_ref)) / np.max(np.abs(x_ref)))
        print(f"nb {nb}: noalloc vs allocating rel {rel:.3g}, bitwise equal {np.array_equal(x, x_ref)}")
        assert rel <= 1e-12, (nb, rel)
        assert float(np.max(np.abs(x - np.linalg.solve(A, b))) / np.max(np.abs(x_ref))) < 1e-10


def test_noalloc_rejects_systems_beyond_its_scratch(sg):
    # a fully coupled 520-block system: its one front alone needs 3120^2 doubles (78 MB)
    nb = 520
    col_ptr = np.zeros(nb + 1, dtype=np.int64)
    rows = []
    for c in range(nb):
        rows.append(np.arange(c, nb, dtype=np.int64))
        col_ptr[c + 1] = col_ptr[c] + (nb - c)
    rows = np.concatenate(rows)
    vals = np.zeros((rows.size, 6, 6))
    rhs = np.zeros(6 * nb)
    x = np.empty(6 * nb)
    lib = sg.library()
    DP, I64P = sg._DP, sg._I64P
    st = lib.sgpu_block_ldlt_solve_noalloc(nb, col_ptr.ctypes.data_as(I64P), rows.ctypes.data_as(I64P),
                                           vals.ctypes.data_as(DP), rhs.ctypes.data_as(DP),
                                           x.ctypes.data_as(DP))
    assert st == 10  # SGPU_ERR_OUT_OF_MEMORY
    msg = lib.sgpu_last_error().decode()
    assert "bytes of device scratch" in msg and str(sg.NOALLOC_SCRATCH_BYTES) in msg, msg
