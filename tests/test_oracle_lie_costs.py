"""Pins the oracle's Lie group maps, cost residuals/Jacobians and Huber kernel against
closed forms that do not share its code: matrix exponentials/logarithms (scipy),
central finite differences, and the oracle's own forward-mode dual numbers.

Reference: crates/sophus_lie/src/groups/rotation3.rs (exp/log/mat_v/inv_left_jacobian),
crates/sophus_lie/src/groups/affine_group_template.rs:350-410,596-618 (SE(3)),
crates/sophus_opt/src/nlls/functor_library/costs/pinhole_reprojection.rs:42-95,
crates/sophus_opt/src/nlls/functor_library/costs/pose_graph3.rs:33-79,
crates/sophus_opt/src/robust_kernel.rs:13-35.
"""
import json
from pathlib import Path

import numpy as np
import pytest
from scipy.linalg import expm

FIX = json.loads((Path(__file__).parent / "golden" / "reference_fixtures.json").read_text())
TANGENTS = np.array(FIX["rotation3_tangent_examples"]["tangents"])


def hat(w):
    return np.array([[0.0, -w[2], w[1]], [w[2], 0.0, -w[0]], [-w[1], w[0], 0.0]])


def twist(xi):
    """4x4 matrix of the SE(3) tangent [omega; upsilon] (rotation part first)."""
    M = np.zeros((4, 4))
    M[:3, :3] = hat(xi[:3])
    M[:3, 3] = xi[3:]
    return M


def se3_ad(xi):
    A = np.zeros((6, 6))
    A[:3, :3] = hat(xi[:3])
    A[3:, :3] = hat(xi[3:])
    A[3:, 3:] = hat(xi[:3])
    return A


def series_jacobian(A):
    """sum_k A^k / (k+1)!  (the left Jacobian of the group whose ad is A)."""
    n = A.shape[0]
    big = np.zeros((2 * n, 2 * n))
    big[:n, :n] = A
    big[:n, n:] = np.eye(n)
    return expm(big)[:n, n:]


def ref_mat_v(w):
    """mat_v with the reference's branch: the first-order I + w^/2 for |w|^2 <= 1e-6
    (rotation3.rs:761-776), else the exact series."""
    if w @ w <= 1e-6:
        return np.eye(3) + 0.5 * hat(w)
    return series_jacobian(hat(w))


def close(a, b, atol):
    return np.allclose(a, b, rtol=0.0, atol=atol)


def pose_matrix(orc, T):
    M = np.eye(4)
    M[:3, :3] = orc.so3_matrix(T[:4])
    M[:3, 3] = T[4:]
    return M


def se3_tangents(rng, k=12):
    out = [np.concatenate([w, rng.normal(0, 1.0, 3)]) for w in TANGENTS]
    out += [rng.normal(0, 0.8, 6) for _ in range(k)]
    return out


def test_so3_exp_matches_matrix_exponential(orc):
    for w in TANGENTS:
        q = orc.so3_exp(w)
        assert abs(np.linalg.norm(q) - 1.0) < 1e-14
        assert close(orc.so3_matrix(q), expm(hat(w)), 1e-14)


def test_so3_log_inverts_exp_on_reference_examples(orc):
    for w in TANGENTS:
        assert close(orc.so3_log(orc.so3_exp(w)), w, 1e-12)


def test_se3_exp_log_match_matrix_exponential(orc):
    rng = np.random.default_rng(0)
    for xi in se3_tangents(rng):
        T = orc.se3_exp(xi)
        M = expm(twist(xi))
        assert close(pose_matrix(orc, T)[:3, :3], M[:3, :3], 1e-13)
        assert close(T[4:], ref_mat_v(xi[:3]) @ xi[3:], 1e-13)
        # log(exp(x)) = x, up to the first-order V of the series branch
        th2 = xi[:3] @ xi[:3]
        slack = th2 * np.linalg.norm(xi[3:]) if th2 <= 1e-6 else 0.0
        assert close(orc.se3_log(T), xi, 1e-10 + slack)


def test_se3_mul_and_inverse_are_matrix_products(orc):
    rng = np.random.default_rng(1)
    for _ in range(10):
        a, b = orc.se3_exp(rng.normal(0, 1, 6)), orc.se3_exp(rng.normal(0, 1, 6))
        assert close(pose_matrix(orc, orc.se3_mul(a, b)), pose_matrix(orc, a) @ pose_matrix(orc, b), 1e-13)
        assert close(pose_matrix(orc, orc.se3_inverse(a)), np.linalg.inv(pose_matrix(orc, a)), 1e-13)


def test_so3_mat_v_is_left_jacobian_series(orc):
    rng = np.random.default_rng(2)
    ws = list(TANGENTS) + [rng.normal(0, 1e-4, 3), rng.normal(0, 1e-2, 3)]
    for w in ws:
        V = series_jacobian(hat(w))
        assert close(orc.so3_mat_v(w), ref_mat_v(w), 1e-14)
        # the series branch truncates V after the first order
        assert close(orc.so3_mat_v(w), V, 1e-12 + (w @ w) / 6.0)
        # mat_v_inverse's series branch keeps the second order (rotation3.rs:785-786)
        assert close(orc.so3_mat_v_inverse(w), np.linalg.inv(V), 1e-12)


def test_se3_inv_left_jacobian_large_angles(orc):
    """Away from the series branch (|omega|^2 > 1e-6) it is the inverse of the SE(3)
    left Jacobian sum_k ad^k/(k+1)!."""
    rng = np.random.default_rng(3)
    for xi in se3_tangents(rng):
        if xi[:3] @ xi[:3] <= 1e-6:
            continue
        J = series_jacobian(se3_ad(xi))
        assert close(orc.se3_inv_left_jacobian(xi), np.linalg.inv(J), 1e-9)


def test_se3_inv_left_jacobian_series_branch_keeps_reference_form(orc):
    """For |omega|^2 <= 1e-6 the reference uses I + omega^/2 for the SO(3) block
    (rotation3.rs:646) and the truncated Q = upsilon^/2 + omega^ upsilon^/6
    (rotation3.rs:993-995), with V^-1 from the exact series (affine_group_template.rs:601-618)."""
    rng = np.random.default_rng(4)
    for _ in range(6):
        w = rng.normal(0, 2e-4, 3)
        u = rng.normal(0, 1.0, 3)
        xi = np.concatenate([w, u])
        jl_inv = np.eye(3) + 0.5 * hat(w)
        v_inv = np.linalg.inv(series_jacobian(hat(w)))
        q = 0.5 * hat(u) + hat(w) @ hat(u) / 6.0
        expect = np.zeros((6, 6))
        expect[:3, :3] = jl_inv
        expect[3:, 3:] = v_inv
        expect[3:, :3] = -(v_inv @ q @ jl_inv)
        assert close(orc.se3_inv_left_jacobian(xi), expect, 1e-13)


@pytest.mark.parametrize("delta", [0.5, 1.0, 3.0])
def test_huber_weight_closed_form(orc, delta):
    for n in (0.0, 1e-9, 0.25, delta, delta * 1.0000001, 2.0, 50.0):
        if n == 0.0:
            expect = 0.0
        elif n <= delta:
            expect = 1.0
        else:
            expect = np.sqrt(2 * delta * n - delta * delta) / n
        assert abs(orc.huber_weight(delta, n) - expect) <= 1e-15


def _pinhole_inputs(rng, n, orc):
    intr = np.tile([500.0, 505.0, 320.0, 240.0], (n, 1))
    poses = np.stack([orc.se3_exp(rng.normal(0, [0.3, 0.3, 0.3, 1, 1, 1])) for _ in range(n)])
    pts = []
    for i in range(n):
        xc = np.array([rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(2, 8)])
        pts.append(orc.so3_matrix(poses[i][:4]) @ xc + poses[i][4:])
    uv = rng.uniform([0, 0], [640, 480], (n, 2))
    return intr, poses, np.array(pts), uv


def _project(intr, T, x, orc):
    """z - K pi(T^-1 x), written with 4x4 matrices."""
    M = np.linalg.inv(pose_matrix(orc, T))
    xc = M[:3, :3] @ x + M[:3, 3]
    return np.array([intr[0] * xc[0] / xc[2] + intr[2], intr[1] * xc[1] / xc[2] + intr[3]])


def test_pinhole_residual_is_projection(orc):
    rng = np.random.default_rng(5)
    intr, poses, pts, uv = _pinhole_inputs(rng, 32, orc)
    o = orc.eval_pinhole_terms(intr, poses, pts, uv)
    for i in range(32):
        assert close(o["residual"][i], uv[i] - _project(intr[i], poses[i], pts[i], orc), 1e-10)


def test_pinhole_jacobians_match_finite_differences_and_duals(orc):
    """J_pose is d r(exp(d) T) / d d at 0 (left update, [omega; upsilon]); J_point is
    d r / d x.  Central differences at h = 1e-6 and forward-mode duals."""
    rng = np.random.default_rng(6)
    n = 24
    intr, poses, pts, uv = _pinhole_inputs(rng, n, orc)
    o = orc.eval_pinhole_terms(intr, poses, pts, uv)
    jp_dual, jx_dual = orc.pinhole_dual_jacobians(intr, poses, pts, uv)
    h = 1e-6
    for i in range(n):
        r = lambda T, x: uv[i] - _project(intr[i], T, x, orc)
        Jp = np.zeros((2, 6))
        Jx = np.zeros((2, 3))
        for k in range(6):
            d = np.zeros(6)
            d[k] = h
            Jp[:, k] = (r(orc.se3_mul(orc.se3_exp(d), poses[i]), pts[i])
                        - r(orc.se3_mul(orc.se3_exp(-d), poses[i]), pts[i])) / (2 * h)
        for k in range(3):
            d = np.zeros(3)
            d[k] = h
            Jx[:, k] = (r(poses[i], pts[i] + d) - r(poses[i], pts[i] - d)) / (2 * h)
        scale = max(1.0, np.max(np.abs(o["jac_pose"][i])))
        assert close(o["jac_pose"][i], Jp, 2e-6 * scale)
        assert close(o["jac_point"][i], Jx, 2e-6 * scale)
        assert close(o["jac_pose"][i], jp_dual[i], 1e-11 * scale)
        assert close(o["jac_point"][i], jx_dual[i], 1e-11 * scale)


def test_pinhole_huber_weighting(orc):
    """cost = |w r|^2 = rho(|r|) with w = sqrt(rho(|r|))/|r|; the entry point returns the
    raw residual and w separately, and the Jacobians are not weighted
    (evaluated_term.rs:286-327)."""
    rng = np.random.default_rng(7)
    intr, poses, pts, uv = _pinhole_inputs(rng, 64, orc)
    plain = orc.eval_pinhole_terms(intr, poses, pts, uv)
    for delta in (1.0, 20.0):
        o = orc.eval_pinhole_terms(intr, poses, pts, uv, delta)
        nrm = np.linalg.norm(plain["residual"], axis=1)
        rho = np.where(nrm <= delta, nrm ** 2, 2 * delta * nrm - delta ** 2)
        assert np.allclose(o["cost"], rho, rtol=1e-12, atol=0.0)
        assert np.array_equal(o["residual"], plain["residual"])
        wr = plain["residual"] * o["weight"][:, None]
        assert np.allclose(np.sum(wr * wr, axis=1), rho, rtol=1e-12, atol=0.0)
        assert np.array_equal(o["jac_pose"], plain["jac_pose"])


def _pose_graph_inputs(orc, rng, n, scale):
    m = np.stack([orc.se3_exp(rng.normal(0, 0.5, 6)) for _ in range(n)])
    nn = np.stack([orc.se3_exp(rng.normal(0, 0.5, 6)) for _ in range(n)])
    mn = np.stack([orc.se3_mul(orc.se3_mul(orc.se3_inverse(m[i]), nn[i]), orc.se3_exp(rng.normal(0, scale, 6)))
                   for i in range(n)])
    return m, nn, mn


def test_pose_graph_residual_is_log_of_error(orc):
    rng = np.random.default_rng(8)
    m, nn, mn = _pose_graph_inputs(orc, rng, 16, 0.3)
    r, _, _ = orc.eval_pose_graph3_terms(m, nn, mn)
    for i in range(16):
        E = np.linalg.inv(pose_matrix(orc, m[i])) @ pose_matrix(orc, nn[i]) @ np.linalg.inv(pose_matrix(orc, mn[i]))
        assert close(expm(twist(r[i])), E, 1e-10)


def test_pose_graph_jacobians_match_duals_for_large_residuals(orc):
    """pose_graph3.rs:33-79: J_m = -J_l^-1(r) Adj(T_m^-1), J_n = +J_l^-1(r) Adj(T_m^-1);
    away from the series branch these equal the true derivative (duals and central
    differences of r(exp(d) T))."""
    rng = np.random.default_rng(9)
    m, nn, mn = _pose_graph_inputs(orc, rng, 16, 0.3)
    r, jm, jn = orc.eval_pose_graph3_terms(m, nn, mn)
    dm, dn = orc.pose_graph3_dual_jacobians(m, nn, mn)
    assert close(jm, dm, 1e-9)
    assert close(jn, dn, 1e-9)
    h = 1e-6
    for i in range(4):
        res = lambda a, b: orc.eval_pose_graph3_terms(a[None], b[None], mn[i][None])[0][0]
        for k in range(6):
            d = np.zeros(6)
            d[k] = h
            col = (res(orc.se3_mul(orc.se3_exp(d), m[i]), nn[i])
                   - res(orc.se3_mul(orc.se3_exp(-d), m[i]), nn[i])) / (2 * h)
            assert close(jm[i][:, k], col, 1e-7)
