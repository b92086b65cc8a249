"""Inverse blocks of the final Hessian (SchurFactor::inverse_block, schur_ldlt.rs:814-958,
through DampedHessian::inverse_block_with_damping, damped_hessian.rs:98) from the GPU's
resident factorization, against dense numpy on the oracle's assembled Hessian.

- Damped (skip_final_hessian, nu > 0): H + nu N is positive definite, so every block of
  every kind must equal the same block of the dense inverse numpy.linalg.inv(H).
- Undamped (nu = 0): with camera 0 fixed the scale of the scene is a gauge direction, S is
  singular, and the reference's formulas use S's Moore-Penrose inverse; numpy forms S and
  S^+ from its eigendecomposition (the one eigenvalue at rounding level dropped).
Bar: 1e-8 relative to the block's largest entry (FP64, condition numbers about 1e6).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _setup(sg, orc, skip_final_hessian):
    p = sg.generate_ba(sg.config_spec(0))
    params = sg.OptParams(num_iterations=6, initial_lm_damping=1.0, skip_final_hessian=skip_final_hessian)
    opt = sg.Optimizer(p.variables(), p.cost_fns(1.0), params)
    while opt.step().termination is None:
        pass
    poses, points = opt.members("poses"), opt.members("points")
    nu = opt.lm_damping() if skip_final_hessian else 0.0
    H, g = orc.linearize_dense(orc.ba_families(p, poses, points), orc.ba_costs(p), nu)
    return p, opt, H


def _queries():
    # (row family, member, col family, member): free/free (same and different cameras),
    # free/marg, marg/free, marg/marg (same and different points)
    return [("poses", 3, "poses", 3), ("poses", 2, "poses", 7), ("poses", 5, "points", 17),
            ("points", 40, "poses", 1), ("points", 9, "points", 9), ("points", 9, "points", 131)]


def _index(name, member, nfree):
    # oracle layout: free poses (members 1..9; member 0 is constant) then points
    return (6 * (member - 1), 6) if name == "poses" else (nfree + 3 * member, 3)


def test_inverse_blocks_damped_match_dense_inverse(sg, orc):
    p, opt, H = _setup(sg, orc, skip_final_hessian=True)
    nfree = 6 * (p.poses_init.shape[0] - 1)
    Hinv = np.linalg.inv(H)
    for rf, rm, cf, cm in _queries():
        blk = opt.inverse_block(rf, rm, cf, cm)
        (r0, dr), (c0, dc) = _index(rf, rm, nfree), _index(cf, cm, nfree)
        ref = Hinv[r0:r0 + dr, c0:c0 + dc]
        rel = np.max(np.abs(blk - ref)) / np.max(np.abs(ref))
        print(f"{rf}[{rm}] x {cf}[{cm}]: rel {rel:.3g}")
        assert rel < 1e-8, (rf, rm, cf, cm, rel)
    opt.close()


def test_inverse_blocks_undamped_use_min_norm_schur_complement(sg, orc):
    p, opt, H = _setup(sg, orc, skip_final_hessian=False)
    nfree = 6 * (p.poses_init.shape[0] - 1)
    Hff, Hfm, Hmm = H[:nfree, :nfree], H[:nfree, nfree:], H[nfree:, nfree:]
    npt = Hmm.shape[0] // 3
    Hmm_inv = np.zeros_like(Hmm)
    for b in range(npt):
        Hmm_inv[3 * b:3 * b + 3, 3 * b:3 * b + 3] = np.linalg.inv(Hmm[3 * b:3 * b + 3, 3 * b:3 * b + 3])
    S = Hff - Hfm @ Hmm_inv @ Hfm.T
    S = 0.5 * (S + S.T)
    w, V = np.linalg.eigh(S)
    keep = w > 1e-9 * w.max()
    assert np.sum(~keep) == 1 and w[1] > 1e3 * abs(w[0])  # one gauge direction, well separated
    Spinv = (V[:, keep] / w[keep]) @ V[:, keep].T
    for rf, rm, cf, cm in _queries():
        blk = opt.inverse_block(rf, rm, cf, cm)
        if rf == "poses" and cf == "poses":
            ref = Spinv[6 * (rm - 1):6 * rm, 6 * (cm - 1):6 * cm]
        elif rf == "poses":
            hb = Hfm[:, 3 * cm:3 * cm + 3]
            ref = -(Spinv @ hb @ Hmm_inv[3 * cm:3 * cm + 3, 3 * cm:3 * cm + 3])[6 * (rm - 1):6 * rm]
        elif cf == "poses":
            hb = Hfm[:, 3 * rm:3 * rm + 3]
            ref = (-(Spinv @ hb @ Hmm_inv[3 * rm:3 * rm + 3, 3 * rm:3 * rm + 3])[6 * (cm - 1):6 * cm]).T
        else:
            hi, hj = Hmm_inv[3 * rm:3 * rm + 3, 3 * rm:3 * rm + 3], Hmm_inv[3 * cm:3 * cm + 3, 3 * cm:3 * cm + 3]
            ref = hi @ Hfm[:, 3 * rm:3 * rm + 3].T @ Spinv @ Hfm[:, 3 * cm:3 * cm + 3] @ hj
            if rm == cm:
                ref = ref + hi
        rel = np.max(np.abs(blk - ref)) / np.max(np.abs(ref))
        print(f"{rf}[{rm}] x {cf}[{cm}]: rel {rel:.3g}")
        assert rel < 1e-8, (rf, rm, cf, cm, rel)
    opt.close()


def test_inverse_block_rejects_constant_and_conditioned_members(sg):
    p = sg.generate_ba(sg.config_spec(0))
    opt = sg.Optimizer(p.variables(), p.cost_fns(1.0), sg.OptParams(num_iterations=2, initial_lm_damping=1.0))
    opt.step()
    with pytest.raises(sg.NllsError):
        opt.inverse_block("poses", 0, "poses", 1)  # camera 0 is constant
    with pytest.raises(sg.NllsError):
        opt.inverse_block("cams", 0, "poses", 1)  # intrinsics are conditioned
    opt.close()


def test_covariance_block_full_rank_matches_dense_inverse(sg, orc):
    """Cameras 0 and 5 fixed: the scale is no longer a gauge, the undamped final Hessian has
    full rank and covariance_block (optimizer_types.rs:335) is its inverse's block."""
    p = sg.generate_ba(sg.config_spec(0))
    const = [0, 5]

    def families(poses, points):
        return (sg.VarBuilder()
                .add_family("cams", sg.VarFamily(sg.VarKind.Conditioned, sg.VarType.Pinhole, p.intrinsics))
                .add_family("poses", sg.VarFamily(sg.VarKind.Free, sg.VarType.Isometry3, poses, const))
                .add_family("points", sg.VarFamily(sg.VarKind.Marginalized, sg.VarType.Vec3, points))
                .build())

    opt = sg.Optimizer(families(p.poses_init, p.points_init), p.cost_fns(1.0),
                       sg.OptParams(num_iterations=6, initial_lm_damping=1.0))
    while opt.step().termination is None:
        pass
    poses, points = opt.members("poses"), opt.members("points")
    ofams = [("cams", 2, 1, p.intrinsics, []), ("poses", 0, 0, poses, const), ("points", 1, 2, points, [])]
    H, _ = orc.linearize_dense(ofams, orc.ba_costs(p), 0.0)
    Hinv = np.linalg.inv(H)
    free = [m for m in range(p.poses_init.shape[0]) if m not in const]
    nfree = 6 * len(free)

    def index(name, member):
        return (6 * free.index(member), 6) if name == "poses" else (nfree + 3 * member, 3)

    for rf, rm, cf, cm in [("poses", 3, "poses", 3), ("poses", 2, "points", 11), ("points", 7, "points", 50)]:
        blk = opt.covariance_block(rf, rm, cf, cm)
        (r0, dr), (c0, dc) = index(rf, rm), index(cf, cm)
        ref = Hinv[r0:r0 + dr, c0:c0 + dc]
        rel = np.max(np.abs(blk - ref)) / np.max(np.abs(ref))
        print(f"{rf}[{rm}] x {cf}[{cm}]: rel {rel:.3g}")
        assert rel < 1e-8
    opt.close()


def test_covariance_block_rejects_rank_deficient_hessian(sg):
    """Camera 0 fixed only: the scale gauge leaves H rank-deficient, where the reference's
    min-norm inverse of the whole of H is not the Schur-complement block."""
    p = sg.generate_ba(sg.config_spec(0))
    opt = sg.Optimizer(p.variables(), p.cost_fns(1.0), sg.OptParams(num_iterations=3, initial_lm_damping=1.0))
    while opt.step().termination is None:
        pass
    with pytest.raises(sg.Unsupported):
        opt.covariance_block("poses", 3, "poses", 3)
    opt.close()
