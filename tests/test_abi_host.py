"""The C ABI library without a GPU: it loads, exports every function that
include/sophus_gpu.h declares, carries the reference's OptParams defaults, and its
host-side parts (validation, symbolic analysis, the synthetic problem generator) behave
as specified.  No compute entry point is called here.
"""
import re
from pathlib import Path

import numpy as np
import pytest

HEADER = Path(__file__).resolve().parents[1] / "include" / "sophus_gpu.h"


def _declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sgpu_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(sg):
    lib = sg.library()
    declared = _declared_functions()
    assert len(declared) >= 20
    for name in declared:
        assert hasattr(lib, name), name
    assert sorted(sg.exported_symbols()) == declared


def test_abi_version_and_default_params(sg):
    import ctypes as C

    lib = sg.library()
    assert lib.sgpu_abi_version() >= 1
    p = sg._OptParams()
    lib.sgpu_opt_params_default(C.byref(p))
    # OptParams::default, crates/sophus_opt/src/nlls/optimizer_types.rs:224-238
    assert p.num_iterations == 100 and p.initial_lm_damping == 10.0 and p.parallelize == 1
    assert p.skip_final_hessian == 0 and p.max_consecutive_rejects == 20
    assert (p.function_tolerance, p.gradient_tolerance, p.parameter_tolerance) == (1e-10, 1e-14, 1e-12)
    assert not p.precond_weights  # NULL by default
    py = sg.OptParams()._c()
    for k, _ in sg._OptParams._fields_:
        if k == "precond_weights":
            assert not py.precond_weights
        else:
            assert getattr(py, k) == getattr(p, k), k


def test_generator_is_deterministic_and_seeded(sg):
    a = sg.generate_ba(sg.config_spec(0))
    b = sg.generate_ba(sg.config_spec(0))
    c = sg.generate_ba(sg.config_spec(0, seed=1))
    for k in ("poses_init", "points_init", "obs_idx", "obs_uv"):
        assert np.array_equal(getattr(a, k), getattr(b, k))
    assert not np.array_equal(a.obs_uv, c.obs_uv)


def test_ring_layout_geometry(sg, orc):
    """10 cameras on a radius-2 circle looking at the origin; 5 distinct cameras per
    point; noise-free projections land on the measurements up to the pixel noise."""
    spec = sg.config_spec(0)
    p = sg.generate_ba(spec)
    assert p.poses_gt.shape == (10, 7) and p.points_gt.shape == (200, 3) and p.num_obs == 1000
    centres = p.poses_gt[:, 4:]
    assert np.allclose(np.linalg.norm(centres[:, :2], axis=1), 2.0) and np.allclose(centres[:, 2], 0.0)
    for T in p.poses_gt:
        R = orc.so3_matrix(T[:4])
        axis = R[:, 2]  # optical axis in the world
        assert np.allclose(axis, -T[4:] / np.linalg.norm(T[4:]), atol=1e-12)
    per_point = np.bincount(p.obs_idx[:, 2], minlength=200)
    assert np.all(per_point == 5)
    for b in range(200):
        cams = p.obs_idx[p.obs_idx[:, 2] == b, 1]
        assert len(set(cams.tolist())) == 5
    assert np.array_equal(p.poses_init[0], p.poses_gt[0])  # camera 0 is the fixed gauge
    o = orc.eval_pinhole_terms(np.repeat(p.intrinsics, p.num_obs, 0), p.poses_gt[p.obs_idx[:, 1]],
                               p.points_gt[p.obs_idx[:, 2]], p.obs_uv)
    assert np.std(o["residual"]) < 0.6 and np.std(o["residual"]) > 0.4  # N(0, 0.5^2) pixels


def test_walk_layout_statistics(sg, orc):
    spec = sg.BaSpec(layout=1, num_cameras=80, num_points=4000, geometric_p=0.1, seed=3)
    p = sg.generate_ba(spec)
    per_point = np.bincount(p.obs_idx[:, 2], minlength=4000)
    assert per_point.min() >= 2 and per_point.max() <= 30
    steps = np.linalg.norm(np.diff(p.poses_gt[:, 4:], axis=0), axis=1)
    assert np.allclose(steps, 0.3)
    # every observation projects inside the 640 x 480 image at the ground truth
    o = orc.eval_pinhole_terms(np.repeat(p.intrinsics, p.num_obs, 0), p.poses_gt[p.obs_idx[:, 1]],
                               p.points_gt[p.obs_idx[:, 2]], np.zeros((p.num_obs, 2)))
    uv = -o["residual"]
    assert np.all((uv[:, 0] >= 0) & (uv[:, 0] <= 640) & (uv[:, 1] >= 0) & (uv[:, 1] <= 480))


def test_symbolic_analysis_counts(sg):
    """The reduced camera system's pattern: one 6x6 block per pair of free cameras that
    share a point (lower triangle, diagonal included); Schur pairs per point k(k+1)/2."""
    spec = sg.BaSpec(layout=1, num_cameras=40, num_points=1500, geometric_p=0.2, seed=5)
    p = sg.generate_ba(spec)
    st = sg.analyze(p.variables(), p.cost_fns(1.0))
    assert st["num_free_blocks"] == 39 and st["num_marg_blocks"] == 1500
    assert st["num_observations"] == p.num_obs
    pairs = set()
    n_pairs = 0
    for b in range(1500):
        cams = sorted(c for c in p.obs_idx[p.obs_idx[:, 2] == b, 1].tolist() if c != 0)
        n_pairs += len(cams) * (len(cams) + 1) // 2
        for i in cams:
            for j in cams:
                if j <= i:
                    pairs.add((i, j))
    pairs |= {(c, c) for c in range(1, 40)}
    assert st["nnz_s_blocks"] == len(pairs)
    assert st["num_schur_pairs"] == n_pairs
    # L's strictly-lower blocks hold S's off-diagonal pattern plus the fill-in
    assert st["nnz_l_blocks"] >= st["nnz_s_blocks"] - st["num_free_blocks"]
    assert st["factor_flops"] > 0 and st["num_supernodes"] >= 1 and st["num_levels"] >= 1


def test_window_generator_and_batch_validation(sg):
    """The batched workload's windows: sizes in range, 2..8 distinct cameras per point;
    windows beyond 31 cameras are rejected before any device work."""
    for i in range(20):
        s = sg.window_spec(i)
        assert 5 <= s.num_cameras <= 30 and 100 <= s.num_points <= 2000
        w = sg.generate_ba(s)
        per = np.bincount(w.obs_idx[:, 2], minlength=s.num_points)
        assert per.min() >= 1 and per.max() <= 8
    big = sg.generate_ba(sg.BaSpec(layout=0, num_cameras=40, num_points=50, obs_per_point_ring=3))
    with pytest.raises(sg.Unsupported):
        sg.Batch([big], sg.OptParams())
