"""The symbolic analysis on the device against the host analysis, element by element.

The device pipeline (stable LSD radix sort of the terms by (point, free argument,
conditioned argument), the S pattern as a bitmap, the per-slot pair lists by count, scan,
fill and a stable radix sort by slot) must give every array the numeric phase reads exactly
as the host analysis does: the sorted terms, the pattern, the pairs and their order within
each slot, the camera-major lists, the ordering and the supernodal structure.  The terms
are handed over shuffled so the sort does real work; the cases cover duplicated terms,
constant poses and points, a camera without observations, pose-graph terms, free and
conditioned EUCM intrinsics and the benchmark's 2,000-camera walk.  An optimizer built
either way follows a bitwise identical trajectory.
"""
import dataclasses

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _shuffled(p, seed):
    rng = np.random.default_rng(seed)
    k = rng.permutation(len(p.obs_idx))
    return dataclasses.replace(p, obs_idx=np.ascontiguousarray(p.obs_idx[k]),
                               obs_uv=np.ascontiguousarray(p.obs_uv[k]))


def _check(sg, v, costs, label):
    r = sg.symbolic_check(v, costs)
    print(f"{label}: {r}", flush=True)
    assert r["mismatches"] == 0, r
    return r


def test_ring_with_duplicates_matches_host(sg):
    p = sg.generate_ba(sg.config_spec(0))
    rng = np.random.default_rng(3)
    dup = rng.choice(len(p.obs_idx), 40, replace=False)
    q = dataclasses.replace(p, obs_idx=np.concatenate([p.obs_idx, p.obs_idx[dup]]),
                            obs_uv=np.concatenate([p.obs_uv, p.obs_uv[dup] + 1.0]))
    q = _shuffled(q, 1)
    _check(sg, q.variables(), q.cost_fns(1.0), "ring + duplicates")


def test_walk_500cams_shuffled_matches_host(sg):
    p = _shuffled(sg.generate_ba(sg.config_spec(1)), 2)
    r = _check(sg, p.variables(), p.cost_fns(1.0), "walk 500")
    assert r["schur_pairs"] > 1_000_000


def test_constant_members_unobserved_camera_and_pose_graph_match_host(sg, orc):
    spec = sg.BaSpec(layout=1, num_cameras=40, num_points=1500, geometric_p=0.2, seed=5)
    p = sg.generate_ba(spec)
    keep = p.obs_idx[:, 1] != 17  # camera 17 is never observed
    p = _shuffled(dataclasses.replace(p, obs_idx=p.obs_idx[keep], obs_uv=p.obs_uv[keep]), 4)
    v = (sg.VarBuilder()
         .add_family("cams", sg.VarFamily(sg.VarKind.Conditioned, sg.VarType.Pinhole, p.intrinsics))
         .add_family("poses", sg.VarFamily(sg.VarKind.Free, sg.VarType.Isometry3, p.poses_init, [0, 9, 30]))
         .add_family("points", sg.VarFamily(sg.VarKind.Marginalized, sg.VarType.Vec3, p.points_init,
                                            list(range(0, 1500, 7))))
         .build())
    pairs = np.array([[i, i + 1] for i in range(39)] + [[3, 25], [10, 2]], dtype=np.int64)
    meas = np.stack([orc.se3_mul(orc.se3_inverse(p.poses_gt[i]), p.poses_gt[j]) for i, j in pairs])
    costs = p.cost_fns(1.0) + [sg.CostFn.pose_graph3(["poses", "poses"], pairs, meas)]
    _check(sg, v, costs, "const members + pose graph")


@pytest.mark.parametrize("cams_kind", ["Free", "Conditioned"])
def test_eucm_ring_matches_host(sg, cams_kind):
    p = sg.generate_ba_eucm(40, 2000, seed=1)
    p = dataclasses.replace(p, obs_idx=p.obs_idx[np.random.default_rng(6).permutation(len(p.obs_idx))])
    kind = getattr(sg.VarKind, cams_kind)
    _check(sg, p.variables(kind), [sg.CostFn.eucm_reprojection(["cams", "poses", "points"], p.obs_idx,
                                                               np.zeros((len(p.obs_idx), 2)))],
           f"eucm {cams_kind}")


def test_walk_2000cams_matches_host(sg):
    """The benchmark's workload (5.95M observations), in the generator's order."""
    p = sg.generate_ba(sg.config_spec(2))
    r = _check(sg, p.variables(), p.cost_fns(1.0), "walk 2000")
    # a second call: the device pipeline without first-use costs
    r = _check(sg, p.variables(), p.cost_fns(1.0), "walk 2000 (warm)")
    assert r["s_blocks"] > 2000


def test_optimizer_trajectory_is_bitwise_identical_either_way(sg):
    p = _shuffled(sg.generate_ba(sg.BaSpec(layout=1, num_cameras=60, num_points=3000, geometric_p=0.15,
                                           seed=7)), 8)
    runs = []
    for host in (True, False):
        opt = sg.Optimizer(p.variables(), p.cost_fns(1.0),
                           sg.OptParams(num_iterations=6, initial_lm_damping=1.0, host_symbolic=host))
        merits = []
        while True:
            info = opt.step()
            merits.append((info.merit, info.lm_damping, info.accepted))
            if info.termination is not None:
                break
        runs.append((merits, opt.members("poses"), opt.members("points")))
        opt.close()
    assert runs[0][0] == runs[1][0]
    assert np.array_equal(runs[0][1], runs[1][1]) and np.array_equal(runs[0][2], runs[1][2])
