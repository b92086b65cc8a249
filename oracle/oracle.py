"""ORACLE — test infrastructure only.

ctypes front end of liboracle.so, the CPU FP64 restatement of sophus-rs's Schur LM path
(oracle/ba_oracle.cpp).  Only tests/, the smoke check and the benchmark's CPU-baseline
leg import this module; the product library never does.  The ABI structs are declared
here independently of the product's Python package, so a marshalling mistake on one
side cannot hide behind the same mistake on the other.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
# variant -> library: "" is the oracle; "fma" its rounding twin (same source, FMA contraction
# on), used to measure how far the oracle itself moves under a rounding-only change
# SOPHUS_ORACLE_LIB replaces the oracle's library, e.g. by its sanitizer build (Makefile)
VARIANTS = {"": Path(os.environ.get("SOPHUS_ORACLE_LIB", LIB)), "fma": HERE / "liboracle_fma.so"}


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


class VarFamilyC(C.Structure):
    _fields_ = [("name", C.c_char_p), ("type", C.c_int32), ("kind", C.c_int32),
                ("num_members", C.c_int64), ("members", C.POINTER(C.c_double)),
                ("num_const_ids", C.c_int64), ("const_ids", C.POINTER(C.c_int64))]


class CostFnC(C.Structure):
    _fields_ = [("kind", C.c_int32), ("robust_kernel", C.c_int32), ("huber_delta", C.c_double),
                ("family_names", C.c_char_p * 3), ("num_terms", C.c_int64),
                ("entity_indices", C.POINTER(C.c_int64)), ("measurements", C.POINTER(C.c_double))]


class OptParamsC(C.Structure):
    _fields_ = [("num_iterations", C.c_int64), ("initial_lm_damping", C.c_double),
                ("parallelize", C.c_int32), ("solver", C.c_int32), ("skip_final_hessian", C.c_int32),
                ("precision", C.c_int32), ("function_tolerance", C.c_double),
                ("gradient_tolerance", C.c_double), ("parameter_tolerance", C.c_double),
                ("max_consecutive_rejects", C.c_int64), ("pcg_relative_tolerance", C.c_double),
                ("pcg_max_iterations", C.c_int64), ("device", C.c_int32), ("reserved0", C.c_int32),
                ("precond_weights", C.c_void_p)]


class EqFnC(C.Structure):
    _fields_ = [("kind", C.c_int32), ("reserved0", C.c_int32), ("family_names", C.c_char_p * 1),
                ("num_terms", C.c_int64), ("entity_indices", C.POINTER(C.c_int64)),
                ("params", C.POINTER(C.c_double))]


class SolutionC(C.Structure):
    _fields_ = [("initial_cost", C.c_double), ("final_cost", C.c_double), ("num_steps", C.c_int64),
                ("termination", C.c_int32), ("reserved0", C.c_int32),
                ("final_max_abs_neg_grad", C.c_double)]


class StepRecordC(C.Structure):
    _fields_ = [("merit", C.c_double), ("lm_damping", C.c_double), ("new_merit", C.c_double),
                ("gain_ratio", C.c_double), ("max_abs_delta", C.c_double),
                ("accepted", C.c_int32), ("termination", C.c_int32)]


DP = C.POINTER(C.c_double)
I64P = C.POINTER(C.c_int64)
_libs: Dict[str, C.CDLL] = {}


def lib(variant: str = "") -> C.CDLL:
    if variant not in _libs:
        path = VARIANTS[variant]
        if not path.exists():
            build()
        L = C.CDLL(str(path))
        L.oracle_last_error.restype = C.c_char_p
        L.oracle_last_step_seconds.restype = C.c_double
        L.oracle_optimize_nlls_eq.restype = C.c_int
        L.oracle_optimize_nlls_eq.argtypes = [C.POINTER(VarFamilyC), C.c_int64, C.POINTER(CostFnC), C.c_int64,
                                              C.POINTER(EqFnC), C.c_int64, C.POINTER(OptParamsC), C.POINTER(DP),
                                              C.POINTER(SolutionC), C.POINTER(StepRecordC), C.c_int64, I64P, DP]
        L.oracle_kkt_step.restype = C.c_int32
        L.oracle_kkt_step.argtypes = [C.POINTER(VarFamilyC), C.c_int64, C.POINTER(CostFnC), C.c_int64,
                                      C.POINTER(EqFnC), C.c_int64, C.c_double, DP, DP, DP, DP, DP]
        L.oracle_optimize_nlls.restype = C.c_int
        L.oracle_optimize_nlls.argtypes = [C.POINTER(VarFamilyC), C.c_int64, C.POINTER(CostFnC), C.c_int64,
                                           C.POINTER(OptParamsC), C.POINTER(DP), C.POINTER(SolutionC),
                                           C.POINTER(StepRecordC), C.c_int64, I64P]
        L.oracle_linearize_dense.restype = C.c_int
        L.oracle_linearize_dense.argtypes = [C.POINTER(VarFamilyC), C.c_int64, C.POINTER(CostFnC), C.c_int64,
                                             C.c_double, DP, DP, I64P]
        L.oracle_schur_step.restype = C.c_int
        L.oracle_schur_step.argtypes = [C.POINTER(VarFamilyC), C.c_int64, C.POINTER(CostFnC), C.c_int64,
                                        C.c_double, DP, C.POINTER(C.c_int32)]
        L.oracle_merit.restype = C.c_double
        L.oracle_merit.argtypes = [C.POINTER(VarFamilyC), C.c_int64, C.POINTER(CostFnC), C.c_int64]
        L.oracle_schur_solve_dense.restype = None
        L.oracle_schur_solve_dense.argtypes = [C.c_int64, I64P, C.c_int64, DP, DP, DP, C.POINTER(C.c_int32)]
        L.oracle_block_ldlt_solve_dense.restype = C.c_int32
        L.oracle_block_ldlt_solve_dense.argtypes = [C.c_int64, I64P, DP, DP, DP]
        L.oracle_sparse_ldlt_solve_dense.restype = C.c_int32
        L.oracle_sparse_ldlt_solve_dense.argtypes = [C.c_int64, I64P, DP, DP, DP]
        L.oracle_dense_ldlt.restype = C.c_int32
        L.oracle_dense_ldlt.argtypes = [C.c_int32, DP, C.c_double, DP, DP, C.POINTER(C.c_int32),
                                        C.POINTER(C.c_int32), DP]
        L.oracle_dense_ldlt_solve.restype = None
        L.oracle_dense_ldlt_solve.argtypes = [C.c_int32, DP, DP, DP]
        L.oracle_eval_pinhole_terms.restype = None
        L.oracle_eval_pinhole_terms.argtypes = [C.c_int64, DP, DP, DP, DP, C.c_int32, C.c_double,
                                                DP, DP, DP, DP, DP]
        L.oracle_pinhole_dual_jacobians.restype = None
        L.oracle_pinhole_dual_jacobians.argtypes = [C.c_int64, DP, DP, DP, DP, DP, DP]
        L.oracle_eval_eucm_terms.restype = None
        L.oracle_eval_eucm_terms.argtypes = [C.c_int64, DP, DP, DP, DP, C.c_int32, C.c_double,
                                             DP, DP, DP, DP, DP, DP]
        L.oracle_eucm_dual_jacobians.restype = None
        L.oracle_eucm_dual_jacobians.argtypes = [C.c_int64, DP, DP, DP, DP, DP, DP, DP]
        for name in ("oracle_eucm_distort", "oracle_eucm_undistort"):
            getattr(L, name).restype = None
            getattr(L, name).argtypes = [C.c_int64, DP, DP, DP]
        L.oracle_eval_pose_graph3_terms.restype = None
        L.oracle_eval_pose_graph3_terms.argtypes = [C.c_int64, DP, DP, DP, DP, DP, DP]
        L.oracle_eval_isometry3_prior_terms.restype = None
        L.oracle_eval_isometry3_prior_terms.argtypes = [C.c_int64, DP, DP, DP, C.c_int32, C.c_double,
                                                        DP, DP, DP]
        L.oracle_pose_graph3_dual_jacobians.restype = None
        L.oracle_pose_graph3_dual_jacobians.argtypes = [C.c_int64, DP, DP, DP, DP, DP]
        for name in ("oracle_so3_exp", "oracle_so3_log", "oracle_so3_matrix", "oracle_se3_exp",
                     "oracle_se3_log", "oracle_so3_mat_v", "oracle_so3_mat_v_inverse",
                     "oracle_se3_inv_left_jacobian", "oracle_se3_inverse"):
            getattr(L, name).restype = None
            getattr(L, name).argtypes = [DP, DP]
        L.oracle_se3_mul.restype = None
        L.oracle_se3_mul.argtypes = [DP, DP, DP]
        L.oracle_huber_weight.restype = C.c_double
        L.oracle_huber_weight.argtypes = [C.c_double, C.c_double]
        L.oracle_inverse3.restype = C.c_int32
        L.oracle_inverse3.argtypes = [DP, DP]
        L.oracle_set_rcs_order.restype = C.c_int32
        L.oracle_set_rcs_order.argtypes = [I64P, C.c_int64]
        L.oracle_factor_stats.restype = C.c_int64
        L.oracle_factor_stats.argtypes = [I64P, I64P, DP, C.c_int64]
        L.oracle_reset_factor_stats.restype = None
        L.oracle_reset_factor_stats.argtypes = []
        L.oracle_set_threads.restype = C.c_int32
        L.oracle_set_threads.argtypes = [C.c_int32]
        _libs[variant] = L
    return _libs[variant]


def _p(a: np.ndarray):
    return a.ctypes.data_as(DP)


class Marshal:
    """Problem as ctypes structs.

    families: list of (name, type, kind, members (N x P), const_ids)
    costs: list of (kind, family_names, entity_indices (N x arity), measurements, huber_delta|None)
    """

    def __init__(self, families, costs):
        self.keep = []
        self.names = [f[0] for f in families]
        self.fams = (VarFamilyC * len(families))()
        for i, (name, typ, kind, members, const_ids) in enumerate(families):
            m = np.ascontiguousarray(members, dtype=np.float64)
            c = np.ascontiguousarray(np.asarray(const_ids, dtype=np.int64).reshape(-1))
            nm = name.encode()
            self.keep += [m, c, nm]
            f = self.fams[i]
            f.name, f.type, f.kind, f.num_members = nm, int(typ), int(kind), m.shape[0]
            f.members, f.num_const_ids, f.const_ids = _p(m), c.size, c.ctypes.data_as(I64P)
        self.costs = (CostFnC * len(costs))()
        for i, (kind, names, idx, meas, delta) in enumerate(costs):
            ix = np.ascontiguousarray(idx, dtype=np.int64)
            ms = np.ascontiguousarray(meas, dtype=np.float64)
            self.keep += [ix, ms]
            c = self.costs[i]
            c.kind = int(kind)
            c.robust_kernel = 0 if delta is None else 1
            c.huber_delta = 0.0 if delta is None else float(delta)
            for a, n in enumerate(names):
                b = n.encode()
                self.keep.append(b)
                c.family_names[a] = b
            c.num_terms = ix.shape[0]
            c.entity_indices = ix.ctypes.data_as(I64P)
            c.measurements = _p(ms)
        self.n_fam, self.n_cost = len(families), len(costs)


def _check(st: int) -> None:
    if st != 0:
        raise RuntimeError(f"oracle status {st}: {lib().oracle_last_error().decode()}")


def default_params(**kw) -> OptParamsC:
    p = OptParamsC()
    p.num_iterations = 100
    p.initial_lm_damping = 10.0
    p.parallelize = 1
    p.function_tolerance = 1e-10
    p.gradient_tolerance = 1e-14
    p.parameter_tolerance = 1e-12
    p.max_consecutive_rejects = 20
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def optimize_nlls(families, costs, params: Optional[OptParamsC] = None, max_trace: int = 1000,
                  variant: str = "", eqs=None):
    """Returns (solution dict, {name: final members}, trace list) or raises with the status.
    variant "fma" runs the oracle's rounding twin.  eqs: equality-constraint sets
    [(kind, family_name, entity_indices, params)]; the solution then holds the final
    multipliers under "lambda"."""
    m = Marshal(families, costs)
    params = params or default_params()
    outs = [np.empty_like(np.asarray(f[3], dtype=np.float64)) for f in families]
    ptrs = (DP * len(outs))(*[_p(o) for o in outs])
    sol = SolutionC()
    trace = (StepRecordC * max_trace)()
    n = C.c_int64()
    eqs = eqs or []
    eq_arr = (EqFnC * max(1, len(eqs)))()
    keep = []
    for i, (kind, name, idx, prm) in enumerate(eqs):
        ix = np.ascontiguousarray(idx, dtype=np.int64).reshape(-1)
        pr = np.ascontiguousarray(prm, dtype=np.float64).reshape(-1)
        nm = name.encode()
        keep += [ix, pr, nm]
        eq_arr[i].kind, eq_arr[i].num_terms = int(kind), ix.size
        eq_arr[i].family_names[0] = nm
        eq_arr[i].entity_indices, eq_arr[i].params = ix.ctypes.data_as(I64P), _p(pr)
    lam = np.zeros(max(1, len(eqs)))
    st = lib(variant).oracle_optimize_nlls_eq(m.fams, m.n_fam, m.costs, m.n_cost, eq_arr, len(eqs),
                                              C.byref(params), ptrs, C.byref(sol), trace, max_trace,
                                              C.byref(n), _p(lam))
    if st != 0:
        return st, None, None
    recs = [dict(merit=t.merit, lm_damping=t.lm_damping, new_merit=t.new_merit, gain_ratio=t.gain_ratio,
                 max_abs_delta=t.max_abs_delta, accepted=bool(t.accepted), termination=t.termination)
            for t in trace[: min(n.value, max_trace)]]
    solution = dict(initial_cost=sol.initial_cost, final_cost=sol.final_cost, num_steps=sol.num_steps,
                    termination=sol.termination, final_max_abs_neg_grad=sol.final_max_abs_neg_grad,
                    **({"lambda": lam[:len(eqs)].copy()} if eqs else {}))
    return solution, {f[0]: o for f, o in zip(families, outs)}, recs


def _eq_structs(eqs):
    arr = (EqFnC * max(1, len(eqs)))()
    keep = []
    for i, (kind, name, idx, prm) in enumerate(eqs):
        ix = np.ascontiguousarray(idx, dtype=np.int64).reshape(-1)
        pr = np.ascontiguousarray(prm, dtype=np.float64).reshape(-1)
        nm = name.encode()
        keep += [ix, pr, nm]
        arr[i].kind, arr[i].num_terms = int(kind), ix.size
        arr[i].family_names[0] = nm
        arr[i].entity_indices, arr[i].params = ix.ctypes.data_as(I64P), _p(pr)
    return arr, keep


def kkt_step(families, costs, eqs, nu: float, lam, num_free_scalars: int, dim: int):
    """One equality-constrained linear step: (G, -c, neg_grad + G^T lambda, [dx; dlambda])."""
    m = Marshal(families, costs)
    arr, keep = _eq_structs(eqs)
    nc = len(eqs)
    lam = np.ascontiguousarray(lam, dtype=np.float64).reshape(-1)
    G, rc = np.zeros((nc, num_free_scalars)), np.zeros(nc)
    g, dx = np.zeros(dim), np.zeros(dim + nc)
    st = lib().oracle_kkt_step(m.fams, m.n_fam, m.costs, m.n_cost, arr, nc, nu, _p(lam), _p(G), _p(rc),
                               _p(g), _p(dx))
    if st != 0:
        raise RuntimeError(f"oracle_kkt_step status {st}")
    return G, rc, g, dx


def set_rcs_order(perm=None, variant: str = "") -> None:
    """Eliminate the RCS blocks in the given order (perm[new] = old free block) in every later
    factorization; None restores the oracle's minimum-degree order."""
    if perm is None:
        lib(variant).oracle_set_rcs_order(None, 0)
        return
    p = np.ascontiguousarray(perm, dtype=np.int64)
    if lib(variant).oracle_set_rcs_order(p.ctypes.data_as(I64P), p.size) != 0:
        raise ValueError("not a permutation")


def set_threads(n: int, variant: str = "") -> int:
    """Host threads for the oracle's per-term and per-point loops; results are bitwise the
    same for any count.  Returns the previous count."""
    return int(lib(variant).oracle_set_threads(int(n)))


def factor_stats(reset: bool = True, variant: str = "") -> List[dict]:
    """Pivot statistics of the block LDLt factorizations since the last reset: zeroed and
    negative pivots, and the smallest kept |d| / max(1, max |d|) of a diagonal block."""
    cap = 100000
    z = np.zeros(cap, dtype=np.int64)
    n_ = np.zeros(cap, dtype=np.int64)
    r = np.zeros(cap)
    k = lib(variant).oracle_factor_stats(z.ctypes.data_as(I64P), n_.ctypes.data_as(I64P), _p(r), cap)
    if reset:
        lib(variant).oracle_reset_factor_stats()
    return [{"zero": int(z[i]), "neg": int(n_[i]), "min_ratio": float(r[i])} for i in range(k)]


def last_step_seconds() -> float:
    """Wall time spent inside LM steps by the last optimize_nlls call."""
    return lib().oracle_last_step_seconds()


def linearize_dense(families, costs, nu: float) -> Tuple[np.ndarray, np.ndarray]:
    m = Marshal(families, costs)
    dim = C.c_int64()
    _check(lib().oracle_linearize_dense(m.fams, m.n_fam, m.costs, m.n_cost, nu, None, None, C.byref(dim)))
    n = dim.value
    H = np.zeros((n, n))
    g = np.zeros(n)
    _check(lib().oracle_linearize_dense(m.fams, m.n_fam, m.costs, m.n_cost, nu, _p(H), _p(g), C.byref(dim)))
    return H, g


def schur_step(families, costs, nu: float, dim: int) -> Optional[np.ndarray]:
    m = Marshal(families, costs)
    dx = np.zeros(dim)
    failed = C.c_int32()
    _check(lib().oracle_schur_step(m.fams, m.n_fam, m.costs, m.n_cost, nu, _p(dx), C.byref(failed)))
    return None if failed.value else dx


def merit(families, costs) -> float:
    m = Marshal(families, costs)
    return lib().oracle_merit(m.fams, m.n_fam, m.costs, m.n_cost)


def schur_solve_dense(H: np.ndarray, dims: Sequence[int], num_free_blocks: int, g: np.ndarray):
    H = np.ascontiguousarray(H, dtype=np.float64)
    d = np.ascontiguousarray(dims, dtype=np.int64)
    g = np.ascontiguousarray(g, dtype=np.float64)
    dx = np.zeros(H.shape[0])
    failed = C.c_int32()
    lib().oracle_schur_solve_dense(len(dims), d.ctypes.data_as(I64P), num_free_blocks, _p(H), _p(g), _p(dx),
                                   C.byref(failed))
    return None if failed.value else dx


def block_ldlt_solve_dense(A: np.ndarray, dims: Sequence[int], b: np.ndarray) -> Optional[np.ndarray]:
    A = np.ascontiguousarray(A, dtype=np.float64)
    d = np.ascontiguousarray(dims, dtype=np.int64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros(A.shape[0])
    if lib().oracle_block_ldlt_solve_dense(len(dims), d.ctypes.data_as(I64P), _p(A), _p(b), _p(x)):
        return None
    return x


def sparse_ldlt_solve_dense(A: np.ndarray, dims: Sequence[int], b: np.ndarray) -> Optional[np.ndarray]:
    """SparseLdlt (sparse_ldlt.rs:571-650, 252-330) on A given dense with the block partition
    dims: the scalar LDLt in the block order, each pivot tested on its own; a zero pivot makes
    the solve divide by zero."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    d = np.ascontiguousarray(dims, dtype=np.int64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros(A.shape[0])
    with np.errstate(all="ignore"):
        if lib().oracle_sparse_ldlt_solve_dense(len(dims), d.ctypes.data_as(I64P), _p(A), _p(b), _p(x)):
            return None
    return x


def dense_ldlt(A: np.ndarray, tol_rel: float = 1e-12):
    A = np.ascontiguousarray(A, dtype=np.float64)
    n = A.shape[0]
    L = np.zeros((n, n))
    d = np.zeros(n)
    rank = C.c_int32()
    definiteness = C.c_int32()
    pc = C.c_double()
    if lib().oracle_dense_ldlt(n, _p(A), tol_rel, _p(L), _p(d), C.byref(rank), C.byref(definiteness),
                               C.byref(pc)):
        return None
    return L, d, rank.value, definiteness.value, pc.value


def dense_ldlt_solve(L, d, b):
    y = np.ascontiguousarray(b, dtype=np.float64).copy()
    L = np.ascontiguousarray(L, dtype=np.float64)
    d = np.ascontiguousarray(d, dtype=np.float64)
    lib().oracle_dense_ldlt_solve(L.shape[0], _p(L), _p(d), _p(y))
    return y


def eval_pinhole_terms(intr, poses, points, uv, huber_delta=None) -> Dict[str, np.ndarray]:
    intr = np.ascontiguousarray(intr, dtype=np.float64).reshape(-1, 4)
    poses = np.ascontiguousarray(poses, dtype=np.float64).reshape(-1, 7)
    points = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
    uv = np.ascontiguousarray(uv, dtype=np.float64).reshape(-1, 2)
    n = uv.shape[0]
    out = dict(residual=np.zeros((n, 2)), jac_pose=np.zeros((n, 2, 6)), jac_point=np.zeros((n, 2, 3)),
               cost=np.zeros(n), weight=np.zeros(n))
    lib().oracle_eval_pinhole_terms(n, _p(intr), _p(poses), _p(points), _p(uv), 0 if huber_delta is None else 1,
                                    0.0 if huber_delta is None else huber_delta,
                                    *[_p(out[k]) for k in ("residual", "jac_pose", "jac_point", "cost", "weight")])
    return out


def pinhole_dual_jacobians(intr, poses, points, uv):
    intr = np.ascontiguousarray(intr, dtype=np.float64).reshape(-1, 4)
    poses = np.ascontiguousarray(poses, dtype=np.float64).reshape(-1, 7)
    points = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
    uv = np.ascontiguousarray(uv, dtype=np.float64).reshape(-1, 2)
    n = uv.shape[0]
    jp, jx = np.zeros((n, 2, 6)), np.zeros((n, 2, 3))
    lib().oracle_pinhole_dual_jacobians(n, _p(intr), _p(poses), _p(points), _p(uv), _p(jp), _p(jx))
    return jp, jx


def eval_eucm_terms(intr, poses, points, uv, huber_delta=None) -> Dict[str, np.ndarray]:
    """EUCM reprojection terms: residual, J_intr (2x6), J_pose (2x6), J_point (2x3), cost, weight."""
    intr = np.ascontiguousarray(intr, dtype=np.float64).reshape(-1, 6)
    poses = np.ascontiguousarray(poses, dtype=np.float64).reshape(-1, 7)
    points = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
    uv = np.ascontiguousarray(uv, dtype=np.float64).reshape(-1, 2)
    n = uv.shape[0]
    out = dict(residual=np.zeros((n, 2)), jac_intr=np.zeros((n, 2, 6)), jac_pose=np.zeros((n, 2, 6)),
               jac_point=np.zeros((n, 2, 3)), cost=np.zeros(n), weight=np.zeros(n))
    lib().oracle_eval_eucm_terms(n, _p(intr), _p(poses), _p(points), _p(uv), 0 if huber_delta is None else 1,
                                 0.0 if huber_delta is None else huber_delta,
                                 *[_p(out[k]) for k in ("residual", "jac_intr", "jac_pose", "jac_point", "cost",
                                                        "weight")])
    return out


def eucm_dual_jacobians(intr, poses, points, uv):
    intr = np.ascontiguousarray(intr, dtype=np.float64).reshape(-1, 6)
    poses = np.ascontiguousarray(poses, dtype=np.float64).reshape(-1, 7)
    points = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
    uv = np.ascontiguousarray(uv, dtype=np.float64).reshape(-1, 2)
    n = uv.shape[0]
    jc, jp, jx = np.zeros((n, 2, 6)), np.zeros((n, 2, 6)), np.zeros((n, 2, 3))
    lib().oracle_eucm_dual_jacobians(n, _p(intr), _p(poses), _p(points), _p(uv), _p(jc), _p(jp), _p(jx))
    return jc, jp, jx


def eucm_distort(intr, xy):
    intr = np.ascontiguousarray(intr, dtype=np.float64).reshape(-1, 6)
    xy = np.ascontiguousarray(xy, dtype=np.float64).reshape(-1, 2)
    out = np.zeros_like(xy)
    lib().oracle_eucm_distort(xy.shape[0], _p(intr), _p(xy), _p(out))
    return out


def eucm_undistort(intr, uv):
    intr = np.ascontiguousarray(intr, dtype=np.float64).reshape(-1, 6)
    uv = np.ascontiguousarray(uv, dtype=np.float64).reshape(-1, 2)
    out = np.zeros_like(uv)
    lib().oracle_eucm_undistort(uv.shape[0], _p(intr), _p(uv), _p(out))
    return out


def eval_pose_graph3_terms(m, n_, mn):
    m = np.ascontiguousarray(m, dtype=np.float64).reshape(-1, 7)
    n_ = np.ascontiguousarray(n_, dtype=np.float64).reshape(-1, 7)
    mn = np.ascontiguousarray(mn, dtype=np.float64).reshape(-1, 7)
    k = m.shape[0]
    r, jm, jn = np.zeros((k, 6)), np.zeros((k, 6, 6)), np.zeros((k, 6, 6))
    lib().oracle_eval_pose_graph3_terms(k, _p(m), _p(n_), _p(mn), _p(r), _p(jm), _p(jn))
    return r, jm, jn


def eval_isometry3_prior_terms(poses, means, precisions, huber_delta=None):
    """Isometry3PriorCostTerm (isometry3_prior.rs:46-88): residual log(T E^-1), its Jacobian
    by duals of log(exp(x) T E^-1) at 0, and the cost r~^T W r~."""
    poses = np.ascontiguousarray(poses, dtype=np.float64).reshape(-1, 7)
    means = np.ascontiguousarray(means, dtype=np.float64).reshape(-1, 7)
    W = np.ascontiguousarray(precisions, dtype=np.float64).reshape(-1, 6, 6)
    k = poses.shape[0]
    r, J, c = np.zeros((k, 6)), np.zeros((k, 6, 6)), np.zeros(k)
    lib().oracle_eval_isometry3_prior_terms(k, _p(poses), _p(means), _p(W), 0 if huber_delta is None else 1,
                                            0.0 if huber_delta is None else float(huber_delta), _p(r), _p(J),
                                            _p(c))
    return r, J, c


def pose_graph3_dual_jacobians(m, n_, mn):
    m = np.ascontiguousarray(m, dtype=np.float64).reshape(-1, 7)
    n_ = np.ascontiguousarray(n_, dtype=np.float64).reshape(-1, 7)
    mn = np.ascontiguousarray(mn, dtype=np.float64).reshape(-1, 7)
    k = m.shape[0]
    jm, jn = np.zeros((k, 6, 6)), np.zeros((k, 6, 6))
    lib().oracle_pose_graph3_dual_jacobians(k, _p(m), _p(n_), _p(mn), _p(jm), _p(jn))
    return jm, jn


def _unary(name: str, x, out_len: int) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros(out_len)
    getattr(lib(), name)(_p(x), _p(out))
    return out


def so3_exp(w):
    return _unary("oracle_so3_exp", w, 4)


def so3_log(q):
    return _unary("oracle_so3_log", q, 3)


def so3_matrix(q):
    return _unary("oracle_so3_matrix", q, 9).reshape(3, 3)


def se3_exp(t):
    return _unary("oracle_se3_exp", t, 7)


def se3_log(T):
    return _unary("oracle_se3_log", T, 6)


def se3_inverse(T):
    return _unary("oracle_se3_inverse", T, 7)


def so3_mat_v(w):
    return _unary("oracle_so3_mat_v", w, 9).reshape(3, 3)


def so3_mat_v_inverse(w):
    return _unary("oracle_so3_mat_v_inverse", w, 9).reshape(3, 3)


def se3_inv_left_jacobian(t):
    return _unary("oracle_se3_inv_left_jacobian", t, 36).reshape(6, 6)


def se3_mul(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    out = np.zeros(7)
    lib().oracle_se3_mul(_p(a), _p(b), _p(out))
    return out


def huber_weight(delta: float, nrm: float) -> float:
    return lib().oracle_huber_weight(delta, nrm)


def inverse3(m):
    m = np.ascontiguousarray(m, dtype=np.float64).reshape(9)
    out = np.zeros(9)
    if lib().oracle_inverse3(_p(m), _p(out)):
        return None
    return out.reshape(3, 3)


def ba_families(problem, poses=None, points=None):
    """Oracle-side families of a BaProblem: cams Conditioned, points Marg, poses Free (camera 0 fixed)."""
    return [("cams", 2, 1, problem.intrinsics, []),
            ("poses", 0, 0, problem.poses_init if poses is None else poses, [0]),
            ("points", 1, 2, problem.points_init if points is None else points, [])]


def ba_costs(problem, huber_delta=1.0):
    return [(0, ["cams", "poses", "points"], problem.obs_idx, problem.obs_uv, huber_delta)]
