"""Python mirror of sophus_opt's NLLS surface over libsophus_gpu's C ABI.

Names follow the reference crate so that calling code reads the same:

- ``VarKind`` (crates/sophus_opt/src/variables.rs:17-24)
- ``VarFamily`` / ``VarBuilder`` / ``VarFamilies``
  (crates/sophus_opt/src/variables/var_family.rs:15-46, var_builder.rs:36)
- ``RobustKernel.huber`` (crates/sophus_opt/src/robust_kernel.rs:28-57)
- ``CostFn`` with ``PinholeCameraReprojectionCostTerm`` / ``PoseGraph3CostTerm`` terms
  (crates/sophus_opt/src/nlls/cost/cost_fn.rs:105-151,
  crates/sophus_opt/src/nlls/functor_library/costs/pinhole_reprojection.rs,
  crates/sophus_opt/src/nlls/functor_library/costs/pose_graph3.rs)
- ``OptParams`` with the reference defaults (crates/sophus_opt/src/nlls/optimizer_types.rs:225-240)
- ``optimize_nlls`` (crates/sophus_opt/src/nlls.rs:73-80), ``Optimizer`` (optimizer.rs:71-645)
- ``NllsError`` variants (optimizer_types.rs:422-504), ``TerminationReason`` (:93-122)

Every compute call goes to the HIP library; there is no CPU fallback.  Importing the
module only loads the shared library (no GPU is touched until a solve is requested).
"""
from __future__ import annotations

import ctypes as C
import enum
import math
from dataclasses import dataclass, field
from pathlib import Path
from typing import Dict, List, Optional, Sequence

import numpy as np

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE.parent / "libsophus_gpu.so"
HEADER_PATH = _HERE.parent.parent / "include" / "sophus_gpu.h"

# ---- ABI structs (include/sophus_gpu.h) -------------------------------------------------


class _VarFamily(C.Structure):
    _fields_ = [
        ("name", C.c_char_p),
        ("type", C.c_int32),
        ("kind", C.c_int32),
        ("num_members", C.c_int64),
        ("members", C.POINTER(C.c_double)),
        ("num_const_ids", C.c_int64),
        ("const_ids", C.POINTER(C.c_int64)),
    ]


class _CostFn(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("robust_kernel", C.c_int32),
        ("huber_delta", C.c_double),
        ("family_names", C.c_char_p * 3),
        ("num_terms", C.c_int64),
        ("entity_indices", C.POINTER(C.c_int64)),
        ("measurements", C.POINTER(C.c_double)),
    ]


class _EqFn(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("reserved0", C.c_int32),
        ("family_names", C.c_char_p * 1),
        ("num_terms", C.c_int64),
        ("entity_indices", C.POINTER(C.c_int64)),
        ("params", C.POINTER(C.c_double)),
    ]


class _OptParams(C.Structure):
    _fields_ = [
        ("num_iterations", C.c_int64),
        ("initial_lm_damping", C.c_double),
        ("parallelize", C.c_int32),
        ("solver", C.c_int32),
        ("skip_final_hessian", C.c_int32),
        ("precision", C.c_int32),
        ("function_tolerance", C.c_double),
        ("gradient_tolerance", C.c_double),
        ("parameter_tolerance", C.c_double),
        ("max_consecutive_rejects", C.c_int64),
        ("pcg_relative_tolerance", C.c_double),
        ("pcg_max_iterations", C.c_int64),
        ("device", C.c_int32),
        ("host_symbolic", C.c_int32),
        ("precond_weights", C.POINTER(C.c_float)),
        ("pcg_forcing", C.c_double),
    ]


class _StepInfo(C.Structure):
    _fields_ = [
        ("merit", C.c_double),
        ("smooth_cost", C.c_double),
        ("inner_step", C.c_int64),
        ("lm_damping", C.c_double),
        ("termination", C.c_int32),
        ("accepted", C.c_int32),
        ("new_merit", C.c_double),
        ("gain_ratio", C.c_double),
        ("max_abs_neg_grad", C.c_double),
        ("max_abs_delta", C.c_double),
        ("pcg_iterations", C.c_int64),
        ("ms_linearize", C.c_double),
        ("ms_factor", C.c_double),
        ("ms_solve", C.c_double),
        ("ms_cost", C.c_double),
        ("ms_total", C.c_double),
    ]


class _Solution(C.Structure):
    _fields_ = [
        ("initial_cost", C.c_double),
        ("final_cost", C.c_double),
        ("num_steps", C.c_int64),
        ("termination", C.c_int32),
        ("reserved0", C.c_int32),
        ("final_max_abs_neg_grad", C.c_double),
    ]


class _RcsStats(C.Structure):
    _fields_ = [
        ("num_free_blocks", C.c_int64),
        ("num_marg_blocks", C.c_int64),
        ("num_observations", C.c_int64),
        ("num_pose_terms", C.c_int64),
        ("nnz_s_blocks", C.c_int64),
        ("nnz_l_blocks", C.c_int64),
        ("num_schur_pairs", C.c_int64),
        ("num_levels", C.c_int64),
        ("factor_flops", C.c_double),
        ("num_supernodes", C.c_int64),
        ("max_front_blocks", C.c_int64),
        ("front_bytes", C.c_int64),
        ("mf_step_flops", C.c_double),
        ("mf_step_bytes", C.c_double),
        ("mf_step_launches", C.c_int64),
    ]


class _BaSpec(C.Structure):
    _fields_ = [
        ("layout", C.c_int32),
        ("num_cameras", C.c_int32),
        ("num_points", C.c_int64),
        ("obs_per_point_ring", C.c_int32),
        ("min_obs", C.c_int32),
        ("max_obs", C.c_int32),
        ("geometric_p", C.c_double),
        ("pixel_noise", C.c_double),
        ("rot_noise", C.c_double),
        ("trans_noise", C.c_double),
        ("point_noise", C.c_double),
        ("step_length", C.c_double),
        ("yaw_jitter", C.c_double),
        ("seed", C.c_uint64),
    ]


class _BaSizes(C.Structure):
    _fields_ = [("num_cameras", C.c_int64), ("num_points", C.c_int64), ("num_obs", C.c_int64)]


class _BatchDesc(C.Structure):
    _fields_ = [("num_problems", C.c_int64), ("cam_ptr", C.POINTER(C.c_int64)),
                ("point_ptr", C.POINTER(C.c_int64)), ("obs_ptr", C.POINTER(C.c_int64)),
                ("intrinsics", C.POINTER(C.c_double)), ("poses", C.POINTER(C.c_double)),
                ("points", C.POINTER(C.c_double)), ("obs_cam", C.POINTER(C.c_int32)),
                ("obs_point", C.POINTER(C.c_int32)), ("obs_uv", C.POINTER(C.c_double)),
                ("huber_delta", C.c_double)]


class _PrecondTrainParams(C.Structure):
    _fields_ = [("num_problems", C.c_int32), ("num_cameras", C.c_int32), ("num_points", C.c_int64),
                ("geometric_p", C.c_double), ("lm_iterations", C.c_int32), ("epochs", C.c_int32),
                ("probes", C.c_int32), ("device", C.c_int32), ("learning_rate", C.c_double),
                ("seed", C.c_uint64), ("batch_snapshots", C.c_int32), ("holdout_problems", C.c_int32)]


_DP = C.POINTER(C.c_double)
_I64P = C.POINTER(C.c_int64)
_FP = C.POINTER(C.c_float)

PRECOND_FEATURES = 32
PRECOND_HIDDEN = 128
PRECOND_OUTPUTS = 21
PRECOND_WEIGHTS = (PRECOND_HIDDEN * PRECOND_FEATURES + PRECOND_HIDDEN + PRECOND_HIDDEN * PRECOND_HIDDEN
                   + PRECOND_HIDDEN + PRECOND_OUTPUTS * PRECOND_HIDDEN + PRECOND_OUTPUTS)

# Every entry point declared in include/sophus_gpu.h with its ctypes signature.
_SIGNATURES = {
    "sgpu_abi_version": (C.c_int32, []),
    "sgpu_last_error": (C.c_char_p, []),
    "sgpu_opt_params_default": (None, [C.POINTER(_OptParams)]),
    "sgpu_optimize_nlls": (C.c_int, [C.POINTER(_VarFamily), C.c_int64, C.POINTER(_CostFn), C.c_int64,
                                     C.POINTER(_OptParams), C.POINTER(_DP), C.POINTER(_Solution)]),
    "sgpu_optimizer_create": (C.c_int, [C.POINTER(_VarFamily), C.c_int64, C.POINTER(_CostFn), C.c_int64,
                                        C.POINTER(_OptParams), C.POINTER(C.c_void_p)]),
    "sgpu_optimize_nlls_eq": (C.c_int, [C.POINTER(_VarFamily), C.c_int64, C.POINTER(_CostFn), C.c_int64,
                                        C.POINTER(_EqFn), C.c_int64, C.POINTER(_OptParams), C.POINTER(_DP),
                                        C.POINTER(_Solution)]),
    "sgpu_optimizer_create_eq": (C.c_int, [C.POINTER(_VarFamily), C.c_int64, C.POINTER(_CostFn), C.c_int64,
                                           C.POINTER(_EqFn), C.c_int64, C.POINTER(_OptParams),
                                           C.POINTER(C.c_void_p)]),
    "sgpu_optimizer_eq_state": (C.c_int, [C.c_void_p, _DP, _DP, C.c_int64]),
    "sgpu_optimizer_step": (C.c_int, [C.c_void_p, C.POINTER(_StepInfo)]),
    "sgpu_optimizer_get_members": (C.c_int, [C.c_void_p, C.c_char_p, _DP, C.c_int64]),
    "sgpu_optimizer_merit": (C.c_double, [C.c_void_p]),
    "sgpu_optimizer_lm_damping": (C.c_double, [C.c_void_p]),
    "sgpu_optimizer_eval_cost_at": (C.c_int, [C.c_void_p, C.POINTER(_DP), _DP]),
    "sgpu_optimizer_build_solution": (C.c_int, [C.c_void_p, C.POINTER(_Solution)]),
    "sgpu_optimizer_rcs_stats": (C.c_int, [C.c_void_p, C.POINTER(_RcsStats)]),
    "sgpu_analyze": (C.c_int, [C.POINTER(_VarFamily), C.c_int64, C.POINTER(_CostFn), C.c_int64,
                               C.POINTER(_RcsStats)]),
    "sgpu_analyze_rcs_order": (C.c_int, [C.POINTER(_VarFamily), C.c_int64, C.POINTER(_CostFn), C.c_int64,
                                         _I64P, C.c_int64, _I64P]),
    "sgpu_symbolic_check": (C.c_int, [C.POINTER(_VarFamily), C.c_int64, C.POINTER(_CostFn), C.c_int64,
                                      C.c_int32, _I64P, _DP]),
    "sgpu_optimizer_inverse_block": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int64, C.c_char_p, C.c_int64, _DP,
                                               C.c_int64]),
    "sgpu_optimizer_covariance_block": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int64, C.c_char_p, C.c_int64,
                                                  _DP, C.c_int64]),
    "sgpu_optimizer_rcs_order": (C.c_int, [C.c_void_p, _I64P, C.c_int64, _I64P]),
    "sgpu_optimizer_rcs_pivots": (C.c_int, [C.c_void_p, _DP, C.c_int64]),
    "sgpu_optimizer_kernel_times": (C.c_int32, [C.c_void_p, C.POINTER(C.c_char_p), _I64P, _DP, C.c_int32]),
    "sgpu_optimizer_reset_kernel_times": (None, [C.c_void_p]),
    "sgpu_optimizer_set_kernel_timing": (None, [C.c_void_p, C.c_int32]),
    "sgpu_optimizer_destroy": (None, [C.c_void_p]),
    "sgpu_eval_pinhole_terms": (C.c_int, [C.c_int32, C.c_int64, _DP, _DP, _DP, _DP, C.c_int32, C.c_double,
                                          _DP, _DP, _DP, _DP, _DP]),
    "sgpu_eval_eucm_terms": (C.c_int, [C.c_int64, _DP, _DP, _DP, _DP, C.c_int32, C.c_double,
                                       _DP, _DP, _DP, _DP, _DP, _DP]),
    "sgpu_eval_pose_graph3_terms": (C.c_int, [C.c_int64, _DP, _DP, _DP, _DP, _DP, _DP]),
    "sgpu_eval_isometry3_prior_terms": (C.c_int, [C.c_int64, _DP, _DP, _DP, C.c_int32, C.c_double, _DP, _DP,
                                                  _DP]),
    "sgpu_block_ldlt_solve": (C.c_int, [C.c_int64, _I64P, _I64P, _DP, _DP, _DP]),
    "sgpu_block_ldlt_solve_noalloc": (C.c_int, [C.c_int64, _I64P, _I64P, _DP, _DP, _DP]),
    "sgpu_sparse_ldlt_solve": (C.c_int, [C.c_int64, _I64P, _I64P, _DP, _DP, _DP]),
    "sgpu_block_pcg_solve": (C.c_int, [C.c_int64, _I64P, _I64P, _DP, _DP, _DP, C.c_double, C.c_int64,
                                       _I64P]),
    "sgpu_precond_features": (C.c_int, [C.c_int64, _I64P, _I64P, _DP, _FP]),
    "sgpu_precond_forward": (C.c_int, [C.c_int64, _FP, _FP, C.c_int32, _FP]),
    "sgpu_precond_blocks": (C.c_int, [C.c_int64, _DP, _FP, _DP]),
    "sgpu_precond_loss_grad": (C.c_int, [C.c_int64, _FP, _DP, _DP, _DP, _FP, _DP, _FP]),
    "sgpu_precond_train": (C.c_int, [C.POINTER(_PrecondTrainParams), _FP, _DP, _DP, _I64P]),
    "sgpu_batch_create": (C.c_int, [C.POINTER(_BatchDesc), C.POINTER(_OptParams), C.POINTER(C.c_void_p)]),
    "sgpu_batch_run": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, _DP]),
    "sgpu_batch_run_persistent": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, _DP]),
    "sgpu_batch_latencies": (C.c_int, [C.c_void_p, _DP, _DP]),
    "sgpu_batch_reset": (C.c_int, [C.c_void_p]),
    "sgpu_batch_results": (C.c_int, [C.c_void_p, _DP, _DP, _DP, _DP, C.POINTER(C.c_int32),
                                     C.POINTER(C.c_int32)]),
    "sgpu_batch_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), _DP, _I64P, C.POINTER(C.c_int32), _DP,
                                  C.POINTER(C.c_int32)]),
    "sgpu_batch_destroy": (None, [C.c_void_p]),
    "sgpu_generate_ba": (C.c_int, [C.POINTER(_BaSpec), C.POINTER(_BaSizes), _DP, _DP, _DP, _DP, _DP,
                                   _I64P, _DP]),
    "sgpu_generate_ba_eucm": (C.c_int, [C.c_int32, C.c_int64, C.c_uint64, C.POINTER(_BaSizes), _DP, _DP,
                                        _DP, _DP, _DP, _DP, _I64P, _DP]),
}

_lib: Optional[C.CDLL] = None
ABI_VERSION = 11  # SGPU_ABI_VERSION of include/sophus_gpu.h these bindings mirror


def library() -> C.CDLL:
    """Load libsophus_gpu.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it with sophus-rs_amd/build.py")
        lib = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        if lib.sgpu_abi_version() != ABI_VERSION:
            raise RuntimeError(f"{LIB_PATH} has ABI {lib.sgpu_abi_version()}, these bindings "
                               f"expect {ABI_VERSION}: rebuild it with sophus-rs_amd/build.py")
        _lib = lib
    return _lib


def exported_symbols() -> List[str]:
    return list(_SIGNATURES)


# ---- errors ---------------------------------------------------------------------------------


class NllsError(RuntimeError):
    """Error raised by the NLLS entry points (NllsError, optimizer_types.rs:422-504)."""

    status = -1

    def __init__(self, status: int, message: str):
        super().__init__(f"[{status}] {message}")
        self.status = status
        self.message = message


class InvalidArgument(NllsError):
    pass


class SchurConstraintOnMarginalizedFamily(NllsError):
    pass


class SchurMissingFreeOrMarg(NllsError):
    pass


class SchurCostConnectsTwoMargFamilies(NllsError):
    pass


class UnknownFamily(NllsError):
    pass


class FamilyTypeError(NllsError):
    pass


class IndexOutOfRange(NllsError):
    pass


class Unsupported(NllsError):
    pass


class DeviceError(NllsError):
    pass


class Terminated(NllsError):
    pass


_ERRORS = {
    1: InvalidArgument,
    2: SchurMissingFreeOrMarg,
    3: SchurCostConnectsTwoMargFamilies,
    4: UnknownFamily,
    5: FamilyTypeError,
    6: IndexOutOfRange,
    7: Unsupported,
    8: DeviceError,
    9: DeviceError,
    10: DeviceError,
    11: NllsError,
    12: Terminated,
    13: SchurConstraintOnMarginalizedFamily,
}


def _check(status: int) -> None:
    if status != 0:
        msg = library().sgpu_last_error().decode(errors="replace")
        raise _ERRORS.get(status, NllsError)(status, msg)


# ---- problem description -----------------------------------------------------------------------


class VarKind(enum.IntEnum):
    Free = 0
    Conditioned = 1
    Marginalized = 2


class VarType(enum.IntEnum):
    Isometry3 = 0
    Vec3 = 1
    Pinhole = 2
    Eucm = 3  # EnhancedUnifiedCameraF64: (fx, fy, cx, cy, alpha, beta)


_PARAMS = {VarType.Isometry3: 7, VarType.Vec3: 3, VarType.Pinhole: 4, VarType.Eucm: 6}


@dataclass
class VarFamily:
    """A family of variables of one type (VarFamily::new / new_with_const_ids)."""

    kind: VarKind
    type: VarType
    members: np.ndarray
    const_ids: Sequence[int] = ()

    def __post_init__(self):
        self.members = np.ascontiguousarray(self.members, dtype=np.float64).reshape(-1, _PARAMS[self.type])
        self.const_ids = np.ascontiguousarray(np.asarray(self.const_ids, dtype=np.int64).reshape(-1))

    @staticmethod
    def new(kind: VarKind, type: VarType, members) -> "VarFamily":
        return VarFamily(kind, type, members)

    @staticmethod
    def new_with_const_ids(kind: VarKind, type: VarType, members, const_ids) -> "VarFamily":
        return VarFamily(kind, type, members, const_ids)


class VarBuilder:
    """VarBuilder::new().add_family(name, family).build() (var_builder.rs:36)."""

    def __init__(self):
        self.families: Dict[str, VarFamily] = {}

    def add_family(self, name: str, family: VarFamily) -> "VarBuilder":
        self.families[name] = family
        return self

    def build(self) -> "VarFamilies":
        return VarFamilies(dict(self.families))


@dataclass
class VarFamilies:
    collection: Dict[str, VarFamily]

    def get_members(self, name: str) -> np.ndarray:
        return self.collection[name].members


@dataclass(frozen=True)
class RobustKernel:
    """RobustKernel::Huber(HuberKernel::new(delta)) (robust_kernel.rs:40-57)."""

    delta: float

    @staticmethod
    def huber(delta: float) -> "RobustKernel":
        return RobustKernel(float(delta))


class CostKind(enum.IntEnum):
    PinholeReprojection = 0
    PoseGraph3 = 1
    EucmReprojection = 2
    Isometry3Prior = 3


@dataclass
class CostFn:
    """CostFn::new_boxed((), CostTerms::new(family_names, terms), robust_kernel).

    Terms are stored vectorised: ``entity_indices`` (N x arity) and ``measurements``
    (N x 2 pixel uv for reprojection, N x 7 pose_m_from_pose_n for the pose graph, N x 43 for
    the Isometry3 prior: the mean's 7 parameters, then the 6x6 precision row-major).
    """

    kind: CostKind
    family_names: Sequence[str]
    entity_indices: np.ndarray
    measurements: np.ndarray
    robust_kernel: Optional[RobustKernel] = None

    def __post_init__(self):
        reproj = self.kind in (CostKind.PinholeReprojection, CostKind.EucmReprojection)
        prior = self.kind == CostKind.Isometry3Prior
        arity = 3 if reproj else 1 if prior else 2
        mlen = 2 if reproj else 43 if prior else 7
        self.entity_indices = np.ascontiguousarray(self.entity_indices, dtype=np.int64).reshape(-1, arity)
        self.measurements = np.ascontiguousarray(self.measurements, dtype=np.float64).reshape(-1, mlen)
        if len(self.family_names) != arity:
            raise ValueError(f"{self.kind.name} needs {arity} family names")

    @staticmethod
    def pinhole_reprojection(family_names, entity_indices, uv, robust_kernel=None) -> "CostFn":
        """PinholeCameraReprojectionCostTerm terms: (intrinsics, world_from_camera, point)."""
        return CostFn(CostKind.PinholeReprojection, family_names, entity_indices, uv, robust_kernel)

    @staticmethod
    def eucm_reprojection(family_names, entity_indices, uv, robust_kernel=None) -> "CostFn":
        """EucmCameraReprojectionCostTerm terms: (intrinsics, world_from_camera, point)."""
        return CostFn(CostKind.EucmReprojection, family_names, entity_indices, uv, robust_kernel)

    @staticmethod
    def pose_graph3(family_names, entity_indices, pose_m_from_pose_n, robust_kernel=None) -> "CostFn":
        """PoseGraph3CostTerm terms: (world_from_pose_m, world_from_pose_n)."""
        return CostFn(CostKind.PoseGraph3, family_names, entity_indices, pose_m_from_pose_n, robust_kernel)

    @staticmethod
    def isometry3_prior(family_names, entity_indices, prior_means, prior_precisions,
                        robust_kernel=None) -> "CostFn":
        """Isometry3PriorCostTerm terms (isometry3_prior.rs:36-44): one pose each, the prior
        mean (N x 7) and precision (N x 6 x 6)."""
        means = np.asarray(prior_means, dtype=np.float64).reshape(-1, 7)
        prec = np.asarray(prior_precisions, dtype=np.float64).reshape(-1, 36)
        if prec.shape[0] == 1 and means.shape[0] > 1:
            prec = np.repeat(prec, means.shape[0], axis=0)
        return CostFn(CostKind.Isometry3Prior, family_names, np.asarray(entity_indices).reshape(-1, 1),
                      np.hstack([means, prec]), robust_kernel)


class Solver(enum.IntEnum):
    SchurBlockSparseLdlt = 0
    SchurPcgJacobi = 1
    SchurPcgLearned = 2
    SchurSparseLdlt = 3  # the same factorization with SparseLdlt's scalar pivot rule


class Precision(enum.IntEnum):
    F64 = 0
    Mixed = 1


class TerminationReason(enum.IntEnum):
    NoTermination = 0
    MaxIterations = 1
    FunctionTolerance = 2
    GradientTolerance = 3
    ParameterTolerance = 4
    MaxConsecutiveRejects = 5


@dataclass
class OptParams:
    """OptParams with the reference defaults (optimizer_types.rs:225-240)."""

    num_iterations: int = 100
    initial_lm_damping: float = 10.0
    parallelize: bool = True
    solver: Solver = Solver.SchurBlockSparseLdlt
    skip_final_hessian: bool = False
    function_tolerance: float = 1e-10
    gradient_tolerance: float = 1e-14
    parameter_tolerance: float = 1e-12
    max_consecutive_rejects: int = 20
    precision: Precision = Precision.F64
    pcg_relative_tolerance: float = 1e-8
    pcg_max_iterations: int = 500
    device: int = 0
    precond_weights: Optional[np.ndarray] = None  # SchurPcgLearned: PRECOND_WEIGHTS float32
    pcg_forcing: float = 0.0  # PCG: forcing-sequence CG tolerance (0: off)
    host_symbolic: bool = False  # build the term sort, S pattern and pair lists on the host

    def _c(self) -> _OptParams:
        p = _OptParams()
        library().sgpu_opt_params_default(C.byref(p))
        p.num_iterations = int(self.num_iterations)
        p.initial_lm_damping = float(self.initial_lm_damping)
        p.parallelize = int(bool(self.parallelize))
        p.solver = int(self.solver)
        p.skip_final_hessian = int(bool(self.skip_final_hessian))
        p.precision = int(self.precision)
        p.function_tolerance = float(self.function_tolerance)
        p.gradient_tolerance = float(self.gradient_tolerance)
        p.parameter_tolerance = float(self.parameter_tolerance)
        p.max_consecutive_rejects = int(self.max_consecutive_rejects)
        p.pcg_relative_tolerance = float(self.pcg_relative_tolerance)
        p.pcg_max_iterations = int(self.pcg_max_iterations)
        p.device = int(self.device)
        p.pcg_forcing = float(self.pcg_forcing)
        p.host_symbolic = int(bool(self.host_symbolic))
        if self.precond_weights is not None:
            w = np.ascontiguousarray(self.precond_weights, dtype=np.float32).reshape(-1)
            if w.size != PRECOND_WEIGHTS:
                raise ValueError(f"precond_weights must hold {PRECOND_WEIGHTS} floats")
            self._w_keep = w  # the library copies it during the call that takes these params
            p.precond_weights = w.ctypes.data_as(C.POINTER(C.c_float))
        return p


@dataclass
class StepInfo:
    merit: float
    smooth_cost: float
    inner_step: int
    lm_damping: float
    termination: Optional[TerminationReason]
    accepted: bool
    new_merit: float
    gain_ratio: float
    max_abs_neg_grad: float
    max_abs_delta: float
    pcg_iterations: int
    ms_linearize: float
    ms_factor: float
    ms_solve: float
    ms_cost: float
    ms_total: float


@dataclass
class OptimizationSolution:
    variables: VarFamilies
    initial_cost: float
    final_cost: float
    num_steps: int
    termination: TerminationReason
    final_max_abs_neg_grad: float


class EqKind(enum.IntEnum):
    TranslationNorm = 0


@dataclass
class EqConstraintFn:
    """EqConstraintFn::new_boxed((), EqConstraints::new(family_names, terms))
    (eq_constraint_fn.rs:61-140).  All terms of one set share one constraint row and one
    multiplier (eq_system.rs:33-37,207-262)."""

    kind: EqKind
    family_name: str
    entity_indices: np.ndarray
    params: np.ndarray

    def __post_init__(self):
        self.entity_indices = np.ascontiguousarray(self.entity_indices, dtype=np.int64).reshape(-1)
        self.params = np.ascontiguousarray(self.params, dtype=np.float64).reshape(-1)

    @staticmethod
    def translation_norm(family_name: str, entity_indices, radii) -> "EqConstraintFn":
        """TranslationNormConstraint terms |t(T)| = r (ba_scale_constraint.rs:44-99)."""
        return EqConstraintFn(EqKind.TranslationNorm, family_name, entity_indices, radii)


class _Marshal:
    """Keeps the ctypes views of one problem alive for the duration of a call."""

    def __init__(self, variables: VarFamilies, cost_fns: Sequence[CostFn],
                 eq_fns: Optional[Sequence[EqConstraintFn]] = None):
        self.names = list(variables.collection)
        self.keep = []
        fams = (_VarFamily * len(self.names))()
        for i, name in enumerate(self.names):
            f = variables.collection[name]
            nm = name.encode()
            self.keep += [nm, f.members, f.const_ids]
            fams[i].name = nm
            fams[i].type = int(f.type)
            fams[i].kind = int(f.kind)
            fams[i].num_members = f.members.shape[0]
            fams[i].members = f.members.ctypes.data_as(_DP)
            fams[i].num_const_ids = f.const_ids.size
            fams[i].const_ids = f.const_ids.ctypes.data_as(_I64P)
        costs = (_CostFn * len(cost_fns))()
        for i, c in enumerate(cost_fns):
            costs[i].kind = int(c.kind)
            costs[i].robust_kernel = 1 if c.robust_kernel is not None else 0
            costs[i].huber_delta = c.robust_kernel.delta if c.robust_kernel is not None else 0.0
            for a, n in enumerate(c.family_names):
                b = n.encode()
                self.keep.append(b)
                costs[i].family_names[a] = b
            costs[i].num_terms = c.entity_indices.shape[0]
            costs[i].entity_indices = c.entity_indices.ctypes.data_as(_I64P)
            costs[i].measurements = c.measurements.ctypes.data_as(_DP)
            self.keep += [c.entity_indices, c.measurements]
        self.fams, self.costs = fams, costs
        eq_fns = list(eq_fns or [])
        self.n_eq = len(eq_fns)
        self.eqs = (_EqFn * max(1, len(eq_fns)))()
        for i, e in enumerate(eq_fns):
            nm = e.family_name.encode()
            self.keep += [nm, e.entity_indices, e.params]
            self.eqs[i].kind = int(e.kind)
            self.eqs[i].family_names[0] = nm
            self.eqs[i].num_terms = e.entity_indices.size
            self.eqs[i].entity_indices = e.entity_indices.ctypes.data_as(_I64P)
            self.eqs[i].params = e.params.ctypes.data_as(_DP)


def _termination(v: int) -> Optional[TerminationReason]:
    return None if v == 0 else TerminationReason(v)


class Optimizer:
    """Stateful LM optimizer on the GPU (Optimizer::new / step / build_solution)."""

    def __init__(self, variables: VarFamilies, cost_fns: Sequence[CostFn], params: Optional[OptParams] = None,
                 eq_fns: Optional[Sequence[EqConstraintFn]] = None):
        self.params = params or OptParams()
        self._variables = variables
        self._m = _Marshal(variables, cost_fns, eq_fns)
        h = C.c_void_p()
        _check(library().sgpu_optimizer_create_eq(self._m.fams, len(self._m.names), self._m.costs,
                                                  len(cost_fns), self._m.eqs, self._m.n_eq,
                                                  C.byref(self.params._c()), C.byref(h)))
        self._h = h

    def eq_state(self):
        """(multipliers, constraint residuals c) at the current variables, one per set."""
        n = self._m.n_eq
        lam, res = np.zeros(max(n, 1)), np.zeros(max(n, 1))
        _check(library().sgpu_optimizer_eq_state(self._h, lam.ctypes.data_as(_DP), res.ctypes.data_as(_DP),
                                                 lam.size))
        return lam[:n], res[:n]

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            library().sgpu_optimizer_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def step(self) -> StepInfo:
        s = _StepInfo()
        _check(library().sgpu_optimizer_step(self._h, C.byref(s)))
        return StepInfo(s.merit, s.smooth_cost, s.inner_step, s.lm_damping, _termination(s.termination),
                        bool(s.accepted), s.new_merit, s.gain_ratio, s.max_abs_neg_grad, s.max_abs_delta,
                        s.pcg_iterations, s.ms_linearize, s.ms_factor, s.ms_solve, s.ms_cost, s.ms_total)

    def merit(self) -> float:
        return library().sgpu_optimizer_merit(self._h)

    def lm_damping(self) -> float:
        return library().sgpu_optimizer_lm_damping(self._h)

    def members(self, name: str) -> np.ndarray:
        f = self._variables.collection[name]
        out = np.empty_like(f.members)
        _check(library().sgpu_optimizer_get_members(self._h, name.encode(), out.ctypes.data_as(_DP), out.size))
        return out

    def variables(self) -> VarFamilies:
        return VarFamilies({n: VarFamily(f.kind, f.type, self.members(n), f.const_ids)
                            for n, f in self._variables.collection.items()})

    def eval_cost_at(self, variables: VarFamilies) -> float:
        arrs = [np.ascontiguousarray(variables.collection[n].members, dtype=np.float64) for n in self._m.names]
        ptrs = (_DP * len(arrs))(*[a.ctypes.data_as(_DP) for a in arrs])
        out = C.c_double()
        _check(library().sgpu_optimizer_eval_cost_at(self._h, ptrs, C.byref(out)))
        return out.value

    def rcs_stats(self) -> dict:
        s = _RcsStats()
        _check(library().sgpu_optimizer_rcs_stats(self._h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in _RcsStats._fields_}

    def inverse_block(self, row_family: str, row_member: int, col_family: str, col_member: int) -> np.ndarray:
        """Block of the inverse of the final linearization's damped Hessian
        (DampedHessian::inverse_block_with_damping / SchurFactor::inverse_block)."""
        dims = {n: _PARAMS[f.type] - (f.type == VarType.Isometry3) for n, f in self._variables.collection.items()}
        out = np.empty((dims[row_family], dims[col_family]))
        _check(library().sgpu_optimizer_inverse_block(self._h, row_family.encode(), int(row_member),
                                                      col_family.encode(), int(col_member),
                                                      out.ctypes.data_as(_DP), out.size))
        return out

    def covariance_block(self, row_family: str, row_member: int, col_family: str, col_member: int) -> np.ndarray:
        """OptimizationSolution::covariance_block: block of H^-1 of the undamped final Hessian
        (full-rank H; a rank-deficient one raises Unsupported)."""
        dims = {n: _PARAMS[f.type] - (f.type == VarType.Isometry3) for n, f in self._variables.collection.items()}
        out = np.empty((dims[row_family], dims[col_family]))
        _check(library().sgpu_optimizer_covariance_block(self._h, row_family.encode(), int(row_member),
                                                         col_family.encode(), int(col_member),
                                                         out.ctypes.data_as(_DP), out.size))
        return out

    def rcs_order(self) -> np.ndarray:
        """Elimination order of the reduced camera system, perm[new] = old free block."""
        n = C.c_int64()
        _check(library().sgpu_optimizer_rcs_order(self._h, None, 0, C.byref(n)))
        out = np.empty(n.value, dtype=np.int64)
        _check(library().sgpu_optimizer_rcs_order(self._h, out.ctypes.data_as(_I64P), out.size, C.byref(n)))
        return out

    def rcs_pivots(self) -> np.ndarray:
        """Pivots d of the last direct factorization, (blocks x 6) in elimination order."""
        n = self.rcs_stats()["num_free_blocks"]
        out = np.empty((n, 6))
        _check(library().sgpu_optimizer_rcs_pivots(self._h, out.ctypes.data_as(_DP), out.size))
        return out

    def set_kernel_timing(self, mode: int) -> None:
        """0 off, 1 per phase kernel, 2 also per panel-step launch ("mf_step")."""
        library().sgpu_optimizer_set_kernel_timing(self._h, int(mode))

    def reset_kernel_times(self) -> None:
        library().sgpu_optimizer_reset_kernel_times(self._h)

    def kernel_times(self) -> Dict[str, tuple]:
        n = 16
        names = (C.c_char_p * n)()
        counts = (C.c_int64 * n)()
        ms = (C.c_double * n)()
        k = library().sgpu_optimizer_kernel_times(self._h, names, counts, ms, n)
        return {names[i].decode(): (int(counts[i]), float(ms[i])) for i in range(k)}

    def build_solution(self, initial_cost: float = math.nan) -> OptimizationSolution:
        s = _Solution()
        _check(library().sgpu_optimizer_build_solution(self._h, C.byref(s)))
        return OptimizationSolution(self.variables(), initial_cost, s.final_cost, s.num_steps,
                                    TerminationReason(s.termination), s.final_max_abs_neg_grad)


def analyze(variables: VarFamilies, cost_fns: Sequence[CostFn]) -> dict:
    """Host-only validation and symbolic analysis (no device needed)."""
    m = _Marshal(variables, cost_fns)
    s = _RcsStats()
    _check(library().sgpu_analyze(m.fams, len(m.names), m.costs, len(cost_fns), C.byref(s)))
    return {k: getattr(s, k) for k, _ in _RcsStats._fields_}


def analyze_rcs_order(variables: VarFamilies, cost_fns: Sequence[CostFn]) -> np.ndarray:
    """Host-only: the elimination order the GPU factorization uses, perm[new] = old free block."""
    m = _Marshal(variables, cost_fns)
    n = C.c_int64()
    _check(library().sgpu_analyze_rcs_order(m.fams, len(m.names), m.costs, len(cost_fns), None, 0,
                                            C.byref(n)))
    out = np.empty(n.value, dtype=np.int64)
    _check(library().sgpu_analyze_rcs_order(m.fams, len(m.names), m.costs, len(cost_fns),
                                            out.ctypes.data_as(_I64P), out.size, C.byref(n)))
    return out


def symbolic_check(variables: VarFamilies, cost_fns: Sequence[CostFn], device: int = 0) -> dict:
    """The host and the device symbolic analysis compared element by element (needs a GPU):
    mismatches, Schur pairs, S blocks and the two wall times in ms."""
    m = _Marshal(variables, cost_fns)
    out = np.zeros(3, dtype=np.int64)
    ms = np.zeros(2, dtype=np.float64)
    _check(library().sgpu_symbolic_check(m.fams, len(m.names), m.costs, len(cost_fns), int(device),
                                         out.ctypes.data_as(_I64P), ms.ctypes.data_as(_DP)))
    return {"mismatches": int(out[0]), "schur_pairs": int(out[1]), "s_blocks": int(out[2]),
            "host_ms": float(ms[0]), "device_ms": float(ms[1])}


def optimize_nlls(variables: VarFamilies, cost_fns: Sequence[CostFn],
                  params: Optional[OptParams] = None,
                  eq_fns: Optional[Sequence[EqConstraintFn]] = None) -> OptimizationSolution:
    """optimize_nlls(variables, cost_fns, eq_fns, params) on the GPU (nlls.rs:73-80)."""
    params = params or OptParams()
    m = _Marshal(variables, cost_fns, eq_fns)
    outs = [np.empty_like(variables.collection[n].members) for n in m.names]
    ptrs = (_DP * len(outs))(*[o.ctypes.data_as(_DP) for o in outs])
    sol = _Solution()
    _check(library().sgpu_optimize_nlls_eq(m.fams, len(m.names), m.costs, len(cost_fns), m.eqs, m.n_eq,
                                           C.byref(params._c()), ptrs, C.byref(sol)))
    fams = {n: VarFamily(variables.collection[n].kind, variables.collection[n].type, o,
                         variables.collection[n].const_ids) for n, o in zip(m.names, outs)}
    return OptimizationSolution(VarFamilies(fams), sol.initial_cost, sol.final_cost, sol.num_steps,
                                TerminationReason(sol.termination), sol.final_max_abs_neg_grad)


# ---- stage-level entry points ---------------------------------------------------------------


def eval_pinhole_terms(intrinsics, poses, points, uv, robust_kernel: Optional[RobustKernel] = None,
                       precision: Precision = Precision.F64) -> dict:
    """Kernel (a) per term on the GPU: residual, J_pose (2x6), J_point (2x3), cost, weight."""
    intr = np.ascontiguousarray(intrinsics, dtype=np.float64).reshape(-1, 4)
    T = np.ascontiguousarray(poses, dtype=np.float64).reshape(-1, 7)
    x = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
    z = np.ascontiguousarray(uv, dtype=np.float64).reshape(-1, 2)
    n = z.shape[0]
    out = {"residual": np.empty((n, 2)), "jac_pose": np.empty((n, 2, 6)), "jac_point": np.empty((n, 2, 3)),
           "cost": np.empty(n), "weight": np.empty(n)}
    _check(library().sgpu_eval_pinhole_terms(
        int(precision), n, intr.ctypes.data_as(_DP), T.ctypes.data_as(_DP), x.ctypes.data_as(_DP),
        z.ctypes.data_as(_DP), 1 if robust_kernel else 0, robust_kernel.delta if robust_kernel else 0.0,
        *[out[k].ctypes.data_as(_DP) for k in ("residual", "jac_pose", "jac_point", "cost", "weight")]))
    return out


def eval_eucm_terms(intrinsics, poses, points, uv, robust_kernel: Optional[RobustKernel] = None) -> dict:
    """EUCM reprojection terms on the GPU (FP64): residual, J_intr (2x6), J_pose (2x6),
    J_point (2x3), cost, weight."""
    intr = np.ascontiguousarray(intrinsics, dtype=np.float64).reshape(-1, 6)
    T = np.ascontiguousarray(poses, dtype=np.float64).reshape(-1, 7)
    x = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
    z = np.ascontiguousarray(uv, dtype=np.float64).reshape(-1, 2)
    n = z.shape[0]
    out = {"residual": np.empty((n, 2)), "jac_intr": np.empty((n, 2, 6)), "jac_pose": np.empty((n, 2, 6)),
           "jac_point": np.empty((n, 2, 3)), "cost": np.empty(n), "weight": np.empty(n)}
    _check(library().sgpu_eval_eucm_terms(
        n, intr.ctypes.data_as(_DP), T.ctypes.data_as(_DP), x.ctypes.data_as(_DP), z.ctypes.data_as(_DP),
        1 if robust_kernel else 0, robust_kernel.delta if robust_kernel else 0.0,
        *[out[k].ctypes.data_as(_DP) for k in ("residual", "jac_intr", "jac_pose", "jac_point", "cost",
                                                "weight")]))
    return out


def eval_pose_graph3_terms(world_from_m, world_from_n, m_from_n) -> dict:
    """Kernel (a') per term on the GPU: residual (6), J_m, J_n (6x6)."""
    a = np.ascontiguousarray(world_from_m, dtype=np.float64).reshape(-1, 7)
    b = np.ascontiguousarray(world_from_n, dtype=np.float64).reshape(-1, 7)
    c = np.ascontiguousarray(m_from_n, dtype=np.float64).reshape(-1, 7)
    n = a.shape[0]
    out = {"residual": np.empty((n, 6)), "jac_m": np.empty((n, 6, 6)), "jac_n": np.empty((n, 6, 6))}
    _check(library().sgpu_eval_pose_graph3_terms(n, a.ctypes.data_as(_DP), b.ctypes.data_as(_DP),
                                                 c.ctypes.data_as(_DP),
                                                 *[out[k].ctypes.data_as(_DP) for k in ("residual", "jac_m", "jac_n")]))
    return out


def eval_isometry3_prior_terms(poses, means, precisions, huber_delta: Optional[float] = None) -> dict:
    """Isometry3PriorCostTerm per term on the GPU: residual (6), Jacobian (6x6, by duals) and
    cost r~^T W r~."""
    a = np.ascontiguousarray(poses, dtype=np.float64).reshape(-1, 7)
    b = np.ascontiguousarray(means, dtype=np.float64).reshape(-1, 7)
    W = np.ascontiguousarray(precisions, dtype=np.float64).reshape(-1, 36)
    n = a.shape[0]
    out = {"residual": np.empty((n, 6)), "jacobian": np.empty((n, 6, 6)), "cost": np.empty(n)}
    _check(library().sgpu_eval_isometry3_prior_terms(
        n, a.ctypes.data_as(_DP), b.ctypes.data_as(_DP), W.ctypes.data_as(_DP),
        0 if huber_delta is None else 1, 0.0 if huber_delta is None else float(huber_delta),
        *[out[k].ctypes.data_as(_DP) for k in ("residual", "jacobian", "cost")]))
    return out


def _lower_blocks(A_dense: np.ndarray):
    """Lower 6x6 blocks of a dense symmetric matrix in CSC-by-block-column (column-major
    blocks); structurally zero blocks are skipped."""
    A = np.asarray(A_dense, dtype=np.float64)
    n = A.shape[0]
    if n % 6:
        raise ValueError("dimension must be a multiple of 6")
    nb = n // 6
    col_ptr, rows, vals = [0], [], []
    for c in range(nb):
        for r in range(c, nb):
            blk = A[6 * r:6 * r + 6, 6 * c:6 * c + 6]
            if r == c or np.any(blk != 0):
                rows.append(r)
                vals.append(blk.T.copy())  # column-major block
        col_ptr.append(len(rows))
    return (nb, np.asarray(col_ptr, dtype=np.int64), np.asarray(rows, dtype=np.int64),
            np.ascontiguousarray(np.stack(vals), dtype=np.float64))


NOALLOC_SCRATCH_BYTES = 64 * 1024 * 1024  # SGPU_NOALLOC_SCRATCH_BYTES


def block_ldlt_solve(A_dense: np.ndarray, b: np.ndarray, noalloc: bool = False) -> np.ndarray:
    """Kernel (d): block-sparse LDLt solve of a symmetric matrix with 6x6 blocks on the GPU.
    noalloc=True calls the allocation-free variant (device buffers from the library's static
    scratch array of NOALLOC_SCRATCH_BYTES per GPU)."""
    nb, col_ptr, rows, vals = _lower_blocks(A_dense)
    rhs = np.ascontiguousarray(b, dtype=np.float64).reshape(-1)
    x = np.empty(6 * nb)
    fn = library().sgpu_block_ldlt_solve_noalloc if noalloc else library().sgpu_block_ldlt_solve
    _check(fn(nb, col_ptr.ctypes.data_as(_I64P), rows.ctypes.data_as(_I64P), vals.ctypes.data_as(_DP),
              rhs.ctypes.data_as(_DP), x.ctypes.data_as(_DP)))
    return x


def sparse_ldlt_solve(A_dense: np.ndarray, b: np.ndarray) -> np.ndarray:
    """SparseLdlt::factorize + solve_inplace (the scalar pivot rule) of a symmetric matrix with
    6x6 blocks on the GPU; a zero pivot's component is NaN (the reference divides by it)."""
    nb, col_ptr, rows, vals = _lower_blocks(A_dense)
    rhs = np.ascontiguousarray(b, dtype=np.float64).reshape(-1)
    x = np.empty(6 * nb)
    _check(library().sgpu_sparse_ldlt_solve(nb, col_ptr.ctypes.data_as(_I64P), rows.ctypes.data_as(_I64P),
                                            vals.ctypes.data_as(_DP), rhs.ctypes.data_as(_DP),
                                            x.ctypes.data_as(_DP)))
    return x


def block_pcg_solve(A_dense: np.ndarray, b: np.ndarray, relative_tolerance: float = 1e-8,
                    max_iterations: int = 500):
    """Block-Jacobi PCG on the GPU (FP64, x0 = 0); returns (x, iterations)."""
    nb, col_ptr, rows, vals = _lower_blocks(A_dense)
    rhs = np.ascontiguousarray(b, dtype=np.float64).reshape(-1)
    x = np.empty(6 * nb)
    it = C.c_int64(0)
    _check(library().sgpu_block_pcg_solve(nb, col_ptr.ctypes.data_as(_I64P), rows.ctypes.data_as(_I64P),
                                          vals.ctypes.data_as(_DP), rhs.ctypes.data_as(_DP),
                                          x.ctypes.data_as(_DP), relative_tolerance, max_iterations,
                                          C.byref(it)))
    return x, it.value


# ---- learned block-Jacobi correction ------------------------------------------------------


def precond_weight_views(w: np.ndarray) -> dict:
    """Views of the flat weight vector: W1 (128x32), b1, W2 (128x128), b2, W3 (21x128), b3."""
    w = np.asarray(w).reshape(-1)
    out, o = {}, 0
    for name, shape in (("W1", (PRECOND_HIDDEN, PRECOND_FEATURES)), ("b1", (PRECOND_HIDDEN,)),
                        ("W2", (PRECOND_HIDDEN, PRECOND_HIDDEN)), ("b2", (PRECOND_HIDDEN,)),
                        ("W3", (PRECOND_OUTPUTS, PRECOND_HIDDEN)), ("b3", (PRECOND_OUTPUTS,))):
        n = int(np.prod(shape))
        out[name] = w[o:o + n].reshape(shape)
        o += n
    return out


def init_precond_weights(seed: int = 0, out_scale: float = 0.0) -> np.ndarray:
    """Scaled-uniform hidden layers; the output layer is zero (times `out_scale` noise), so
    the untrained correction is exactly block-Jacobi."""
    rng = np.random.default_rng(seed)
    w = np.zeros(PRECOND_WEIGHTS, dtype=np.float32)
    v = precond_weight_views(w)
    v["W1"][:] = rng.uniform(-1, 1, v["W1"].shape) * np.sqrt(3.0 / PRECOND_FEATURES)
    v["W2"][:] = rng.uniform(-1, 1, v["W2"].shape) * np.sqrt(3.0 / PRECOND_HIDDEN)
    v["W3"][:] = rng.uniform(-1, 1, v["W3"].shape) * out_scale
    return w


def precond_features(A_dense: np.ndarray) -> np.ndarray:
    """Features (num_blocks x 32) of every 6x6 block row of a symmetric block matrix."""
    nb, col_ptr, rows, vals = _lower_blocks(A_dense)
    f = np.empty((nb, PRECOND_FEATURES), dtype=np.float32)
    _check(library().sgpu_precond_features(nb, col_ptr.ctypes.data_as(_I64P), rows.ctypes.data_as(_I64P),
                                           vals.ctypes.data_as(_DP), f.ctypes.data_as(_FP)))
    return f


def precond_forward(features: np.ndarray, weights: np.ndarray, bf16: bool = False) -> np.ndarray:
    f = np.ascontiguousarray(features, dtype=np.float32).reshape(-1, PRECOND_FEATURES)
    w = np.ascontiguousarray(weights, dtype=np.float32).reshape(-1)
    out = np.empty((f.shape[0], PRECOND_OUTPUTS), dtype=np.float32)
    _check(library().sgpu_precond_forward(f.shape[0], f.ctypes.data_as(_FP), w.ctypes.data_as(_FP),
                                          int(bool(bf16)), out.ctypes.data_as(_FP)))
    return out


def precond_blocks(diag_blocks: np.ndarray, outputs: np.ndarray) -> np.ndarray:
    D = np.ascontiguousarray(diag_blocks, dtype=np.float64).reshape(-1, 6, 6)
    o = np.ascontiguousarray(outputs, dtype=np.float32).reshape(-1, PRECOND_OUTPUTS)
    out = np.empty_like(D)
    _check(library().sgpu_precond_blocks(D.shape[0], D.ctypes.data_as(_DP), o.ctypes.data_as(_FP),
                                         out.ctypes.data_as(_DP)))
    return out


def precond_loss_grad(features, diag_blocks, u, w, weights):
    f = np.ascontiguousarray(features, dtype=np.float32).reshape(-1, PRECOND_FEATURES)
    D = np.ascontiguousarray(diag_blocks, dtype=np.float64).reshape(-1, 6, 6)
    uu = np.ascontiguousarray(u, dtype=np.float64).reshape(-1, 6)
    ww = np.ascontiguousarray(w, dtype=np.float64).reshape(-1, 6)
    wt = np.ascontiguousarray(weights, dtype=np.float32).reshape(-1)
    g = np.empty(PRECOND_WEIGHTS, dtype=np.float32)
    loss = C.c_double()
    _check(library().sgpu_precond_loss_grad(f.shape[0], f.ctypes.data_as(_FP), D.ctypes.data_as(_DP),
                                            uu.ctypes.data_as(_DP), ww.ctypes.data_as(_DP),
                                            wt.ctypes.data_as(_FP), C.byref(loss), g.ctypes.data_as(_FP)))
    return loss.value, g


@dataclass
class PrecondTrainParams:
    num_problems: int = 16
    num_cameras: int = 80
    num_points: int = 3000
    geometric_p: float = 0.15
    lm_iterations: int = 10
    epochs: int = 30
    probes: int = 4
    device: int = 0
    learning_rate: float = 1e-3
    seed: int = 1000
    batch_snapshots: int = 8
    holdout_problems: int = 2

    def _c(self) -> _PrecondTrainParams:
        c = _PrecondTrainParams()
        for k, _ in _PrecondTrainParams._fields_:
            setattr(c, k, getattr(self, k))
        return c


def train_preconditioner(params: PrecondTrainParams, weights: Optional[np.ndarray] = None):
    """Train on the GPU; returns (weights, training loss history (epochs + 1), held-out loss
    history (epochs + 1), snapshots generated)."""
    w = (init_precond_weights(params.seed) if weights is None
         else np.array(weights, dtype=np.float32).reshape(-1).copy())
    hist = np.zeros(params.epochs + 1)
    val = np.zeros(params.epochs + 1)
    n = C.c_int64()
    _check(library().sgpu_precond_train(C.byref(params._c()), w.ctypes.data_as(_FP), hist.ctypes.data_as(_DP),
                                        val.ctypes.data_as(_DP), C.byref(n)))
    return w, hist, val, n.value


# ---- batched small windows -----------------------------------------------------------------


class Batch:
    """Many independent BA windows solved together on the GPU (sgpu_batch_*).

    ``windows``: BaProblem objects (one shared pinhole each, camera 0 constant, Huber).
    """

    def __init__(self, windows: Sequence["BaProblem"], params: Optional[OptParams] = None,
                 huber_delta: Optional[float] = 1.0):
        params = params or OptParams()
        self.windows = list(windows)
        nc = np.array([w.poses_init.shape[0] for w in self.windows], dtype=np.int64)
        npt = np.array([w.points_init.shape[0] for w in self.windows], dtype=np.int64)
        no = np.array([w.num_obs for w in self.windows], dtype=np.int64)
        self.cam_ptr = np.concatenate([[0], np.cumsum(nc)]).astype(np.int64)
        self.point_ptr = np.concatenate([[0], np.cumsum(npt)]).astype(np.int64)
        self.obs_ptr = np.concatenate([[0], np.cumsum(no)]).astype(np.int64)
        self.intr = np.ascontiguousarray(np.concatenate([w.intrinsics.reshape(-1) for w in self.windows]))
        self.poses = np.ascontiguousarray(np.concatenate([w.poses_init for w in self.windows]))
        self.points = np.ascontiguousarray(np.concatenate([w.points_init for w in self.windows]))
        self.obs_cam = np.ascontiguousarray(np.concatenate([w.obs_idx[:, 1] for w in self.windows]),
                                            dtype=np.int32)
        self.obs_point = np.ascontiguousarray(np.concatenate([w.obs_idx[:, 2] for w in self.windows]),
                                              dtype=np.int32)
        self.obs_uv = np.ascontiguousarray(np.concatenate([w.obs_uv for w in self.windows]))
        d = _BatchDesc()
        d.num_problems = len(self.windows)
        d.cam_ptr = self.cam_ptr.ctypes.data_as(_I64P)
        d.point_ptr = self.point_ptr.ctypes.data_as(_I64P)
        d.obs_ptr = self.obs_ptr.ctypes.data_as(_I64P)
        d.intrinsics = self.intr.ctypes.data_as(_DP)
        d.poses = self.poses.ctypes.data_as(_DP)
        d.points = self.points.ctypes.data_as(_DP)
        d.obs_cam = self.obs_cam.ctypes.data_as(C.POINTER(C.c_int32))
        d.obs_point = self.obs_point.ctypes.data_as(C.POINTER(C.c_int32))
        d.obs_uv = self.obs_uv.ctypes.data_as(_DP)
        d.huber_delta = -1.0 if huber_delta is None else float(huber_delta)
        h = C.c_void_p()
        _check(library().sgpu_batch_create(C.byref(d), C.byref(params._c()), C.byref(h)))
        self._h = h

    @property
    def num_problems(self) -> int:
        return len(self.windows)

    def run(self, iterations: int, graph: bool = True) -> float:
        secs = C.c_double()
        _check(library().sgpu_batch_run(self._h, int(iterations), int(bool(graph)), C.byref(secs)))
        return secs.value

    def run_persistent(self, iterations: int, graph: bool = True) -> float:
        """Every window's LM run (up to `iterations` iterations) in one persistent launch."""
        secs = C.c_double()
        _check(library().sgpu_batch_run_persistent(self._h, int(iterations), int(bool(graph)), C.byref(secs)))
        return secs.value

    def latencies(self) -> dict:
        """Per window, from the device clock of the last persistent run: when it was taken
        and when its run ended (ms after the first window was taken)."""
        P = self.num_problems
        t0, t1 = np.empty(P), np.empty(P)
        _check(library().sgpu_batch_latencies(self._h, t0.ctypes.data_as(_DP), t1.ctypes.data_as(_DP)))
        return {"start_ms": t0, "end_ms": t1}

    def reset(self) -> None:
        _check(library().sgpu_batch_reset(self._h))

    def results(self) -> dict:
        P = self.num_problems
        poses = np.empty_like(self.poses)
        points = np.empty_like(self.points)
        c0, c1 = np.empty(P), np.empty(P)
        steps = np.empty(P, dtype=np.int32)
        term = np.empty(P, dtype=np.int32)
        _check(library().sgpu_batch_results(self._h, poses.ctypes.data_as(_DP), points.ctypes.data_as(_DP),
                                            c0.ctypes.data_as(_DP), c1.ctypes.data_as(_DP),
                                            steps.ctypes.data_as(C.POINTER(C.c_int32)),
                                            term.ctypes.data_as(C.POINTER(C.c_int32))))
        return {"poses": poses, "points": points, "initial_cost": c0, "final_cost": c1, "steps": steps,
                "termination": term}

    def termination(self) -> np.ndarray:
        """Termination code per window (0 while running); no variable copies."""
        term = np.empty(self.num_problems, dtype=np.int32)
        _check(library().sgpu_batch_results(self._h, None, None, None, None, None,
                                            term.ctypes.data_as(C.POINTER(C.c_int32))))
        return term

    def info(self) -> dict:
        t = C.c_int32()
        ms = (C.c_double * 3)()
        lds = C.c_int64()
        pt = C.c_int32()
        pms = (C.c_double * 3)()
        grid = C.c_int32()
        _check(library().sgpu_batch_info(self._h, C.byref(t), ms, C.byref(lds), C.byref(pt), pms, C.byref(grid)))
        return {"threads": t.value, "tune_ms": {128: ms[0], 256: ms[1], 512: ms[2]}, "lds_bytes": lds.value,
                "persistent_threads": pt.value, "persistent_tune_ms": {128: pms[0], 256: pms[1], 512: pms[2]},
                "persistent_grid": grid.value}

    def window_slice(self, p: int) -> dict:
        return {"poses": slice(self.cam_ptr[p], self.cam_ptr[p + 1]),
                "points": slice(self.point_ptr[p], self.point_ptr[p + 1])}

    def close(self) -> None:
        if getattr(self, "_h", None):
            library().sgpu_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def window_spec(index: int, seed_base: int = 0) -> "BaSpec":
    """The batched workload's window `index`: the ring geometry with N_c ~ U{5..30}
    cameras, N_p ~ U{100..2000} points and U{2..8} distinct cameras per point, drawn
    from a generator seeded by the window id."""
    rng = np.random.default_rng(seed_base * 1_000_003 + index)
    nc = int(rng.integers(5, 31))
    npt = int(rng.integers(100, 2001))
    return BaSpec(layout=0, num_cameras=nc, num_points=npt, obs_per_point_ring=0, min_obs=2, max_obs=8,
                  seed=seed_base * 1_000_003 + index)


# ---- synthetic problems ------------------------------------------------------------------------


@dataclass
class BaSpec:
    layout: int = 0
    num_cameras: int = 10
    num_points: int = 200
    obs_per_point_ring: int = 5
    min_obs: int = 2
    max_obs: int = 30
    geometric_p: float = 0.1
    pixel_noise: float = 0.5
    rot_noise: float = 0.05
    trans_noise: float = 0.1
    point_noise: float = 0.1
    step_length: float = 0.3
    yaw_jitter: float = math.radians(5.0)
    seed: int = 0

    def _c(self) -> _BaSpec:
        s = _BaSpec()
        for k, _ in _BaSpec._fields_:
            setattr(s, k, getattr(self, k))
        return s


@dataclass
class BaProblem:
    intrinsics: np.ndarray
    poses_gt: np.ndarray
    poses_init: np.ndarray
    points_gt: np.ndarray
    points_init: np.ndarray
    obs_idx: np.ndarray
    obs_uv: np.ndarray

    @property
    def num_obs(self) -> int:
        return self.obs_uv.shape[0]

    def variables(self, poses=None, points=None) -> VarFamilies:
        """Families 'cams' (Conditioned), 'points' (Marginalized), 'poses' (Free, camera 0 fixed)."""
        return (VarBuilder()
                .add_family("cams", VarFamily(VarKind.Conditioned, VarType.Pinhole, self.intrinsics))
                .add_family("poses", VarFamily(VarKind.Free, VarType.Isometry3,
                                               self.poses_init if poses is None else poses, [0]))
                .add_family("points", VarFamily(VarKind.Marginalized, VarType.Vec3,
                                                self.points_init if points is None else points))
                .build())

    def cost_fns(self, huber_delta: Optional[float] = 1.0) -> List[CostFn]:
        rk = RobustKernel.huber(huber_delta) if huber_delta is not None else None
        return [CostFn.pinhole_reprojection(["cams", "poses", "points"], self.obs_idx, self.obs_uv, rk)]


def generate_ba(spec: BaSpec) -> BaProblem:
    """Seeded synthetic BA problem (host code of the library)."""
    lib = library()
    cs = spec._c()
    sizes = _BaSizes()
    _check(lib.sgpu_generate_ba(C.byref(cs), C.byref(sizes), *([None] * 6), None))
    nc, npt, no = sizes.num_cameras, sizes.num_points, sizes.num_obs
    intr = np.empty(4)
    pg, pi = np.empty((nc, 7)), np.empty((nc, 7))
    xg, xi = np.empty((npt, 3)), np.empty((npt, 3))
    idx = np.empty((no, 3), dtype=np.int64)
    uv = np.empty((no, 2))
    _check(lib.sgpu_generate_ba(C.byref(cs), C.byref(sizes), intr.ctypes.data_as(_DP), pg.ctypes.data_as(_DP),
                                pi.ctypes.data_as(_DP), xg.ctypes.data_as(_DP), xi.ctypes.data_as(_DP),
                                idx.ctypes.data_as(_I64P), uv.ctypes.data_as(_DP)))
    return BaProblem(intr.reshape(1, 4), pg, pi, xg, xi, idx, uv)


@dataclass
class EucmBaProblem:
    """The reference's BaProblem (example_problems/ba_problem.rs:131-412): one shared EUCM camera,
    a ring of poses (0 and 1 held constant), points marginalized."""

    intrinsics_gt: np.ndarray
    intrinsics_init: np.ndarray
    poses_gt: np.ndarray
    poses_init: np.ndarray
    points_gt: np.ndarray
    points_init: np.ndarray
    obs_idx: np.ndarray
    obs_uv: np.ndarray

    @property
    def num_obs(self) -> int:
        return self.obs_uv.shape[0]

    def variables(self, cams_kind: VarKind = VarKind.Free, intrinsics=None, poses=None,
                  points=None, const_poses=(0, 1)) -> VarFamilies:
        """BaProblem::build_variables: 'cams' (the EUCM camera; from the ground truth when
        Conditioned), 'poses' (Free, 0 and 1 constant; BaScaleConstraintProblem holds only 0,
        ba_scale_constraint.rs:124-155), 'points' (Marginalized)."""
        if intrinsics is None:
            intrinsics = self.intrinsics_gt if cams_kind == VarKind.Conditioned else self.intrinsics_init
        return (VarBuilder()
                .add_family("cams", VarFamily(cams_kind, VarType.Eucm, intrinsics))
                .add_family("poses", VarFamily(VarKind.Free, VarType.Isometry3,
                                               self.poses_init if poses is None else poses,
                                               list(const_poses)))
                .add_family("points", VarFamily(VarKind.Marginalized, VarType.Vec3,
                                                self.points_init if points is None else points))
                .build())

    def cost_fns(self, huber_delta: Optional[float] = None) -> List[CostFn]:
        rk = RobustKernel.huber(huber_delta) if huber_delta is not None else None
        return [CostFn.eucm_reprojection(["cams", "poses", "points"], self.obs_idx, self.obs_uv, rk)]

    def gt_errors(self, variables: VarFamilies):
        """BaProblem::gt_errors: RMS pose translation error and RMS point error."""
        P = variables.get_members("poses")
        X = variables.get_members("points")
        pose_rms = float(np.sqrt(np.mean(np.sum((P[:, 4:] - self.poses_gt[:, 4:]) ** 2, axis=1))))
        pt_rms = float(np.sqrt(np.mean(np.sum((X - self.points_gt) ** 2, axis=1))))
        return pose_rms, pt_rms


def generate_ba_eucm(num_cameras: int, num_points: int, seed: int = 0) -> EucmBaProblem:
    """The reference BaProblem's geometry with the library's seeded RNG (host code)."""
    lib = library()
    sizes = _BaSizes()
    _check(lib.sgpu_generate_ba_eucm(num_cameras, num_points, seed, C.byref(sizes), *([None] * 8)))
    nc, npt, no = sizes.num_cameras, sizes.num_points, sizes.num_obs
    ig, ii = np.empty(6), np.empty(6)
    pg, pi = np.empty((nc, 7)), np.empty((nc, 7))
    xg, xi = np.empty((npt, 3)), np.empty((npt, 3))
    idx = np.empty((no, 3), dtype=np.int64)
    uv = np.empty((no, 2))
    _check(lib.sgpu_generate_ba_eucm(num_cameras, num_points, seed, C.byref(sizes),
                                     *[a.ctypes.data_as(_DP) for a in (ig, ii, pg, pi, xg, xi)],
                                     idx.ctypes.data_as(_I64P), uv.ctypes.data_as(_DP)))
    return EucmBaProblem(ig.reshape(1, 6), ii.reshape(1, 6), pg, pi, xg, xi, idx, uv)


def config_spec(index: int, seed: int = 0) -> BaSpec:
    """The workloads: 0 = 10-camera ring / 200 points, 1 = 500-camera walk / 100k points,
    2 = 2000-camera walk / 800k points (about 6M observations)."""
    if index == 0:
        return BaSpec(layout=0, num_cameras=10, num_points=200, obs_per_point_ring=5, seed=seed)
    if index == 1:
        return BaSpec(layout=1, num_cameras=500, num_points=100_000, geometric_p=0.1, seed=seed)
    if index == 2:
        return BaSpec(layout=1, num_cameras=2000, num_points=800_000, geometric_p=0.135, seed=seed)
    raise ValueError(index)
