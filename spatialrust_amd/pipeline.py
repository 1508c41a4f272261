"""Policy-aware registration stage of the MVP pipeline.

The reference's MVP pipeline (crates/spatialrust-pipeline/src/mvp.rs:193-279)
runs voxel -> normals -> plane -> clusters -> optional registration, and its
registration step always reports `ExecutionPolicy::CpuSingle`
(mvp.rs:258-260). This module is the registration part made policy-aware on
the HIP device: the voxel stage and the GICP stage run with the caller's
policies, and each returns the receipt of what actually ran. The normals, plane
and cluster stages are not part of this path.
"""
import dataclasses
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from .registration import (ExecutionPolicy, ExecutionReceipt, GicpConfig, GicpRegistration,
                           InvalidArgument, Isometry3, RegistrationResult, VoxelGridDownsample,
                           VoxelGridDownsampleConfig)


class MvpRegistrationMethod:
    """MvpRegistrationMethod (mvp.rs:20-30). Only GICP runs on this backend."""
    POINT_TO_POINT = "point_to_point"
    POINT_TO_PLANE = "point_to_plane"
    GICP = "gicp"


@dataclass
class IcpConfig:
    """IcpConfig and its defaults (crates/spatialrust-registration/src/icp.rs:25-36)."""
    max_iterations: int = 50
    max_correspondence_distance: float = 1.0
    transformation_epsilon: float = 1e-8
    fitness_epsilon: float = 1e-6
    min_correspondences: int = 3
    initial_guess: Isometry3 = field(default_factory=Isometry3.identity)


@dataclass
class MvpIcpConfig:
    """MvpIcpConfig (mvp.rs:33-52)."""
    icp: IcpConfig = field(default_factory=IcpConfig)
    source_transform: Optional[Isometry3] = None
    method: str = MvpRegistrationMethod.POINT_TO_POINT


def gicp_config(icp: IcpConfig) -> GicpConfig:
    """The shared ICP settings mapped onto GicpConfig, other fields default
    (mvp.rs:294-305)."""
    return GicpConfig(max_iterations=icp.max_iterations,
                      max_correspondence_distance=icp.max_correspondence_distance,
                      transformation_epsilon=icp.transformation_epsilon,
                      fitness_epsilon=icp.fitness_epsilon,
                      min_correspondences=icp.min_correspondences,
                      initial_guess=icp.initial_guess)


def transform_point_cloud(points, transform: Isometry3) -> np.ndarray:
    """transform_point_cloud positions (crates/spatialrust-registration/src/transform.rs:8-25):
    Isometry3<f32>::transform_point, i.e. the normalized quaternion's f32
    rotation matrix (quat.rs:24-32, 53-71) times p, plus t, all in f32."""
    f = np.float32
    x, y, z, w = (f(v) for v in transform.rotation)
    ln = np.sqrt(x * x + y * y + z * z + w * w, dtype=f)
    if ln == f(0):
        x, y, z, w = f(0), f(0), f(0), f(1)
    else:
        x, y, z, w = x / ln, y / ln, z / ln, w / ln
    xx, yy, zz = x * x, y * y, z * z
    xy, xz, yz = x * y, x * z, y * z
    wx, wy, wz = w * x, w * y, w * z
    one, two = f(1), f(2)
    m = [[one - two * (yy + zz), two * (xy - wz), two * (xz + wy)],
         [two * (xy + wz), one - two * (xx + zz), two * (yz - wx)],
         [two * (xz - wy), two * (yz + wx), one - two * (xx + yy)]]
    p = np.asarray(points, dtype=f)
    t = [f(v) for v in transform.translation]
    out = np.empty_like(p)
    for a in range(3):
        out[:, a] = (m[a][0] * p[:, 0] + m[a][1] * p[:, 1] + m[a][2] * p[:, 2]) + t[a]
    return out


def registration_stage(downsampled, config: MvpIcpConfig,
                       policy: ExecutionPolicy = ExecutionPolicy.auto()):
    """The optional registration step of MvpPipeline::run (mvp.rs:232-262) with a
    policy: the source is the target transformed by config.source_transform
    (or a copy of it), GICP aligns it to the target, and the receipt carries the
    requested and resolved policies plus the "registration" stage. Returns
    (RegistrationResult, ExecutionReceipt)."""
    if config.method != MvpRegistrationMethod.GICP:
        raise InvalidArgument(f"registration method {config.method!r} has no HIP backend; "
                              "only gicp runs on this device")
    target = np.asarray(downsampled, dtype=np.float32)
    if config.source_transform is not None:
        source = transform_point_cloud(target, config.source_transform)
    else:
        source = target.copy()
    out = GicpRegistration(gicp_config(config.icp)).align_with_policy_and_receipt(
        source, target, policy)
    receipt = dataclasses.replace(out.receipt, stages=out.receipt.stages + ("registration",))
    return out.output, receipt


@dataclass
class MvpRegistrationResult:
    downsampled: np.ndarray
    registration: Optional[RegistrationResult]
    voxel_receipt: ExecutionReceipt
    registration_receipt: Optional[ExecutionReceipt]


class MvpRegistrationPipeline:
    """The voxel and registration stages of MvpPipeline with per-stage policies
    (mvp.rs:193-262): voxel downsample, then the optional GICP step."""

    def __init__(self, voxel: VoxelGridDownsampleConfig, icp: Optional[MvpIcpConfig] = None,
                 voxel_policy: ExecutionPolicy = ExecutionPolicy.auto(),
                 registration_policy: ExecutionPolicy = ExecutionPolicy.auto()):
        self.voxel = voxel
        self.icp = icp
        self.voxel_policy = voxel_policy
        self.registration_policy = registration_policy

    def run(self, points) -> MvpRegistrationResult:
        vox = VoxelGridDownsample(self.voxel).filter_with_policy_and_receipt(points, self.voxel_policy)
        reg, rc = (None, None)
        if self.icp is not None:
            reg, rc = registration_stage(vox.output.points, self.icp, self.registration_policy)
        return MvpRegistrationResult(vox.output.points, reg, vox.receipt, rc)
