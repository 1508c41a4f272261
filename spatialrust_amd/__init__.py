"""spatialrust_amd — MI355X (gfx950) GICP scan-registration path.

Drop-in GPU backend for SpatialRust's GicpRegistration / ExecutionPolicy
surface: voxel-centroid downsample, grid k-NN / radius covariances, grid
nearest-neighbour correspondences and the Gauss-Newton iteration, as HIP
kernels behind the C ABI in include/sr_gicp.h.
"""
from .registration import (DeviceKind, DeviceUnavailable, ExecutionOutput, ExecutionPolicy,
                           ExecutionReceipt, GicpConfig, GicpRegistration, GicpScanMatcher,
                           GpuNormalEstimator, InvalidArgument, NormalEstimationConfig,
                           Isometry3, RegistrationResult, RollingMapOdometry, Server, SpatialError,
                           TargetMap,
                           VoxelGridDownsample, VoxelGridDownsampleConfig, VoxelOutput,
                           batch_align, batch_align_pairs, device_available, device_view, get_option,
                           gicp_normal_equations, knn, plane_covariances, register_gicp,
                           set_option, tune_describe)

__all__ = [
    "DeviceKind", "DeviceUnavailable", "ExecutionOutput", "ExecutionPolicy", "ExecutionReceipt",
    "GicpConfig", "GicpRegistration", "GicpScanMatcher", "GpuNormalEstimator", "InvalidArgument",
    "NormalEstimationConfig", "Isometry3", "RegistrationResult",
    "RollingMapOdometry", "Server", "SpatialError", "TargetMap", "VoxelGridDownsample", "VoxelGridDownsampleConfig", "VoxelOutput",
    "batch_align", "batch_align_pairs", "device_available", "device_view", "gicp_normal_equations", "knn",
    "plane_covariances", "register_gicp", "get_option", "set_option", "tune_describe",
]
