"""ctypes view of the C ABI declared in include/sr_gicp.h.

The shared library `spatialrust_amd/lib/libsr_gicp.so` is built by
`spatialrust_amd.build` (hipcc, gfx950). Loading it never falls back to
anything else: if the library is missing, `load()` raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# SR_GICP_LIBRARY selects another build of the library (e.g. a tuning variant).
LIB_PATH = os.environ.get("SR_GICP_LIBRARY") or os.path.join(_HERE, "lib", "libsr_gicp.so")

SR_OK = 0
SR_ERR_INVALID_ARGUMENT = 3
SR_ERR_LENGTH_MISMATCH = 5
SR_STATUS_UNAVAILABLE = 100

SR_MEM_HOST = 0
SR_MEM_DEVICE = 1

SR_CPU_SINGLE, SR_CPU_PARALLEL, SR_GPU, SR_AUTO = 0, 1, 2, 3
SR_DEV_CPU, SR_DEV_WGPU, SR_DEV_CUDA, SR_DEV_HIP = 0, 1, 2, 3

SR_STAGE_VOXEL = 0x1
SR_STAGE_COVARIANCE = 0x2
SR_STAGE_ITERATE = 0x4
SR_STAGE_READBACK = 0x8
SR_STAGE_GRID = 0x10


class XyzView(ctypes.Structure):
    _fields_ = [("x", ctypes.c_void_p), ("y", ctypes.c_void_p), ("z", ctypes.c_void_p),
                ("n", ctypes.c_uint64), ("mem", ctypes.c_uint32)]


class GicpConfigC(ctypes.Structure):
    _fields_ = [("max_iterations", ctypes.c_uint32),
                ("max_correspondence_distance", ctypes.c_float),
                ("k_neighbors", ctypes.c_uint32),
                ("epsilon", ctypes.c_double),
                ("transformation_epsilon", ctypes.c_double),
                ("fitness_epsilon", ctypes.c_double),
                ("min_correspondences", ctypes.c_uint32),
                ("init_q", ctypes.c_float * 4),
                ("init_t", ctypes.c_float * 3),
                ("covariance_radius", ctypes.c_float),
                ("voxel_leaf", ctypes.c_float)]


class PolicyC(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("device", ctypes.c_int32)]


class ResultC(ctypes.Structure):
    _fields_ = [("q", ctypes.c_float * 4), ("t", ctypes.c_float * 3), ("fitness", ctypes.c_double),
                ("iterations", ctypes.c_uint64), ("converged", ctypes.c_uint8),
                ("pad_", ctypes.c_uint8 * 7), ("q64", ctypes.c_double * 4),
                ("t64", ctypes.c_double * 3), ("source_points", ctypes.c_uint64),
                ("target_points", ctypes.c_uint64), ("correspondences", ctypes.c_uint64)]


class ReceiptC(ctypes.Structure):
    _fields_ = [("requested", PolicyC), ("resolved", PolicyC), ("h2d_bytes", ctypes.c_uint64),
                ("d2d_bytes", ctypes.c_uint64), ("d2h_bytes", ctypes.c_uint64),
                ("stage_mask", ctypes.c_uint32), ("pad_", ctypes.c_uint32)]


class OdometryConfigC(ctypes.Structure):
    _fields_ = [("voxel_leaf", ctypes.c_float), ("pad_", ctypes.c_uint32),
                ("max_map_points", ctypes.c_uint64)]


class ServerConfigC(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_uint64), ("max_active", ctypes.c_uint32),
                ("max_iterations", ctypes.c_uint32), ("delta_pose_epsilon", ctypes.c_double)]


# Every symbol the header declares, with (restype, argtypes).
_P = ctypes.c_void_p
_I32 = ctypes.c_int32
SIGNATURES = {
    "sr_last_error": (ctypes.c_char_p, []),
    "sr_device_available": (_I32, []),
    "sr_version": (_I32, []),
    "sr_gicp_config_default": (None, [ctypes.POINTER(GicpConfigC)]),
    "sr_option_set": (_I32, [ctypes.c_char_p, ctypes.c_int64]),
    "sr_option_get": (_I32, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int64)]),
    "sr_tune_describe": (_I32, [ctypes.c_char_p, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]),
    "sr_profile_enable": (None, [_I32]),
    "sr_profile_reset": (None, []),
    "sr_profile_read": (_I32, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_double),
                               ctypes.POINTER(ctypes.c_uint64)]),
    "sr_profile_read_each": (ctypes.c_int64, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_double),
                                              ctypes.c_uint64]),
    "sr_profile_read_stats": (_I32, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_uint64)]),
    "sr_gicp_align": (_I32, [ctypes.POINTER(XyzView), ctypes.POINTER(XyzView),
                             ctypes.POINTER(GicpConfigC), PolicyC, ctypes.POINTER(ResultC),
                             ctypes.POINTER(ReceiptC)]),
    "sr_gicp_align_noalloc": (_I32, [ctypes.POINTER(XyzView), ctypes.POINTER(XyzView),
                                     ctypes.POINTER(GicpConfigC), ctypes.POINTER(ResultC)]),
    "sr_grid_normals": (_I32, [ctypes.POINTER(XyzView), ctypes.c_float, _P, _P, _P]),
    "sr_batch_align_pairs": (_I32, [ctypes.POINTER(XyzView), ctypes.POINTER(XyzView), ctypes.c_uint32,
                                    ctypes.POINTER(GicpConfigC), ctypes.POINTER(ResultC)]),
    "sr_gicp_normal_equations": (_I32, [ctypes.POINTER(XyzView), ctypes.POINTER(XyzView),
                                        ctypes.POINTER(GicpConfigC), _P, _P, _P, _P, _P, _P, _P]),
    "sr_voxel_downsample": (_I32, [ctypes.POINTER(XyzView), ctypes.c_float, _P, PolicyC, _P, _P, _P,
                                   _P, _P, ctypes.POINTER(ctypes.c_uint64),
                                   ctypes.POINTER(ReceiptC)]),
    "sr_plane_covariances": (_I32, [ctypes.POINTER(XyzView), ctypes.c_uint32, ctypes.c_float,
                                    ctypes.c_double, _P, _P]),
    "sr_knn": (_I32, [ctypes.POINTER(XyzView), ctypes.POINTER(XyzView), ctypes.c_uint32, _P, _P]),
    "sr_map_create": (_I32, [ctypes.POINTER(XyzView), ctypes.POINTER(GicpConfigC),
                             ctypes.POINTER(_P)]),
    "sr_map_destroy": (None, [_P]),
    "sr_map_points": (ctypes.c_uint64, [_P]),
    "sr_map_device_bytes": (ctypes.c_uint64, [_P]),
    "sr_map_align": (_I32, [_P, ctypes.POINTER(XyzView), ctypes.POINTER(GicpConfigC),
                            ctypes.POINTER(ResultC)]),
    "sr_batch_align": (_I32, [ctypes.POINTER(_P), ctypes.POINTER(XyzView), ctypes.c_uint32,
                              ctypes.POINTER(GicpConfigC), ctypes.POINTER(ResultC)]),
    "sr_server_create": (_I32, [ctypes.POINTER(ServerConfigC), ctypes.POINTER(GicpConfigC),
                                ctypes.POINTER(_P)]),
    "sr_server_destroy": (None, [_P]),
    "sr_server_submit": (_I32, [_P, _P, ctypes.POINTER(XyzView), ctypes.POINTER(ctypes.c_uint64)]),
    "sr_server_step": (_I32, [_P, ctypes.POINTER(ctypes.c_uint32)]),
    "sr_server_poll": (_I32, [_P, ctypes.c_uint64, ctypes.POINTER(ResultC),
                              ctypes.POINTER(ctypes.c_uint8)]),
    "sr_server_latency": (_I32, [_P, ctypes.c_uint64, ctypes.POINTER(ctypes.c_double)]),
    "sr_server_device_bytes": (ctypes.c_uint64, [_P]),
    "sr_server_finished": (_I32, [_P, _P, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]),
    "sr_odometry_create": (_I32, [ctypes.POINTER(OdometryConfigC), ctypes.POINTER(GicpConfigC),
                                  ctypes.POINTER(_P)]),
    "sr_odometry_destroy": (None, [_P]),
    "sr_odometry_push": (_I32, [_P, ctypes.POINTER(XyzView), ctypes.POINTER(ResultC)]),
    "sr_odometry_push_many": (_I32, [_P, ctypes.POINTER(XyzView), ctypes.c_uint32,
                                     ctypes.POINTER(ResultC)]),
    "sr_odometry_map_points": (ctypes.c_uint64, [_P]),
    "sr_odometry_map_stats": (_I32, [_P, ctypes.POINTER(ctypes.c_uint64)]),
    "sr_odometry_map_read": (_I32, [_P, _P, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]),
}

_lib = None


class NativeLibraryMissing(RuntimeError):
    pass


def load():
    """Loads the HIP library (built beforehand by spatialrust_amd.build)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise NativeLibraryMissing(
            f"{LIB_PATH} is missing: build it with `python -m spatialrust_amd.build` "
            "(hipcc --offload-arch=gfx950); there is no CPU fallback")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def last_error() -> str:
    msg = load().sr_last_error()
    return msg.decode() if msg else ""
