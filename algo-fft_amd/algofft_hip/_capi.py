"""ctypes declarations of include/algofft_hip.h."""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libalgofft_hip.so")

c_i64 = ctypes.c_int64
c_int = ctypes.c_int
c_void_p = ctypes.c_void_p


class DeviceInfo(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 128), ("vendor", ctypes.c_char * 64), ("driver", ctypes.c_char * 64),
                ("memory_mb", c_i64), ("compute_cap", ctypes.c_char * 64)]


class BackendInfo(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 64), ("version", ctypes.c_char * 64),
                ("description", ctypes.c_char * 256)]


class PlanOptions(ctypes.Structure):
    _fields_ = [("device_index", c_int), ("stream_count", c_int), ("in_place", c_int)]


class PlanDesc(ctypes.Structure):
    _fields_ = [("kind", c_int), ("precision", c_int), ("rank", c_int), ("dims", c_i64 * 8), ("batch", c_i64),
                ("istride", c_i64), ("idist", c_i64), ("ostride", c_i64), ("odist", c_i64), ("flags", c_int),
                ("strategy", c_int), ("split", c_i64)]


# (name, restype, argtypes) for every exported entry point of the header.
SIGNATURES = [
    ("afft_last_error", ctypes.c_char_p, []),
    ("afft_status_string", ctypes.c_char_p, [c_int]),
    ("afft_backend_get_info", c_int, [ctypes.POINTER(BackendInfo)]),
    ("afft_backend_available", c_int, []),
    ("afft_device_count", c_int, [ctypes.POINTER(c_int)]),
    ("afft_device_get_info", c_int, [c_int, ctypes.POINTER(DeviceInfo)]),
    ("afft_context_create", c_int, [c_int, ctypes.POINTER(c_void_p)]),
    ("afft_context_device", c_int, [c_void_p, ctypes.POINTER(DeviceInfo)]),
    ("afft_context_destroy", c_int, [c_void_p]),
    ("afft_buffer_create", c_int, [c_void_p, c_i64, c_int, ctypes.POINTER(c_void_p)]),
    ("afft_buffer_len", c_i64, [c_void_p]),
    ("afft_buffer_precision", c_int, [c_void_p]),
    ("afft_buffer_upload", c_int, [c_void_p, c_void_p, c_i64]),
    ("afft_buffer_download", c_int, [c_void_p, c_void_p, c_i64]),
    ("afft_buffer_device_ptr", c_void_p, [c_void_p]),
    ("afft_buffer_destroy", c_int, [c_void_p]),
    ("afft_host_alloc", c_int, [c_void_p, c_i64, ctypes.POINTER(c_void_p)]),
    ("afft_host_free", c_int, [c_void_p]),
    ("afft_stream_create", c_int, [c_void_p, ctypes.POINTER(c_void_p)]),
    ("afft_stream_synchronize", c_int, [c_void_p]),
    ("afft_stream_destroy", c_int, [c_void_p]),
    ("afft_stream_native", c_void_p, [c_void_p]),
    ("afft_plan_create", c_int, [c_void_p, c_i64, c_int, ctypes.POINTER(PlanOptions), ctypes.POINTER(c_void_p)]),
    ("afft_plan_create_desc", c_int, [c_void_p, ctypes.POINTER(PlanDesc), ctypes.POINTER(c_void_p)]),
    ("afft_plan_len", c_i64, [c_void_p]),
    ("afft_plan_precision", c_int, [c_void_p]),
    ("afft_plan_io_len", c_int, [c_void_p, c_int, ctypes.POINTER(c_i64), ctypes.POINTER(c_i64)]),
    ("afft_plan_forward", c_int, [c_void_p, c_void_p, c_i64, c_void_p, c_i64]),
    ("afft_plan_inverse", c_int, [c_void_p, c_void_p, c_i64, c_void_p, c_i64]),
    ("afft_plan_exec_device", c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    ("afft_plan_exec_device_noalloc", c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    ("afft_noalloc_scratch_bytes", c_i64, []),
    ("afft_plan_forward_full", c_int, [c_void_p, c_void_p, c_i64, c_void_p, c_i64]),
    ("afft_plan_inverse_full", c_int, [c_void_p, c_void_p, c_i64, c_void_p, c_i64]),
    ("afft_plan_exec_full_device", c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    ("afft_plan_exec_buffers", c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    ("afft_plan_synchronize", c_int, [c_void_p]),
    ("afft_plan_describe", c_int, [c_void_p, ctypes.c_char_p, c_int]),
    ("afft_plan_workspace_bytes", c_i64, [c_void_p]),
    ("afft_plan_destroy", c_int, [c_void_p]),
    ("afft_plan_clone", c_int, [c_void_p, ctypes.POINTER(c_void_p)]),
    ("afft_fast_plan_create", c_int, [c_void_p, c_i64, c_i64, c_int, c_void_p, ctypes.POINTER(c_void_p)]),
    ("afft_fast_plan_len", c_i64, [c_void_p]),
    ("afft_fast_forward", None, [c_void_p, c_void_p, c_void_p]),
    ("afft_fast_inverse", None, [c_void_p, c_void_p, c_void_p]),
    ("afft_fast_plan_destroy", c_int, [c_void_p]),
    ("afft_ragged_create", c_int, [c_void_p, c_int, ctypes.POINTER(c_i64), c_i64, c_int, ctypes.POINTER(c_void_p)]),
    ("afft_ragged_total_len", c_i64, [c_void_p]),
    ("afft_ragged_offsets", c_int, [c_void_p, ctypes.POINTER(c_i64), c_i64]),
    ("afft_ragged_exec", c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    ("afft_ragged_exec_graph", c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    ("afft_ragged_exec_per_size", c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    ("afft_ragged_exec_per_job", c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    ("afft_ragged_destroy", c_int, [c_void_p]),
    ("afft_plan_candidates", c_int, [ctypes.POINTER(PlanDesc), ctypes.POINTER(c_int), ctypes.POINTER(c_i64), c_int]),
    ("afft_plan_features", c_int, [ctypes.POINTER(PlanDesc), c_int, c_i64, ctypes.POINTER(ctypes.c_float)]),
    ("afft_tune_time", c_int, [c_void_p, ctypes.POINTER(PlanDesc), c_int, c_int, ctypes.POINTER(ctypes.c_double)]),
    ("afft_tune_plan", c_int, [c_void_p, ctypes.POINTER(PlanDesc), c_int, c_int, ctypes.POINTER(c_int),
                               ctypes.POINTER(c_i64), ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                               c_int]),
    ("afft_wisdom_features", ctypes.c_uint64, [c_void_p, ctypes.POINTER(PlanDesc)]),
    ("afft_wisdom_export", c_int, [c_void_p, ctypes.c_char_p, c_i64, ctypes.POINTER(c_i64)]),
    ("afft_wisdom_import", c_int, [c_void_p, ctypes.c_char_p]),
    ("afft_wisdom_clear", c_int, [c_void_p]),
    ("afft_mlp_param_count", c_int, []),
    ("afft_mlp_create", c_int, [c_void_p, ctypes.c_uint64, ctypes.POINTER(c_void_p)]),
    ("afft_mlp_set_normalization", c_int, [c_void_p, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float),
                                           ctypes.c_float, ctypes.c_float]),
    ("afft_mlp_get_params", c_int, [c_void_p, ctypes.POINTER(ctypes.c_float)]),
    ("afft_mlp_set_params", c_int, [c_void_p, ctypes.POINTER(ctypes.c_float)]),
    ("afft_mlp_loss_grad", c_int, [c_void_p, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float), c_i64,
                                   ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]),
    ("afft_mlp_train", c_int, [c_void_p, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float), c_i64, c_int,
                               c_int, ctypes.c_float, ctypes.c_uint64, ctypes.POINTER(ctypes.c_float)]),
    ("afft_mlp_predict", c_int, [c_void_p, ctypes.POINTER(ctypes.c_float), c_i64, ctypes.POINTER(ctypes.c_float)]),
    ("afft_mlp_save", c_int, [c_void_p, ctypes.c_char_p]),
    ("afft_mlp_load", c_int, [c_void_p, ctypes.c_char_p]),
    ("afft_mlp_destroy", c_int, [c_void_p]),
    ("afft_context_set_cost_model", c_int, [c_void_p, c_void_p]),
    ("afft_conv_create", c_int, [c_void_p, c_int, c_i64, c_i64, c_i64, c_int, ctypes.POINTER(c_void_p)]),
    ("afft_conv_exec", c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    ("afft_conv_destroy", c_int, [c_void_p]),
    ("afft_convolve_host", c_int, [c_void_p, c_int, c_int, c_void_p, c_i64, c_void_p, c_i64, c_void_p, c_i64]),
]

_lib = None


def load():
    """Load the HIP library.  Raises RuntimeError when it has not been built:
    there is no CPU fallback behind this API."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python algo-fft_amd/build.py`")
        lib = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib
