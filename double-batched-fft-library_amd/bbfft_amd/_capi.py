"""ctypes binding of the C ABI declared in include/bbfft_hip.h.

The shared library is double-batched-fft-library_amd/lib/libbbfft_hip.so (built by
__graft_entry__.build() or `make -C double-batched-fft-library_amd`); BBFFT_LIB overrides the
path. Loading fails loudly when it is missing: there is no fallback implementation.
"""
import ctypes
import os

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEFAULT_LIB = os.path.join(_PKG, "lib", "libbbfft_hip.so")

# Every symbol include/bbfft_hip.h declares (checked by the CPU test suite).
EXPORTED = [
    "bbfft_last_error", "bbfft_version", "bbfft_parse_fft_descriptor",
    "bbfft_configuration_to_string", "bbfft_default_istride", "bbfft_default_ostride",
    "bbfft_device_info_normalize", "bbfft_factor", "bbfft_iroot", "bbfft_is_prime",
    "bbfft_root_of_unity",
    "bbfft_num_passes", "bbfft_kernel_configuration", "bbfft_generate_source",
    "bbfft_compile_kernel", "bbfft_kernel_configuration_tuned", "bbfft_generate_source_tuned",
    "bbfft_plan_create_tuned", "bbfft_adjoint_configuration", "bbfft_plan_create_adjoint",
    "bbfft_cache_create", "bbfft_cache_load_aot", "bbfft_cache_size", "bbfft_cache_destroy",
    "bbfft_plan_create", "bbfft_plan_create_from_descriptor", "bbfft_plan_execute",
    "bbfft_plan_num_launches", "bbfft_plan_launch_info", "bbfft_plan_destroy",
    "bbfft_device_info_query", "bbfft_aot_generate",
    "bbfft_ragged_kernel_configuration", "bbfft_ragged_variant_configuration",
    "bbfft_ragged_job_variants", "bbfft_ragged_job_table", "bbfft_generate_ragged_source",
    "bbfft_ragged_plan_create", "bbfft_aot_generate_ragged", "bbfft_plan_create_half_input",
    "bbfft_fused_configuration", "bbfft_generate_fused_source",
    "bbfft_half_input_kernel_configuration", "bbfft_cache_load_aot_memory",
    "bbfft_aot_archive_kernels", "bbfft_plan_execute_static_scratch",
]

SUCCESS = 0
ERROR_BAD_CONFIGURATION = 1
ERROR_RUNTIME = 2
ERROR_INVALID_ARGUMENT = 3
ERROR_PARSE = 4
ERROR_INTERNAL = 5

_sz5 = ctypes.c_size_t * 5


class CConfiguration(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_uint),
        ("shape", _sz5),
        ("precision", ctypes.c_int),
        ("direction", ctypes.c_int),
        ("type", ctypes.c_int),
        ("istride", _sz5),
        ("ostride", _sz5),
        ("callback_source", ctypes.c_char_p),
        ("callback_length", ctypes.c_size_t),
        ("load_function", ctypes.c_char_p),
        ("store_function", ctypes.c_char_p),
        ("callback_language", ctypes.c_int),
    ]


class CKernelConfig(ctypes.Structure):
    _fields_ = [
        ("M", ctypes.c_size_t), ("Mb", ctypes.c_size_t), ("N", ctypes.c_size_t),
        ("K", ctypes.c_size_t),
        ("Nt", ctypes.c_size_t), ("Kb", ctypes.c_size_t), ("lds_pad", ctypes.c_size_t),
        ("lds_bytes", ctypes.c_size_t), ("work_group_size", ctypes.c_size_t),
        ("factorization", ctypes.c_int * 8), ("num_factors", ctypes.c_int),
        ("inplace_unsupported", ctypes.c_int),
    ]


class CTuning(ctypes.Structure):
    _fields_ = [
        ("factorization", ctypes.c_int * 8), ("num_factors", ctypes.c_int),
        ("Nt", ctypes.c_size_t), ("Mb", ctypes.c_size_t), ("Kb", ctypes.c_size_t),
        ("lds_pad", ctypes.c_size_t), ("no_pad", ctypes.c_int), ("flags", ctypes.c_uint),
        ("has_flags", ctypes.c_int), ("waves_per_eu", ctypes.c_uint),
        ("groups_per_cu", ctypes.c_uint),
    ]


class CRaggedJob(ctypes.Structure):
    _fields_ = [
        ("M", ctypes.c_size_t), ("N", ctypes.c_size_t), ("K", ctypes.c_size_t),
        ("direction", ctypes.c_int),
        ("in_offset", ctypes.c_size_t), ("out_offset", ctypes.c_size_t),
    ]


class CRaggedInfo(ctypes.Structure):
    _fields_ = [
        ("group_size", ctypes.c_size_t), ("lds_bytes", ctypes.c_size_t),
        ("num_variants", ctypes.c_size_t), ("total_tiles", ctypes.c_ulonglong),
    ]


class CFusedInfo(ctypes.Structure):
    _fields_ = [
        ("first_pass", ctypes.c_uint), ("group_size", ctypes.c_size_t),
        ("lds_bytes", ctypes.c_size_t), ("planes", ctypes.c_size_t), ("ka", ctypes.c_size_t),
        ("kb", ctypes.c_size_t), ("plane_elems", ctypes.c_size_t),
        ("in_plane_stride", ctypes.c_size_t), ("out_plane_stride", ctypes.c_size_t),
    ]


RAGGED_GRAPH = 1
INPUT_F16 = 8
INPUT_BF16 = 16
MFMA = 32
PREFETCH = 64

_lib = None


def lib_path():
    return os.environ.get("BBFFT_LIB", DEFAULT_LIB)


def _init_device_runtime():
    """Let torch (if present) initialise the HIP runtime before the library is loaded: the
    library links the system HIP runtime, and a process that loaded it before torch touched
    the GPU found no device on the first torch GPU call (hipErrorNoDevice on MI355X). torch
    initialising first makes both use one runtime, as in every torch-first program."""
    try:
        import torch
    except ImportError:  # pragma: no cover - the C ABI works without torch
        return
    torch.cuda.is_available()


def load():
    """Load libbbfft_hip.so (once) and declare the prototypes."""
    global _lib
    if _lib is not None:
        return _lib
    path = lib_path()
    if not os.path.exists(path):
        raise ImportError(
            f"libbbfft_hip.so not found at {path}; build it with __graft_entry__.build() "
            "or `make -C double-batched-fft-library_amd`")
    _init_device_runtime()
    L = ctypes.CDLL(path)
    c_int, c_uint, c_size, c_char_p, c_void_p = (ctypes.c_int, ctypes.c_uint, ctypes.c_size_t,
                                                 ctypes.c_char_p, ctypes.c_void_p)
    P = ctypes.POINTER
    protos = {
        "bbfft_last_error": (c_char_p, []),
        "bbfft_version": (c_char_p, []),
        "bbfft_parse_fft_descriptor": (c_int, [c_char_p, P(CConfiguration)]),
        "bbfft_configuration_to_string": (c_int, [P(CConfiguration), c_char_p, c_size]),
        "bbfft_default_istride": (c_int, [c_uint, _sz5, c_int, c_int, _sz5]),
        "bbfft_default_ostride": (c_int, [c_uint, _sz5, c_int, c_int, _sz5]),
        "bbfft_device_info_normalize": (c_int, [c_char_p, c_char_p, c_size]),
        "bbfft_factor": (c_int, [c_uint, c_uint, P(c_uint)]),
        "bbfft_iroot": (c_uint, [c_uint, c_uint]),
        "bbfft_is_prime": (c_int, [c_uint]),
        "bbfft_root_of_unity": (c_int, [ctypes.c_long, ctypes.c_long, c_int,
                                        P(ctypes.c_double), P(ctypes.c_double)]),
        "bbfft_num_passes": (c_int, [P(CConfiguration), P(c_uint)]),
        "bbfft_kernel_configuration": (c_int, [P(CConfiguration), c_uint, c_char_p,
                                               P(CKernelConfig), c_char_p, c_size]),
        "bbfft_generate_source": (c_int, [P(CConfiguration), c_uint, c_char_p, c_char_p, c_size,
                                          P(c_size)]),
        "bbfft_compile_kernel": (c_int, [P(CConfiguration), c_uint, c_char_p, c_char_p,
                                         P(c_size)]),
        "bbfft_kernel_configuration_tuned": (c_int, [P(CConfiguration), P(CTuning), c_char_p,
                                                     P(CKernelConfig), c_char_p, c_size]),
        "bbfft_generate_source_tuned": (c_int, [P(CConfiguration), P(CTuning), c_char_p,
                                                c_char_p, c_size, P(c_size)]),
        "bbfft_plan_create_tuned": (c_int, [P(CConfiguration), P(CTuning), c_void_p, c_void_p,
                                            P(c_void_p)]),
        "bbfft_adjoint_configuration": (c_int, [P(CConfiguration), P(CConfiguration)]),
        "bbfft_plan_create_adjoint": (c_int, [P(CConfiguration), c_void_p, c_void_p,
                                              P(c_void_p)]),
        "bbfft_cache_create": (c_int, [P(c_void_p)]),
        "bbfft_cache_load_aot": (c_int, [c_void_p, c_char_p, c_int, P(c_size)]),
        "bbfft_cache_size": (c_int, [c_void_p, P(c_size)]),
        "bbfft_cache_load_aot_memory": (c_int, [c_void_p, c_void_p, c_size, c_int, P(c_size)]),
        "bbfft_aot_archive_kernels": (c_int, [c_void_p, c_size, c_char_p, c_size, P(c_size)]),
        "bbfft_cache_destroy": (None, [c_void_p]),
        "bbfft_plan_create": (c_int, [P(CConfiguration), c_void_p, c_void_p, P(c_void_p)]),
        "bbfft_plan_create_from_descriptor": (c_int, [c_char_p, c_void_p, c_void_p,
                                                      P(c_void_p)]),
        "bbfft_plan_execute": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_uint32,
                                       P(c_void_p)]),
        "bbfft_plan_execute_static_scratch": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p,
                                                      ctypes.c_uint32, P(c_void_p)]),
        "bbfft_plan_num_launches": (c_int, [c_void_p, P(c_uint)]),
        "bbfft_plan_launch_info": (c_int, [c_void_p, c_uint, c_char_p, c_size, P(c_uint),
                                           P(c_uint), P(c_size)]),
        "bbfft_plan_destroy": (None, [c_void_p]),
        "bbfft_device_info_query": (c_int, [c_int, c_char_p, c_size]),
        "bbfft_aot_generate": (c_int, [P(c_char_p), c_size, c_char_p, c_char_p, c_char_p, c_uint,
                                       c_uint, P(c_size)]),
        "bbfft_ragged_kernel_configuration": (c_int, [c_int, P(CRaggedJob), c_size, c_char_p,
                                                      P(CRaggedInfo), c_char_p, c_size]),
        "bbfft_ragged_variant_configuration": (c_int, [c_int, P(CRaggedJob), c_size, c_char_p,
                                                       c_size, P(CKernelConfig)]),
        "bbfft_ragged_job_variants": (c_int, [c_int, P(CRaggedJob), c_size, c_char_p,
                                              P(c_uint)]),
        "bbfft_ragged_job_table": (c_int, [c_int, P(CRaggedJob), c_size, c_char_p, c_size,
                                           P(ctypes.c_ulonglong), c_size, P(c_size)]),
        "bbfft_generate_ragged_source": (c_int, [c_int, P(CRaggedJob), c_size, c_char_p,
                                                 c_char_p, c_size, P(c_size)]),
        "bbfft_ragged_plan_create": (c_int, [c_int, P(CRaggedJob), c_size, c_uint, c_void_p,
                                             c_void_p, P(c_void_p)]),
        "bbfft_aot_generate_ragged": (c_int, [c_int, P(CRaggedJob), c_size, c_char_p, c_char_p,
                                              c_char_p]),
        "bbfft_plan_create_half_input": (c_int, [P(CConfiguration), c_uint, c_void_p, c_void_p,
                                                 P(c_void_p)]),
        "bbfft_half_input_kernel_configuration": (c_int, [P(CConfiguration), c_uint, c_char_p,
                                                          P(CKernelConfig), c_char_p, c_size]),
        "bbfft_fused_configuration": (c_int, [P(CConfiguration), c_char_p, P(c_int),
                                              P(CFusedInfo), c_char_p, c_size]),
        "bbfft_generate_fused_source": (c_int, [P(CConfiguration), c_char_p, c_char_p, c_size,
                                                P(c_size)]),
    }
    for name, (res, args) in protos.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L
