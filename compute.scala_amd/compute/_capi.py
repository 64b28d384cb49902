"""ctypes binding of libcsmi (include/csmi.h).

This module is the Python side of the drop-in boundary: it loads the in-tree gfx950
library and declares every entry point. Failures raise the exception classes of the
reference's `OpenCL.Exceptions` (OpenCL.scala:143-312), chosen by the status code.
There is no fallback: if the library is missing, importing this module fails.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libcsmi.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        "libcsmi.so is not built (expected at %s); run `python compute.scala_amd/build.py`" % LIB_PATH)

lib = ctypes.CDLL(LIB_PATH)

MAX_RANK = 8


class CsNode(ctypes.Structure):
    """cs_node: one node of a serialised expression DAG."""
    _fields_ = [
        ("op", ctypes.c_int32),
        ("a", ctypes.c_int32),
        ("b", ctypes.c_int32),
        ("param", ctypes.c_int32),
        ("value", ctypes.c_float),
        ("rank", ctypes.c_int32),
        ("shape", ctypes.c_int32 * MAX_RANK),
        ("matrix", ctypes.c_double * (MAX_RANK * (MAX_RANK + 1))),
    ]


class CsFusedOptions(ctypes.Structure):
    _fields_ = [
        ("math_mode", ctypes.c_int32),
        ("vector_width", ctypes.c_int32),
        ("items_per_thread", ctypes.c_int32),
        ("block_size", ctypes.c_int32),
        ("nontemporal_store", ctypes.c_int32),
        ("launch_order", ctypes.c_int32),
    ]


class CsFusedDesc(ctypes.Structure):
    _fields_ = [
        ("nodes", ctypes.POINTER(CsNode)),
        ("nnodes", ctypes.c_int),
        ("outputs", ctypes.POINTER(ctypes.c_int32)),
        ("noutputs", ctypes.c_int),
        ("out_rank", ctypes.c_int),
        ("out_shape", ctypes.POINTER(ctypes.c_int32)),
        ("options", CsFusedOptions),
    ]


# cs_op
OP_LITERAL, OP_PARAMETER, OP_TRANSFORM, OP_FILL, OP_EXTRACT = 1, 2, 3, 4, 5
OP_NEG, OP_ABS, OP_SQRT, OP_TANH, OP_EXP, OP_LOG = 10, 11, 12, 13, 14, 15
CHECK_EXP_BOUNDED = 114  # cs_math_check: exp where its argument cannot exceed 88
OP_ADD, OP_SUB, OP_MUL, OP_DIV, OP_MOD, OP_MIN, OP_MAX = 20, 21, 22, 23, 24, 25, 26

MATH_PRECISE, MATH_FAST, MATH_ACCURATE = 0, 1, 2
ORDER_AUTO, ORDER_LINEAR = 0, 1
PROBE_READ, PROBE_WRITE, PROBE_COPY = 0, 1, 2
PROBE_STORE_PLAIN, PROBE_STORE_NT, PROBE_STORE_NT_WRITE_THROUGH = 0, 1, 2

GEMM_FP32, GEMM_BF16, GEMM_BF16X3, GEMM_FP8 = 0, 1, 2, 3

INFO_DEVICE_TYPE = 1
INFO_COMPUTE_UNITS = 2
INFO_MAX_WORK_GROUP_SIZE = 3
INFO_WAVEFRONT_SIZE = 4
INFO_LDS_BYTES = 5
INFO_GLOBAL_MEM_BYTES = 6
INFO_L2_BYTES = 7
INFO_CLOCK_KHZ = 8
INFO_ARENA_RESERVED_BYTES = 9
INFO_ARENA_IN_USE_BYTES = 10
INFO_ARENA_SLABS = 11
INFO_FUSED_CACHE_ENTRIES = 12


# ---------------------------------------------------------------- exceptions
class OpenCLException(Exception):
    """Base of the status exceptions (OpenCL.scala:143)."""
    status = None


def _exc(name, status):
    cls = type(name, (OpenCLException,), {"status": status})
    globals()[name] = cls
    return cls


_BY_STATUS = {}
for _name, _code in [
    ("DeviceNotFound", -1),
    ("MemObjectAllocationFailure", -4),
    ("OutOfResources", -5),
    ("OutOfHostMemory", -6),
    ("BuildProgramFailure", -11),
    ("InvalidValue", -30),
    ("InvalidDevice", -33),
    ("InvalidMemObject", -38),
    ("InvalidProgram", -44),
    ("InvalidKernelName", -46),
    ("InvalidKernelArgs", -52),
    ("InvalidWorkGroupSize", -54),
    ("InvalidEvent", -58),
    ("InvalidOperation", -59),
    ("InvalidBufferSize", -61),
]:
    _BY_STATUS[_code] = _exc(_name, _code)


class UnknownErrorCode(OpenCLException):
    pass


def last_error():
    n = lib.cs_last_error(None, 0)
    buf = ctypes.create_string_buffer(n + 1)
    lib.cs_last_error(buf, n + 1)
    return buf.value.decode("utf-8", "replace")


def check(status, what=""):
    if status == 0:
        return
    cls = _BY_STATUS.get(status, UnknownErrorCode)
    msg = last_error()
    err = cls("%s%s" % ((what + ": ") if what else "", msg))
    err.status = status
    raise err


# ---------------------------------------------------------------- prototypes
P = ctypes.c_void_p
I = ctypes.c_int
I32 = ctypes.c_int32
I64 = ctypes.c_int64
U32 = ctypes.c_uint32
U64 = ctypes.c_uint64
SZ = ctypes.c_size_t
PP = ctypes.POINTER(ctypes.c_void_p)
F = ctypes.c_float
D = ctypes.c_double
CP = ctypes.c_char_p

_PROTOS = {
    "cs_last_error": (SZ, [ctypes.c_char_p, SZ]),
    "cs_version": (CP, []),
    "cs_status_string": (CP, [I]),
    "cs_device_count": (I, [ctypes.POINTER(I)]),
    "cs_device_open": (I, [I, PP]),
    "cs_device_close": (I, [P]),
    "cs_device_info": (I, [P, I, ctypes.POINTER(I64)]),
    "cs_device_name": (I, [P, ctypes.POINTER(CP)]),
    "cs_stream_create": (I, [P, PP]),
    "cs_stream_destroy": (I, [P]),
    "cs_stream_synchronize": (I, [P]),
    "cs_stream_device": (I, [P, PP]),
    "cs_stream_native": (I, [P, PP]),
    "cs_event_record": (I, [P, PP]),
    "cs_event_retain": (I, [P]),
    "cs_event_release": (I, [P]),
    "cs_event_wait": (I, [P]),
    "cs_event_query": (I, [P, ctypes.POINTER(I)]),
    "cs_stream_wait_event": (I, [P, P]),
    "cs_event_elapsed_ms": (I, [P, P, ctypes.POINTER(F)]),
    "cs_event_on_complete": (I, [P, P, P]),
    "cs_buffer_alloc": (I, [P, SZ, PP]),
    "cs_buffer_alloc_on": (I, [P, SZ, PP]),
    "cs_buffer_retain": (I, [P]),
    "cs_buffer_release": (I, [P]),
    "cs_buffer_device_ptr": (I, [P, PP]),
    "cs_buffer_size": (I, [P, ctypes.POINTER(SZ)]),
    "cs_buffer_write": (I, [P, P, SZ, P, SZ, I]),
    "cs_buffer_read": (I, [P, P, SZ, P, SZ]),
    "cs_memcpy_htod": (I, [P, P, P, SZ]),
    "cs_memcpy_dtoh": (I, [P, P, P, SZ]),
    "cs_memcpy_dtod": (I, [P, P, P, SZ]),
    "cs_arena_trim": (I, [P]),
    "cs_program_build": (I, [P, CP, CP, PP, PP]),
    "cs_program_release": (I, [P]),
    "cs_program_get_kernel": (I, [P, CP, PP]),
    "cs_kernel_release": (I, [P]),
    "cs_kernel_max_block": (I, [P, ctypes.POINTER(ctypes.c_int)]),
    "cs_launch": (I, [P, P, ctypes.POINTER(U32), ctypes.POINTER(U32), U32, PP, I]),
    "cs_free": (None, [P]),
    "cs_graph_begin_capture": (I, [P]),
    "cs_graph_end_capture": (I, [P, PP]),
    "cs_graph_launch": (I, [P, P]),
    "cs_graph_release": (I, [P]),
    "cs_fused_compile": (I, [P, ctypes.POINTER(CsNode), I, ctypes.POINTER(I32), I, I,
                             ctypes.POINTER(I32), ctypes.POINTER(CsFusedOptions), PP]),
    "cs_fused_retain": (I, [P]),
    "cs_fused_release": (I, [P]),
    "cs_fused_source": (I, [P, ctypes.POINTER(CP)]),
    "cs_fused_signature": (I, [P, ctypes.POINTER(I), ctypes.POINTER(I64)]),
    "cs_fused_key": (I, [P, ctypes.POINTER(U64)]),
    "cs_fused_launch": (I, [P, P, PP, I, P]),
    "cs_fused_launch_timed": (I, [P, P, PP, I, P, PP, PP]),
    "cs_fused_compile_sum": (I, [P, ctypes.POINTER(CsNode), I, ctypes.POINTER(I32), I, I,
                                 ctypes.POINTER(I32), ctypes.POINTER(CsFusedOptions), PP]),
    "cs_fused_sum": (I, [P, P, PP, I, P, P]),
    "cs_fused_lookup": (I, [P, ctypes.POINTER(CsNode), I, ctypes.POINTER(I32), I, I,
                            ctypes.POINTER(I32), ctypes.POINTER(CsFusedOptions), ctypes.POINTER(I)]),
    "cs_fused_cache_clear": (I, [P]),
    "cs_fused_compile_many": (I, [P, ctypes.POINTER(CsFusedDesc), I, PP]),
    "cs_fused_launch_block": (I, [P, P, PP, I, P, I]),
    "cs_fused_generate_source": (I, [ctypes.POINTER(CsNode), I, ctypes.POINTER(I32), I, I,
                                     ctypes.POINTER(I32), ctypes.POINTER(CsFusedOptions), PP,
                                     ctypes.POINTER(U64)]),
    "cs_program_check": (I, [CP, CP, PP]),
    "cs_fused_precompile": (I, [ctypes.POINTER(CsNode), I, ctypes.POINTER(I32), I, I, ctypes.POINTER(I32),
                                ctypes.POINTER(CsFusedOptions)]),
    "cs_batch_create": (I, [P, PP, I, I, PP]),
    "cs_batch_bind": (I, [P, P, PP, I]),
    "cs_batch_launch": (I, [P, P]),
    "cs_batch_info": (I, [P, ctypes.POINTER(I64), ctypes.POINTER(I), ctypes.POINTER(I)]),
    "cs_batch_release": (I, [P]),
    "cs_batch_generate_source": (I, [ctypes.POINTER(CsFusedDesc), I, I, PP, ctypes.POINTER(I64)]),
    "cs_reduce_sum": (I, [P, P, I64, P, P]),
    "cs_reduce_max": (I, [P, P, I64, P]),
    "cs_reduce_rows": (I, [P, P, I64, I64, P, P, P, P]),
    "cs_reduce_rows_noalloc": (I, [P, P, I64, I64, P, P, P, P]),
    "cs_reduce_cols": (I, [P, P, I64, I64, P]),
    "cs_random": (I, [P, P, I64, U32]),
    "cs_random_normal": (I, [P, P, I64, U32]),
    "cs_fill_uniform": (I, [P, P, I64, U32, D, D]),
    "cs_fill_normal": (I, [P, P, I64, U32]),
    "cs_gemm": (I, [P, I, I, I, I, P, I, P, I, P, I]),
    "cs_math_check": (I, [P, I, I, U32, U32, P]),
    "cs_hbm_probe": (I, [P, I, I, P, P, I64, I]),
    "cs_mlp_create": (I, [P, ctypes.POINTER(I32), I, U32, PP]),
    "cs_mlp_param_count": (I, [P, ctypes.POINTER(I64)]),
    "cs_mlp_get_params": (I, [P, P, P, I64]),
    "cs_mlp_set_params": (I, [P, P, P, I64]),
    "cs_mlp_forward": (I, [P, P, P, I64, P]),
    "cs_mlp_backward": (I, [P, P, P, P, I64, P, P]),
    "cs_mlp_adam": (I, [P, P, P, F, I]),
    "cs_mlp_release": (I, [P]),
}

for _name, (_res, _args) in _PROTOS.items():
    _f = getattr(lib, _name)  # AttributeError here means the library lacks a declared symbol
    _f.restype = _res
    _f.argtypes = _args

EXPORTED = sorted(_PROTOS)


def take_string(ptr):
    """Copies a malloc'ed C string returned through a void** and frees it."""
    if not ptr:
        return ""
    s = ctypes.cast(ptr, ctypes.c_char_p).value.decode("utf-8", "replace")
    lib.cs_free(ptr)
    return s


def device_count():
    n = ctypes.c_int(0)
    check(lib.cs_device_count(ctypes.byref(n)), "cs_device_count")
    return n.value
