"""
ctypes binding of libfastbarnes.so (the C ABI declared in include/fastbarnes.h).

The library is built in-tree by ``make -C fast-barnes-py_amd`` (or
``__graft_entry__.build()``).  There is no fallback: if the shared object is
missing or fails to load, every entry point raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# FASTBARNES_LIB selects another build of the library (e.g. a tuning variant)
LIB_PATH = os.environ.get('FASTBARNES_LIB') or os.path.join(_HERE, 'libfastbarnes.so')

FB_OK = 0
FB_ERR_INVALID_ARGUMENT = 1
FB_ERR_HIP = 2
FB_ERR_UNSUPPORTED = 3
FB_ERR_OUT_OF_MEMORY = 4

FB_METHOD_OPTIMIZED_CONVOLUTION = 0
FB_METHOD_CONVOLUTION = 1
FB_METHOD_NAIVE = 2
FB_METHOD_RADIUS = 3

FB_PREC_FP64 = 0
FB_PREC_FP32C = 1

FB_FLAG_DEVICE_PTRS = 1

PRECISIONS = {'fp64': FB_PREC_FP64, 'fp32c': FB_PREC_FP32C}

_dp = ctypes.POINTER(ctypes.c_double)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_vp = ctypes.c_void_p


class FbField(ctypes.Structure):
    """Mirror of struct fb_field."""
    _fields_ = [
        ('dim', ctypes.c_int32),
        ('method', ctypes.c_int32),
        ('num_iter', ctypes.c_int32),
        ('reserved', ctypes.c_int32),
        ('n', ctypes.c_int64),
        ('pt_offset', ctypes.c_int64),
        ('out_offset', ctypes.c_int64),
        ('size', ctypes.c_int64 * 3),
        ('sigma', ctypes.c_double * 3),
        ('x0', ctypes.c_double * 3),
        ('step', ctypes.c_double * 3),
        ('max_dist_weight', ctypes.c_double),
    ]


# name -> (restype, argtypes); the exported symbols of include/fastbarnes.h
SIGNATURES = {
    'fb_version': (ctypes.c_int, []),
    'fb_last_error': (ctypes.c_char_p, []),
    'fb_device_count': (ctypes.c_int, [ctypes.POINTER(ctypes.c_int)]),
    'fb_device_arch': (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
    'fb_half_kernel_size': (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_int, _i32p]),
    'fb_tail_value': (ctypes.c_int, [ctypes.c_int, _dp, _dp, ctypes.c_int, _dp]),
    'fb_sigma_effective': (ctypes.c_int, [ctypes.c_int, _dp, _dp, ctypes.c_int, _dp]),
    'fb_threshold': (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _dp, _dp, _i32p, ctypes.c_int, _dp, ctypes.c_double,
                                    _dp]),
    'fb_barnes': (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int64, _vp, _vp, _dp, _dp, _dp,
                                 _i64p, ctypes.c_int, ctypes.c_double, ctypes.c_double, _vp, ctypes.c_uint, _vp]),
    'fb_barnes_noalloc': (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int64, _vp, _vp, _dp, _dp,
                                         _dp, _i64p, ctypes.c_int, ctypes.c_double, _vp, _vp]),
    'fb_barnes_s2': (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int64, _vp, _vp, _dp, _dp, _dp, _i64p,
                                    ctypes.c_int, ctypes.c_double, ctypes.c_int, _vp, ctypes.c_uint, _vp]),
    'fb_inject': (ctypes.c_int, [ctypes.c_int, ctypes.c_int64, _vp, _vp, _dp, _dp, _i64p, _vp, _vp, ctypes.c_uint,
                                 _vp]),
    'fb_accumulate_lines': (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int64, ctypes.c_int64, ctypes.c_int,
                                           ctypes.c_int, ctypes.c_double, _vp, _vp, ctypes.c_uint, _vp]),
    'fb_convolve_grid': (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, _i64p, _i32p, ctypes.c_int, _dp,
                                        ctypes.c_double, _vp, _vp, ctypes.c_uint, _vp]),
    'fb_plan_create': (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.POINTER(FbField), ctypes.c_int, ctypes.c_int,
                                      ctypes.c_uint]),
    'fb_plan_execute': (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp]),
    'fb_plan_execute_graph': (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp]),
    'fb_plan_forward': (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    'fb_plan_backward': (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    'fb_plan_autotune': (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp]),
    'fb_plan_choose_pass': (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_int]),
    'fb_plan_pass_choices': (ctypes.c_int, [_vp, ctypes.c_int, _i64p, _i32p, _i32p, _i32p, ctypes.c_int,
                                            ctypes.POINTER(ctypes.c_int), _i32p]),
    'fb_plan_pass_groups': (ctypes.c_int, [_vp, ctypes.c_int, _i32p, _i32p, _i32p, _i32p, ctypes.c_int,
                                           ctypes.POINTER(ctypes.c_int)]),
    'fb_plan_set_timing': (ctypes.c_int, [_vp, ctypes.c_int]),
    'fb_plan_stage_times': (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_float), ctypes.c_int,
                                           ctypes.POINTER(ctypes.c_int), ctypes.c_char_p, ctypes.c_int]),
    'fb_plan_info': (ctypes.c_int, [_vp, _i64p, _i64p]),
    'fb_plan_field_layout': (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
    'fb_plan_set_blocked': (ctypes.c_int, [_vp, ctypes.c_int]),
    'fb_pass_layout': (ctypes.c_int, [ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                      ctypes.POINTER(ctypes.c_int), _i64p]),
    'fb_plan_destroy': (ctypes.c_int, [_vp]),
}

_lib = None
# which HIP runtime instance the library is bound to: 'torch' or 'system'
RUNTIME = None


def _bind_runtime():
    """One HIP runtime per process.  PyTorch-ROCm ships its own libamdhip64 (same soname as
    the system ROCm one); two instances in one process do not share devices, streams or
    allocations.  When PyTorch is importable, its runtime is loaded first, so that the
    library's libamdhip64.so.7 dependency resolves to it and torch's streams and tensors
    can be handed to the C ABI.  Otherwise the system ROCm runtime (the library's RUNPATH)
    is used.  FASTBARNES_SYSTEM_HIP=1 forces the system runtime (no torch interop)."""
    if os.environ.get('FASTBARNES_SYSTEM_HIP') == '1':
        return 'system'
    try:
        import torch
    except ImportError:
        return 'system'
    if not getattr(torch.version, 'hip', None):
        return 'system'
    path = os.path.join(os.path.dirname(torch.__file__), 'lib', 'libamdhip64.so')
    if not os.path.exists(path):
        return 'system'
    ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
    return 'torch'


def lib():
    """Loads libfastbarnes.so once; raises RuntimeError if it is not built."""
    global _lib, RUNTIME
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError('libfastbarnes.so is not built (expected at %s); run `make -C fast-barnes-py_amd` '
                               'or __graft_entry__.build()' % LIB_PATH)
        RUNTIME = _bind_runtime()
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc):
    """Raises RuntimeError with the library's message for a non-zero status."""
    if rc != FB_OK:
        msg = lib().fb_last_error()
        raise RuntimeError(msg.decode() if msg else 'fastbarnes error %d' % rc)


def device_count():
    n = ctypes.c_int(0)
    check(lib().fb_device_count(ctypes.byref(n)))
    return n.value
