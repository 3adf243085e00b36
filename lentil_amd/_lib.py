"""ctypes binding of libmftx.so (C ABI declared in include/mftx.h).

The library is built in-tree (``python -m lentil_amd.build``) and loaded from
this package directory.  There is no CPU fallback: every entry point raises
if the library or a HIP device is missing.

Status codes map onto the exception types lentil raises for the same class
of error: MFTX_EINVAL -> ValueError, MFTX_ENOMEM -> MemoryError,
MFTX_EDEVICE / MFTX_EHANDLE -> RuntimeError.
"""
import ctypes
import os
import threading

import torch

HERE = os.path.dirname(os.path.abspath(__file__))
# MFTX_LIBRARY names another build of the library (e.g. to compare two builds
# on one device); by default the in-tree build is loaded.
LIB_PATH = os.environ.get('MFTX_LIBRARY') or os.path.join(HERE, 'libmftx.so')

OK, EINVAL, ENOMEM, EDEVICE, EHANDLE = 0, -1, -2, -3, -4
DT_C64, DT_F32, DT_C128, DT_F64 = 0, 1, 2, 3
PREC = {'fp32': 0, 'bf16x3': 1}

c_i64 = ctypes.c_int64
c_int = ctypes.c_int
c_dbl = ctypes.c_double
c_vp = ctypes.c_void_p


class MftxTile(ctypes.Structure):
    _fields_ = [('amp', c_vp), ('opd', c_vp), ('amp_dtype', ctypes.c_int32), ('opd_dtype', ctypes.c_int32),
                ('ld_amp', c_i64), ('ld_opd', c_i64), ('m', c_i64), ('n', c_i64)]


class MftxProblem(ctypes.Structure):
    _fields_ = [('tile', ctypes.c_int32), ('wave', ctypes.c_int32), ('group', ctypes.c_int32),
                ('M', ctypes.c_int32), ('N', ctypes.c_int32),
                ('alpha_r', c_dbl), ('alpha_c', c_dbl), ('shift_r', c_dbl), ('shift_c', c_dbl),
                ('offset_r', c_dbl), ('offset_c', c_dbl), ('out_r0', c_i64), ('out_c0', c_i64),
                ('weight', c_dbl)]


class MftxPlanInfo(ctypes.Structure):
    _fields_ = [('flops', c_dbl), ('gemm_flops', c_dbl * 2), ('mfma_flops', c_dbl * 2), ('bytes_hbm_min', c_dbl),
                ('bytes_phasor', c_dbl), ('bytes_reduce', c_dbl),
                ('n_launch', ctypes.c_int32), ('coherent_fast', ctypes.c_int32),
                ('n_stage1', ctypes.c_int32), ('n_stage2', ctypes.c_int32),
                ('n_work1', ctypes.c_int32), ('n_work2', ctypes.c_int32)]


class MftxMft(ctypes.Structure):
    _fields_ = [('f', c_vp), ('amp', c_vp), ('opd', c_vp), ('f_dtype', ctypes.c_int32), ('amp_dtype', ctypes.c_int32),
                ('opd_dtype', ctypes.c_int32), ('conj', ctypes.c_int32), ('m', c_i64), ('n', c_i64), ('ldf', c_i64),
                ('ld_amp', c_i64), ('ld_opd', c_i64), ('inv_wl', c_dbl), ('F', c_vp), ('M', c_i64), ('N', c_i64),
                ('ldF', c_i64), ('alpha_r', c_dbl), ('alpha_c', c_dbl), ('shift_r', c_dbl), ('shift_c', c_dbl),
                ('offset_r', c_dbl), ('offset_c', c_dbl), ('scale', c_dbl), ('w', c_vp), ('w_dtype', ctypes.c_int32),
                ('pad_', ctypes.c_int32), ('ldw', c_i64)]


# name -> (restype, argtypes); exactly the exports of include/mftx.h
SIGNATURES = {
    'mftx_version': (c_int, []),
    'mftx_create': (c_int, [c_int, ctypes.POINTER(c_vp)]),
    'mftx_destroy': (c_int, [c_vp]),
    'mftx_last_error': (ctypes.c_char_p, [c_vp]),
    'mftx_set_stream': (c_int, [c_vp, c_vp]),
    'mftx_synchronize': (c_int, [c_vp]),
    'mftx_dft2': (c_int, [c_vp, c_vp, c_int, c_i64, c_i64, c_i64, c_vp, c_i64, c_i64, c_i64,
                          c_dbl, c_dbl, c_dbl, c_dbl, c_dbl, c_dbl, c_dbl, c_int, c_int]),
    'mftx_dft2_noalloc': (c_int, [c_vp, c_vp, c_int, c_i64, c_i64, c_i64, c_vp, c_i64, c_i64, c_i64,
                                  c_dbl, c_dbl, c_dbl, c_dbl, c_dbl, c_dbl, c_dbl, c_int, c_int]),
    'mftx_phasor': (c_int, [c_vp, c_vp, c_int, c_i64, c_vp, c_int, c_i64, c_i64, c_i64, c_dbl, c_vp]),
    'mftx_insert': (c_int, [c_vp, c_vp, c_i64, c_i64, c_i64, c_i64, c_i64, c_vp, c_int, c_i64,
                            c_i64, c_i64, c_int, c_dbl]),
    'mftx_multiply': (c_int, [c_vp, c_vp, c_i64, c_vp, c_i64, c_vp, c_i64, c_i64, c_i64]),
    'mftx_dft2_many': (c_int, [c_vp, ctypes.POINTER(MftxMft), ctypes.c_int32, ctypes.c_int32]),
    'mftx_mft_plan_create': (c_int, [c_vp, ctypes.POINTER(MftxMft), ctypes.c_int32, ctypes.c_int32,
                                     ctypes.POINTER(c_vp)]),
    'mftx_mft_plan_execute': (c_int, [c_vp]),
    'mftx_mft_plan_destroy': (c_int, [c_vp]),
    'mftx_phase_grad_many': (c_int, [c_vp, c_vp, c_int, c_i64, c_vp, c_int, c_i64, c_i64, c_i64, c_vp,
                                     ctypes.c_int32, c_vp, c_i64, c_i64, c_vp, c_i64, c_vp, c_i64]),
    'mftx_zero': (c_int, [c_vp, c_vp, c_i64]),
    'mftx_rebin': (c_int, [c_vp, c_vp, c_int, c_i64, c_i64, c_i64, c_i64, c_i64, c_vp, c_int]),
    'mftx_fft_filter': (c_int, [c_vp, c_vp, c_int, c_i64, c_i64, c_i64, c_vp, c_int, c_i64, c_int, c_vp, c_int,
                                c_i64, c_int]),
    'mftx_rescale': (c_int, [c_vp, c_vp, c_int, c_i64, c_i64, c_i64, c_vp, c_int, c_i64, c_dbl, c_i64, c_i64,
                             c_int, c_int, c_int, c_vp, c_int, c_i64]),
    'mftx_ptt_fit': (c_int, [c_vp, c_vp, c_int, c_i64, c_vp, c_int, c_i64, c_i64, ctypes.c_int32, c_i64, c_i64,
                             c_dbl, c_dbl, c_vp]),
    'mftx_remove_tilt': (c_int, [c_vp, c_vp, c_int, c_i64, c_vp, c_int, c_i64, c_i64, ctypes.c_int32,
                                 ctypes.c_int32, c_i64, c_i64, c_dbl, c_dbl, c_vp, c_vp, c_i64]),
    'mftx_zernike_basis': (c_int, [c_vp, c_vp, c_vp, c_vp, c_int, c_i64, c_vp, ctypes.c_int32, c_int, c_vp]),
    'mftx_weight_field': (c_int, [c_vp, c_vp, c_i64, c_vp, c_int, c_i64, c_dbl, c_vp, c_i64, c_i64, c_i64]),
    'mftx_phase_grad': (c_int, [c_vp, c_vp, c_int, c_i64, c_vp, c_int, c_i64, c_i64, c_i64, c_dbl, c_vp, c_i64,
                                c_vp, c_i64, c_vp, c_i64]),
    'mftx_plan_psf': (c_int, [c_vp, ctypes.POINTER(MftxTile), ctypes.c_int32, ctypes.POINTER(c_dbl),
                              ctypes.c_int32, ctypes.POINTER(MftxProblem), ctypes.c_int32, c_i64, c_i64,
                              ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(c_vp)]),
    'mftx_plan_execute': (c_int, [c_vp, c_vp, ctypes.c_int32, c_i64, ctypes.c_int32]),
    'mftx_plan_set_fields': (c_int, [c_vp, ctypes.POINTER(c_vp), c_i64, ctypes.POINTER(c_vp), ctypes.c_int32, c_i64,
                                     ctypes.POINTER(c_dbl)]),
    'mftx_plan_execute_fields': (c_int, [c_vp]),
    'mftx_plan_set_graph': (c_int, [c_vp, ctypes.c_int32]),
    'mftx_plan_get_info': (c_int, [c_vp, ctypes.POINTER(MftxPlanInfo)]),
    'mftx_plan_profile': (c_int, [c_vp, c_vp, ctypes.c_int32, c_i64, ctypes.c_int32, ctypes.POINTER(c_dbl)]),
    'mftx_plan_destroy': (c_int, [c_vp]),
}

_lib = None
_lock = threading.Lock()
_ctx = {}


def load():
    """Load libmftx.so and bind every exported symbol (no GPU access)."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(f'{LIB_PATH} is not built; run `python -m lentil_amd.build` '
                                   '(lentil_amd has no CPU fallback)')
            lib = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(lib, name)
                fn.restype = res
                fn.argtypes = args
            _lib = lib
    return _lib


def check(rc, ctx=None, what=''):
    if rc == OK:
        return
    msg = load().mftx_last_error(ctx).decode() if ctx else ''
    msg = f'{what}: {msg}' if what else msg
    if rc == EINVAL:
        raise ValueError(msg)
    if rc == ENOMEM:
        raise MemoryError(msg)
    raise RuntimeError(f'mftx error {rc}: {msg}')


def device():
    """The HIP device the path runs on (torch 'cuda' is HIP on ROCm)."""
    if not torch.cuda.is_available():
        raise RuntimeError('lentil_amd needs a HIP device; none is visible')
    return torch.device('cuda', torch.cuda.current_device())


def context(dev=None):
    """Per-device mftx context bound to torch's current stream."""
    lib = load()
    dev = dev or device()
    idx = dev.index if dev.index is not None else torch.cuda.current_device()
    ctx = _ctx.get(idx)
    if ctx is None:
        with _lock:  # one context per device, also when threads race here
            ctx = _ctx.get(idx)
            if ctx is None:
                h = c_vp()
                check(lib.mftx_create(idx, ctypes.byref(h)), None, 'mftx_create')
                ctx = h.value
                _ctx[idx] = ctx
    stream = torch.cuda.current_stream(idx).cuda_stream
    check(lib.mftx_set_stream(ctx, c_vp(stream)), ctx, 'mftx_set_stream')
    return ctx


DTYPE_CODE = {torch.complex64: DT_C64, torch.float32: DT_F32, torch.complex128: DT_C128, torch.float64: DT_F64}


def dtype_code(t):
    try:
        return DTYPE_CODE[t.dtype]
    except KeyError:
        raise TypeError(f'unsupported dtype {t.dtype}') from None


def precision_code(precision):
    try:
        return PREC[precision]
    except KeyError:
        raise ValueError(f"precision must be one of {sorted(PREC)}, got {precision!r}") from None
