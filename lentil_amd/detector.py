"""Detector-grid post-processing on the device (lentil detector.py:167-251).

``pixel`` applies the square-pixel MTF and ``pixelate`` follows it with the
cubic-spline rescale to native sampling.  The MTF convolution is
``mftx_fft_filter`` (two MFTs with alpha = 1/N, the MTF multiplied into the
forward transform's store, then |.|); the MTF itself is built on the host
with the reference's numpy expressions, so its values are identical, and is
kept on the device per (shape, parameters).

NumPy in gives float64 NumPy out (the reference's dtype); a device tensor in
gives a device tensor of its dtype (float32 or float64).
"""
import ctypes
import functools

import numpy as np
import torch

from . import _lib
from ._arrays import is_dev, to_dev


def real_image(img):
    """(contiguous 2-D float32/float64 device tensor, came-from-host flag)."""
    host = not is_dev(img)
    if host and np.iscomplexobj(np.asarray(img)):
        raise TypeError('expected a real image')
    t = to_dev(img, real=True)
    if t.dtype not in (torch.float32, torch.float64):
        t = t.to(torch.float64)
    if t.dim() != 2:
        raise ValueError(f'expected a 2-D image, got shape {tuple(t.shape)}')
    return t.contiguous(), host


def finish(t, host, dtype=np.float64):
    return t.cpu().numpy().astype(dtype, copy=False) if host else t


@functools.lru_cache(maxsize=16)
def _device_kernel(kind, shape, params, index):
    from . import convolvable
    if kind == 'pixel':
        k = mtf_pixel(shape, *params)
    elif kind == 'jitter':
        k = convolvable.mtf_jitter(shape, *params)
    else:
        k = convolvable.mtf_smear(shape, *params)
    return torch.from_numpy(np.ascontiguousarray(k, dtype=np.float64)).to(torch.device('cuda', index))


def device_kernel(kind, shape, *params):
    return _device_kernel(kind, tuple(int(v) for v in shape), tuple(float(p) for p in params),
                          torch.cuda.current_device())


def fft_filter(t, kernel, normalize, precision=None):
    """|ifft2(fft2(t) * kernel)| (times sum(t) / sum(out) when normalize) of a
    device image; the output has the image's dtype."""
    lib = _lib.load()
    ctx = _lib.context()
    rows, cols = t.shape
    out = torch.empty_like(t)
    prec = _lib.precision_code(precision or 'fp32')
    _lib.check(lib.mftx_fft_filter(ctx, ctypes.c_void_p(t.data_ptr()), _lib.dtype_code(t), rows, cols, cols,
                                   ctypes.c_void_p(kernel.data_ptr()), _lib.dtype_code(kernel), kernel.stride(0),
                                   1 if normalize else 0, ctypes.c_void_p(out.data_ptr()), _lib.dtype_code(out),
                                   cols, prec), ctx, 'mftx_fft_filter')
    return out


def mtf_pixel(shape, oversample):
    """The pixel MTF as detector.py:216-222 builds it: a (cols, rows) array
    outer(sinc(fftfreq(cols) * os), sinc(fftfreq(rows) * os)), which meets
    the (rows, cols) spectrum by broadcasting."""
    x = np.fft.fftfreq(shape[1])
    y = np.fft.fftfreq(shape[0])
    return np.dot(np.sinc(x * oversample)[:, np.newaxis], np.sinc(y * oversample)[np.newaxis, :])


def _broadcast_check(shape, kshape):
    # numpy broadcasting of the (rows, cols) spectrum with the (cols, rows)
    # kernel: it fails unless the image is square
    if np.broadcast_shapes(tuple(shape), tuple(kshape)) != tuple(shape):
        raise ValueError(f'operands could not be broadcast together with shapes {tuple(shape)} {tuple(kshape)}')


def pixel(img, oversample=1, *, precision=None):
    """Apply the aperture effect of a square pixel (detector.py:167-225)."""
    t, host = real_image(img)
    rows, cols = t.shape
    _broadcast_check((rows, cols), (cols, rows))
    k = device_kernel('pixel', (rows, cols), oversample)
    return finish(fft_filter(t, k, False, precision), host)


def pixelate(img, oversample, *, precision=None):
    """Pixel MTF, then rescale to native sampling (detector.py:228-251)."""
    from .util import rescale
    t, host = real_image(img)
    rows, cols = t.shape
    _broadcast_check((rows, cols), (cols, rows))
    p = fft_filter(t, device_kernel('pixel', (rows, cols), oversample), False, precision)
    return finish(rescale(p, 1 / oversample, order=3, mode='nearest', unitary=True), host)
