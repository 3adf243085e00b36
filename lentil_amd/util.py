"""Array utilities used on the propagation path (lentil util.py)."""
import numpy as np


def boundary(x, threshold=0):
    """Bounding (rmin, rmax, cmin, cmax) of entries > threshold (util.py:190-218)."""
    x = np.asarray(x) > threshold
    rows = np.flatnonzero(np.any(x, axis=1))
    cols = np.flatnonzero(np.any(x, axis=0))
    return rows[0], rows[-1], cols[0], cols[-1]


def centroid(img):
    """(r, c) intensity centroid (util.py:6-28)."""
    img = np.asarray(img)
    img = img / np.sum(img)
    nr, nc = img.shape
    rr, cc = np.mgrid[0:nr, 0:nc]
    return np.dot(rr.ravel(), img.ravel()), np.dot(cc.ravel(), img.ravel())


def pad(array, shape):
    """Centre-pad or centre-crop a 2-D (or 3-D stacked) array (util.py:31-90)."""
    array = np.asarray(array)
    lead = 1 if array.ndim == 3 else 0
    src_r, src_c = array.shape[lead], array.shape[lead + 1]

    def spans(src, dst):
        if dst - src <= 0:
            a = (src - dst) // 2
            return (a, a + dst), (0, dst)
        b = (dst - src) // 2
        return (0, src), (b, b + src)

    (r0, r1), (R0, R1) = spans(src_r, shape[0])
    (c0, c1), (C0, C1) = spans(src_c, shape[1])
    if array.ndim < 3:
        out = np.zeros((shape[0], shape[1]), dtype=array.dtype)
        out[R0:R1, C0:C1] = array[r0:r1, c0:c1]
    else:
        out = np.zeros((array.shape[0], shape[0], shape[1]), dtype=array.dtype)
        out[:, R0:R1, C0:C1] = array[:, r0:r1, c0:c1]
    return out


def normalize_power(array, power=1):
    """Scale so that sum(|array|^2) == power (util.py:378-411)."""
    array = np.asarray(array)
    return array * np.sqrt(power / np.sum(np.abs(array) ** 2))


def sanitize_shape(shape):
    """util.py:414-421."""
    shape = np.asarray(shape)
    if shape.size == 0:
        return ()
    if shape.shape == ():
        shape = np.append(shape, shape)
    return tuple(shape)


def sanitize_bandpass(vec):
    """util.py:424-428."""
    vec = np.asarray(vec)
    return vec[np.newaxis, ...] if vec.shape == () else vec


def rebin(img, factor):
    """Rebin an image (or a cube of images) by an integer factor (util.py:221-259),
    as block sums on the device.  NumPy in gives NumPy out with the reference's
    dtype; a float32/float64 device tensor gives a device tensor."""
    import ctypes
    import torch
    from . import _lib
    from ._arrays import is_dev, to_dev
    host = not is_dev(img)
    if host:
        a = np.asarray(img)
        if np.iscomplexobj(a):
            raise ValueError('rebin is not defined for complex data')
        ref_dtype = a.dtype if a.ndim == 3 else np.zeros(1, a.dtype).sum().dtype
    elif img.is_complex():
        raise ValueError('rebin is not defined for complex data')
    t = to_dev(img)
    if t.dtype not in (torch.float32, torch.float64):
        t = t.to(torch.float64)
    t = t.contiguous()
    if t.dim() not in (2, 3):
        raise ValueError(f'expected an image or a cube of images, got shape {tuple(t.shape)}')
    factor = int(factor)
    rows, cols = t.shape[-2:]
    if factor <= 0 or rows % factor or cols % factor:
        raise ValueError(f'cannot rebin shape {tuple(t.shape)} by {factor}')
    nimg = t.shape[0] if t.dim() == 3 else 1
    out = torch.empty(tuple(t.shape[:-2]) + (rows // factor, cols // factor), dtype=t.dtype, device=t.device)
    lib = _lib.load()
    ctx = _lib.context()
    _lib.check(lib.mftx_rebin(ctx, ctypes.c_void_p(t.data_ptr()), _lib.dtype_code(t), nimg, rows, cols, cols, factor,
                              ctypes.c_void_p(out.data_ptr()), _lib.dtype_code(out)), ctx, 'mftx_rebin')
    if host:
        return out.cpu().numpy().astype(ref_dtype)
    return out


# scipy.ndimage's boundary mode names and the device's codes (MFTX_MODE_*)
_RESCALE_MODES = {'nearest': 0, 'reflect': 1, 'grid-mirror': 2, 'mirror': 3, 'grid-wrap': 4, 'wrap': 5,
                  'constant': 6, 'grid-constant': 7}


def rescale(img, scale, shape=None, mask=None, order=3, mode='nearest', unitary=True):
    """Rescale an image by spline interpolation (util.py:262-349) on the device.

    Follows scipy.ndimage.map_coordinates as the reference calls it: the
    sample coordinates (i - out/2) / scale + in/2, spline order 0-5 and any of
    scipy's boundary modes (cval 0, as the reference never passes one), the
    mask (nonzero pixels when None) resampled with order 1 and zeroed below
    its dtype's eps, the unitary weight correction, then the product with the
    mask.  A complex image has its real and imaginary parts resampled
    separately and the complex ratio of the sums applied (util.py:339-345); it
    gives a complex128 result.  NumPy in gives NumPy out (the image's float
    dtype, or complex128); a device tensor gives a device tensor.
    """
    import ctypes
    import torch
    from . import _lib
    from ._arrays import is_dev, to_dev
    # scipy.ndimage's own errors for an order or mode it does not have
    if int(order) != order or not 0 <= order <= 5:
        raise RuntimeError('spline order not supported')
    if mode not in _RESCALE_MODES:
        raise RuntimeError('boundary mode not supported')
    host = not is_dev(img)
    t = to_dev(img)
    cplx = t.is_complex()
    if cplx:
        if t.dtype not in (torch.complex64, torch.complex128):
            t = t.to(torch.complex128)
    elif t.dtype not in (torch.float32, torch.float64):
        t = t.to(torch.float64)
    t = t.contiguous()
    if t.dim() != 2:
        raise ValueError(f'expected a 2-D image, got shape {tuple(t.shape)}')
    rows, cols = t.shape
    if shape is None:
        oshape = np.ceil((rows * scale, cols * scale)).astype(int)
    elif np.isscalar(shape):
        oshape = np.ceil((shape * scale, shape * scale)).astype(int)
    else:
        oshape = np.ceil((shape[0] * scale, shape[1] * scale)).astype(int)
    m = None
    if mask is not None:
        m = to_dev(mask)
        if m.is_complex():
            m = m.real
        if m.dtype not in (torch.float32, torch.float64):
            m = m.to(torch.float64)
        m = m.contiguous()
        if tuple(m.shape) != (rows, cols):
            raise ValueError(f'mask shape {tuple(m.shape)} != image shape {(rows, cols)}')
    out = torch.empty((int(oshape[0]), int(oshape[1])), dtype=torch.complex128 if cplx else t.dtype,
                      device=t.device)
    lib = _lib.load()
    ctx = _lib.context()
    _lib.check(lib.mftx_rescale(ctx, ctypes.c_void_p(t.data_ptr()), _lib.dtype_code(t), rows, cols, cols,
                                ctypes.c_void_p(m.data_ptr()) if m is not None else None,
                                _lib.dtype_code(m) if m is not None else 0, cols, float(scale),
                                out.shape[0], out.shape[1], int(order), _RESCALE_MODES[mode], 1 if unitary else 0,
                                ctypes.c_void_p(out.data_ptr()), _lib.dtype_code(out), out.shape[1]),
               ctx, 'mftx_rescale')
    return out.cpu().numpy() if host else out
