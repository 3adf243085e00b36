"""Fraunhofer propagation by DFT (lentil propagate.py:147-272).

The per-Field geometry (tilt shift, integer/sub-pixel split, output
clipping) is host bookkeeping shared with the batched planner
(``field_geometry``); each surviving Field is transformed on the device by
``fourier.dft2``.
"""
import numpy as np

from . import extent as _ext, fourier
from .ptype import pupil as PT_PUPIL, image as PT_IMAGE
from ._arrays import to_dev
from .field import Field
from .util import boundary
from .wavefront import Wavefront


def _dft_alpha(dx, du, wavelength, z, oversample):
    """propagate.py:121-123."""
    return ((dx[0] * du[0]) / (wavelength * z * oversample), (dx[1] * du[1]) / (wavelength * z * oversample))


def _propagate_ptype(ptype, method='fraunhofer'):
    """pupil <-> image (propagate.py:263-272)."""
    if method == 'fraunhofer':
        if ptype not in (PT_PUPIL, PT_IMAGE):
            raise TypeError("Wavefront must have ptype 'pupil' or 'image'")
        return PT_IMAGE if ptype == PT_PUPIL else PT_PUPIL
    raise ValueError(f'unknown propagation method {method!r}')


def _mask_shape(x, threshold=0):
    rmin, rmax, cmin, cmax = boundary(x, threshold)
    return rmax - rmin + 1, cmax - cmin + 1


def _mask_shift(x, threshold=0):
    """propagate.py:249-260."""
    rc, cc = x.shape[0] // 2, x.shape[1] // 2
    rmin, rmax, cmin, cmax = boundary(x, threshold)
    shp = (rmax - rmin + 1, cmax - cmin + 1)
    return rmin + shp[0] // 2 - rc, cmin + shp[1] // 2 - cc


def output_extent(shape, oversample, mask):
    """Output grid shape and the extent propagation may write (propagate.py:175-190)."""
    shape = np.broadcast_to(shape, (2,))
    shape_out = shape * oversample
    if mask is not None:
        mask = np.asarray(mask)
        if np.all(mask.shape != shape_out):
            raise ValueError(f'shape mismatch: mask shape {mask.shape} != output shape {tuple(shape_out)}')
        out_ext = _ext.array_extent(_mask_shape(mask), _mask_shift(mask))
    else:
        out_ext = _ext.array_extent(shape_out, shift=(0, 0))
    return shape_out, out_ext


def field_geometry(field, wavelength, dx, du, z, prop_shape_out, out_ext, oversample):
    """dft2 arguments of one Field, or None when it misses the output (propagate.py:204-238).

    Returns dict(alpha, shape, shift, offset, out_offset).
    """
    sh = field.shift(z=z, wavelength=wavelength, pixelscale=du, oversample=oversample, indexing='ij')
    fix = np.fix(sh)
    sub = np.asarray(sh) - fix
    pext = _ext.array_extent(prop_shape_out, fix)
    if not _ext.intersect(out_ext, pext):
        return None
    ishape = _ext.intersection_shape(out_ext, pext)
    ishift = _ext.intersection_shift(out_ext, pext)
    iext = _ext.array_extent(ishape, ishift)
    pshift = np.array(_ext.array_center(pext)) - np.array(_ext.array_center(iext))
    return dict(alpha=_dft_alpha(dx, du, wavelength, z, oversample), shape=ishape, shift=pshift + sub,
                offset=field.offset, out_offset=ishift)


def propagate_dft(wavefront, pixelscale, shape=None, prop_shape=None, oversample=2, mask=None, *,
                  precision=None):
    """Propagate a Wavefront to the far field with the MFT (propagate.py:147-242)."""
    ptype_out = _propagate_ptype(wavefront.ptype, method='fraunhofer')
    shape = np.asarray(wavefront.shape) if shape is None else np.broadcast_to(shape, (2,))
    prop_shape = np.asarray(shape) if prop_shape is None else np.broadcast_to(prop_shape, (2,))
    shape_out, out_ext = output_extent(shape, oversample, mask)
    prop_shape_out = prop_shape * oversample
    dx = wavefront.pixelscale
    du = np.broadcast_to(pixelscale, (2,))
    z = wavefront.focal_length
    out = Wavefront.empty(wavelength=wavefront.wavelength, pixelscale=du / oversample,
                          focal_length=wavefront.focal_length, shape=shape_out, ptype=ptype_out)
    for field in wavefront.data:
        g = field_geometry(field, wavefront.wavelength, dx, du, z, prop_shape_out, out_ext, oversample)
        if g is None:
            continue
        data = fourier.dft2(to_dev(_as_2d(field.data)), alpha=g['alpha'], shape=g['shape'], shift=g['shift'],
                            offset=g['offset'], unitary=True, precision=precision)
        out.data.append(Field(data=data, pixelscale=du / oversample, offset=g['out_offset']))
    return out


def _as_2d(data):
    if len(data.shape) == 2:
        return data
    return data.reshape(1, 1)
