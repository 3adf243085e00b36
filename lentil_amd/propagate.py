"""Fraunhofer propagation by DFT and by FFT (lentil propagate.py:9-272).

The per-Field geometry (tilt shift, integer/sub-pixel split, output
clipping) is host bookkeeping shared with the batched planner
(``field_geometry``); each surviving Field is transformed on the device by
``fourier.dft2``.

``propagate_fft``'s zero-padded, shifted FFT is the same transform as an MFT
with alpha = 1/N on the unpadded field: the zero padding drops out of the
sum, and the fftshift/ifftshift pair becomes integer coordinate offsets
(reduced modulo N), so no padded grid is materialised.
"""
import numpy as np
import torch

from . import _lib, extent as _ext, field as _field, fourier
from .ptype import pupil as PT_PUPIL, image as PT_IMAGE
from ._arrays import is_dev, to_dev, write_back
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


def _fft_shape(dx, du, z, wavelength, oversample):
    """FFT grid round(1/alpha) per axis and the propagation wavelength
    recomputed for that integer grid (propagate.py:126-131; the reference
    passes z and the wavelength to _dft_alpha in swapped positions, which
    leaves their product, and so alpha, unchanged)."""
    alpha = np.asarray(_dft_alpha(dx, du, wavelength, z, oversample))
    n = np.round(np.reciprocal(alpha)).astype(int)
    return n, np.min((n / oversample * dx * du) / z)


def scratch_shape(wavelength, dx, du, z, oversample):
    """Scratch shape an FFT propagation needs (propagate.py:87-118)."""
    dx = np.broadcast_to(dx, (2,))
    du = np.broadcast_to(du, (2,))
    n, _ = _fft_shape(dx, du, z, np.max(wavelength), oversample)
    return tuple(n)


def _wrap(v, n):
    """v modulo n, in (-n/2, n/2]."""
    v = int(v) % int(n)
    return v - n if v > n / 2 else v


def centred_fft_geometry(n, m, pad_offset):
    """MFT (offset, shift) along one axis that reproduce
    ifftshift(fft(fftshift(p))) (propagate.py:134-135) on an N-point grid p
    holding an m-point array at pad_offset.

    With h = N // 2, that transform is sum_p p[j] exp(-2 pi i (j + h)(u + h) / N)
    for both parities of N.  An MFT with alpha = 1/N sums over the m-point
    array with input coordinate x = r - m // 2 + offset and output coordinate
    y = (u - h) - shift, so offset = pad_offset + h + m // 2 and shift = -2h.
    Both coordinates are integers, so each is taken modulo N (the phase is
    periodic); for even N both vanish, and for odd N they are -1 and +1
    when the array fills the grid."""
    h = n // 2
    return _wrap(pad_offset + h + m // 2, n), _wrap(-2 * h, n)


def propagate_fft(wavefront, pixelscale, shape=None, oversample=2, scratch=None, *, precision=None):
    """Propagate a Wavefront to the far field with the zero-padded FFT
    (propagate.py:9-84), run on the device as the equivalent MFT.

    The output Field holds the whole FFT grid round(1/alpha); the Wavefront's
    shape is ``shape * oversample`` (default: the FFT grid).  With
    ``scratch`` (an ndarray or device tensor larger than the FFT grid in both
    dimensions) the Fields are inserted into the zeroed scratch[:N0, :N1]
    region, which keeps the padded input on return, as in the reference."""
    if any(f.tilt for f in wavefront.data):
        raise NotImplementedError('propagate_fft does not support Wavefronts with fitted tilt. '
                                  'Use propagate_dft instead.')
    ptype_out = _propagate_ptype(wavefront.ptype, method='fraunhofer')
    pixelscale = np.broadcast_to(pixelscale, (2,))
    n, prop_wavelength = _fft_shape(wavefront.pixelscale, pixelscale, wavefront.focal_length,
                                    wavefront.wavelength, oversample)
    if shape is None:
        shape_out = tuple(int(v) for v in n)
    else:
        shape = tuple(np.broadcast_to(shape, (2,)))
        if np.any(shape > n / oversample):
            raise ValueError(f'requested shape {tuple(shape)} is larger in at least one dimension than '
                             f'maximum propagation shape {tuple(n // oversample)}')
        shape_out = (shape[0] * oversample, shape[1] * oversample)
    out = Wavefront.empty(wavelength=prop_wavelength, pixelscale=pixelscale / oversample,
                          focal_length=wavefront.focal_length, shape=shape_out, ptype=ptype_out)
    n0, n1 = (int(v) for v in n)
    if scratch is not None:
        if not all(np.asarray(scratch.shape) > n):
            raise ValueError(f'scratch must have shape greater than or equal to {tuple(n)}')
        on_device = is_dev(scratch)
        if on_device:
            buf = scratch[0:n0, 0:n1]
            buf.zero_()
        else:
            buf = torch.zeros((n0, n1), dtype=torch.complex128, device=_lib.device())
        for f in wavefront.data:
            _field.insert(f, buf)
        if not on_device:
            write_back(scratch[0:n0, 0:n1], buf)
        src, pad_off = buf, (0, 0)
    else:
        full = wavefront.field_tensor()
        s0, s1 = full.shape
        # lentil.pad: centred pad, or centred crop along an axis that is larger (util.py:31-90)
        a0, a1 = max(s0 - n0, 0) // 2, max(s1 - n1, 0) // 2
        src = full[a0:a0 + min(s0, n0), a1:a1 + min(s1, n1)]
        pad_off = (max(n0 - s0, 0) // 2, max(n1 - s1, 0) // 2)
    (o0, sh0), (o1, sh1) = (centred_fft_geometry(n0, src.shape[0], pad_off[0]),
                            centred_fft_geometry(n1, src.shape[1], pad_off[1]))
    data = fourier.dft2(src, alpha=(1.0 / n0, 1.0 / n1), shape=(n0, n1), shift=(sh0, sh1), offset=(o0, o1),
                        unitary=True, precision=precision)
    out.data.append(Field(data=data, pixelscale=pixelscale / oversample))
    return out


def _as_2d(data):
    if len(data.shape) == 2:
        return data
    return data.reshape(1, 1)
