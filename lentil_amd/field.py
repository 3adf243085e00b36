"""Two-dimensional sampled complex fields (lentil field.py).

Bookkeeping (extents, offsets, grouping) runs on the host exactly as the
reference does; the array work of ``insert`` and coherent ``merge`` runs on
the HIP device through ``mftx_insert``.  Field data produced by the device
path is a complex64 torch tensor; data given by the caller is kept as the
reference keeps it (a complex NumPy array) until a device op needs it.
"""
import ctypes
import sys
from itertools import combinations

import numpy as np
import torch

from . import _lib, extent as _ext
from ._arrays import is_dev, shape_of, size_of, to_dev, to_host, row_stride, write_back


class Field:
    """field.py:7-78: data, pixelscale, integer offset, tilt list, extent."""

    __slots__ = ('data', 'offset', 'tilt', 'pixelscale', 'extent')

    def __init__(self, data, pixelscale=None, offset=None, tilt=None):
        if is_dev(data):
            self.data = data if data.is_complex() else data.to(torch.complex64)
        else:
            self.data = np.asarray(data, dtype=complex)
        self.pixelscale = pixelscale
        self.offset = offset if offset is not None else [0, 0]
        self.tilt = tilt if tilt else []
        self.extent = _ext.array_extent(self.shape, self.offset)

    @property
    def shape(self):
        return shape_of(self.data)

    @property
    def size(self):
        return size_of(self.data)

    def __mul__(self, other):
        """Product over the intersection of extents (field.py:80-147)."""
        tilt = self.tilt + other.tilt
        if self.size == 1 and other.size == 1:
            if np.array_equal(self.offset, other.offset):
                data, offset = _scalar_product(self.data, other.data), self.offset
            else:
                data, offset = [], None
        else:
            ad, ao, bd, bo = _mul_broadcast(self.data, self.offset, other.data, other.offset)
            ae = _ext.array_extent(shape_of(ad), ao)
            be = _ext.array_extent(shape_of(bd), bo)
            if _ext.intersect(ae, be):
                sa, sb = _ext.intersection_slices(ae, be)
                data = _array_product(ad[sa], bd[sb])
                offset = _ext.intersection_shift(ae, be)
            else:
                data, offset = [], None
        return Field(data=data, offset=offset, tilt=tilt)

    def shift(self, z, wavelength, pixelscale, oversample, indexing='ij'):
        """Output-plane shift in pixels from the Field's tilts (field.py:149-194)."""
        return tilt_shift(self.tilt, z, wavelength, pixelscale, oversample, indexing)


def tilt_shift(tilts, z, wavelength, pixelscale, oversample, indexing='ij'):
    """Sum of tilt displacements, in output pixels, (r, c) for 'ij' (field.py:177-194)."""
    if indexing not in ('xy', 'ij'):
        raise ValueError("Valid values for `indexing` are 'xy' and 'ij'")
    if pixelscale is None:
        raise ValueError('pixelscale must be defined to compute shift')
    x, y = 0, 0
    for t in tilts:
        x, y = t.__shift__(xs=x, ys=y, z=z, wavelength=wavelength)
    ps = np.broadcast_to(pixelscale, (2,))
    out = x / ps[0] * oversample, y / ps[1] * oversample
    if indexing == 'ij':
        out = -out[1], out[0]
    return out


def _scalar_product(a, b):
    if is_dev(a) or is_dev(b):
        return (to_dev(a).reshape(-1)[0] * to_dev(b).reshape(-1)[0]).reshape(shape_of(a))
    return a * b


def _array_product(a, b):
    """Elementwise product of two equally shaped slices (field.py:142)."""
    if not is_dev(a) and not is_dev(b):
        if a.size == 1 or b.size == 1 or a.ndim < 2:
            return a * b
    # A broadcast scalar times a device array is the array itself (times the scalar).
    if _is_broadcast_one(a):
        return b if is_dev(b) else to_dev(b).to(torch.complex64)
    if _is_broadcast_one(b):
        return a if is_dev(a) else to_dev(a).to(torch.complex64)
    ta = _c64(a)
    tb = _c64(b)
    out = torch.empty(ta.shape, dtype=torch.complex64, device=ta.device)
    ctx = _lib.context()
    rc = _lib.load().mftx_multiply(ctx, ctypes.c_void_p(ta.data_ptr()), ta.stride(0), ctypes.c_void_p(tb.data_ptr()),
                                   tb.stride(0), ctypes.c_void_p(out.data_ptr()), out.stride(0),
                                   out.shape[0], out.shape[1])
    _lib.check(rc, ctx, 'mftx_multiply')
    return out


def _c64(x):
    """Contiguous complex64 device copy of a 2-D slice (host or device)."""
    if is_dev(x):
        t = x if x.dtype == torch.complex64 else x.to(torch.complex64)
    else:
        t = to_dev(np.ascontiguousarray(x)).to(torch.complex64)
    return t.contiguous()


def _is_broadcast_one(x):
    if is_dev(x):
        return False
    x = np.asarray(x)
    if x.ndim and all(s == 0 for s in x.strides):
        v = x.reshape(-1)[0] if x.size else None
        return v == 1
    return False


def _mul_broadcast(a_data, a_offset, b_data, b_offset):
    """A size-1 operand adopts the other's shape and offset (field.py:464-484)."""
    if shape_of(a_data) != shape_of(b_data):
        if size_of(a_data) == 1:
            a_data = _broadcast(a_data, shape_of(b_data))
            a_offset = b_offset
        if size_of(b_data) == 1:
            b_data = _broadcast(b_data, shape_of(a_data))
            b_offset = a_offset
    return a_data, a_offset, b_data, b_offset


def _broadcast(x, shape):
    if is_dev(x):
        return x.reshape(1, 1).expand(shape)
    return np.broadcast_to(x, shape)


def boundary(fields):
    """Bounding extent of several Fields; rmax/cmax start at 0 (field.py:196-229)."""
    return extents_boundary([f.extent for f in fields])


def extents_boundary(extents):
    """field.boundary over bare extents, keeping its 0 initial rmax/cmax."""
    rmin, rmax, cmin, cmax = sys.maxsize, 0, sys.maxsize, 0
    for a, b, c, d in extents:
        rmin = a if a < rmin else rmin
        rmax = b if b > rmax else rmax
        cmin = c if c < cmin else cmin
        cmax = d if d > cmax else cmax
    return rmin, rmax, cmin, cmax


def insert_slices(field_shape, field_offset, out_shape):
    """Clipped placement of a Field in an output array (field.py:254-299).

    Returns (out_r0, out_c0, field_r0, field_c0, h, w); h or w <= 0 means
    the Field misses the output entirely.
    """
    if tuple(field_shape) == tuple(out_shape) and np.array_equal(field_offset, [0, 0]):
        return 0, 0, 0, 0, int(out_shape[0]), int(out_shape[1])
    oh, ow = int(out_shape[0]), int(out_shape[1])
    fh, fw = int(field_shape[0]), int(field_shape[1])
    ur = oh // 2 - fh // 2 + int(field_offset[0])
    uc = ow // 2 - fw // 2 + int(field_offset[1])
    fr0, fr1, fc0, fc1 = 0, fh, 0, fw
    or0, or1, oc0, oc1 = ur, ur + fh, uc, uc + fw
    if or0 < 0:
        fr0, or0 = -or0, 0
    if or1 > oh:
        fr1 -= or1 - oh
        or1 = oh
    if oc0 < 0:
        fc0, oc0 = -oc0, 0
    if oc1 > ow:
        fc1 -= oc1 - ow
        oc1 = ow
    return or0, oc0, fr0, fc0, fr1 - fr0, fc1 - fc0


def _insert_host(field, out, intensity, weight):
    """Size-1 / 0-d bookkeeping fields (the scalar 1+0j wavefront placeholder)."""
    data = to_host(field.data)
    if data.shape == np.shape(out) and np.array_equal(field.offset, [0, 0]):
        oslc, fslc = Ellipsis, Ellipsis
    else:
        or0, oc0, fr0, fc0, h, w = insert_slices(data.shape, field.offset, out.shape)
        h, w = max(h, 0), max(w, 0)
        oslc = (slice(or0, or0 + h), slice(oc0, oc0 + w))
        fslc = (slice(fr0, fr0 + h), slice(fc0, fc0 + w))
    if intensity:
        out[oslc] += np.abs(data[fslc] ** 2) * weight
    else:
        out[oslc] += data[fslc] * weight
    return out


def insert(field, out, intensity=False, weight=1):
    """Accumulate a Field (or its intensity) into ``out`` (field.py:231-305).

    ``out`` may be a NumPy array (returned, updated in place) or a device
    tensor (accumulated on the device).  Intensity is |E|^2 * weight; the
    complex form adds E * weight.
    """
    if len(field.shape) < 2 or len(np.shape(out)) < 2:
        if is_dev(out):
            raise ValueError('scalar fields insert into host arrays only')
        return _insert_host(field, out, intensity, weight)
    or0, oc0, fr0, fc0, h, w = insert_slices(field.shape, field.offset, np.shape(out))
    if h <= 0 or w <= 0:
        return out
    if is_dev(out):
        _device_insert(field.data, fr0, fc0, h, w, out, or0, oc0, intensity, weight)
        return out
    out_arr = np.asarray(out)
    dev_out = to_dev(out_arr)
    _device_insert(field.data, fr0, fc0, h, w, dev_out, or0, oc0, intensity, weight)
    write_back(out_arr, dev_out)
    return out


def _device_insert(data, fr0, fc0, h, w, out, or0, oc0, intensity, weight):
    lib = _lib.load()
    src = data if is_dev(data) and data.dtype == torch.complex64 else to_dev(data).to(torch.complex64)
    src = src.contiguous()
    code = _lib.dtype_code(out)
    if intensity and code not in (_lib.DT_F32, _lib.DT_F64):
        raise TypeError(f'cannot accumulate intensity into dtype {out.dtype}')
    if not intensity and code not in (_lib.DT_C64, _lib.DT_C128):
        raise TypeError(f'cannot accumulate a complex field into dtype {out.dtype}')
    ctx = _lib.context()
    rc = lib.mftx_insert(ctx, ctypes.c_void_p(src.data_ptr()), src.stride(0), fr0, fc0, h, w,
                         ctypes.c_void_p(out.data_ptr()), code, row_stride(out), or0, oc0,
                         1 if intensity else 0, float(np.real(weight)))
    _lib.check(rc, ctx, 'mftx_insert')


def merge(a, b, enforce_overlap=True):
    """Coherently merge two Fields (field.py:308-327)."""
    if enforce_overlap and not overlap((a, b)):
        raise ValueError("Can't merge non-overlapping fields")
    return _merge((a, b))


def _merge(fields):
    """Sum Fields into their union box (field.py:331-346, 349-386)."""
    if not np.all([f.pixelscale == fields[0].pixelscale for f in fields]):
        raise ValueError("Can't merge: pixelscales must be equal")
    rmin, rmax, cmin, cmax = boundary(fields)
    nrow, ncol = rmax - rmin + 1, cmax - cmin + 1
    offset = (rmin + nrow // 2, cmin + ncol // 2)
    if (rmin == 0 and rmax == 0 and cmin == 0 and cmax == 0) or any(len(f.shape) < 2 for f in fields):
        # scalar placeholders: host bookkeeping
        out = np.zeros(() if rmin == rmax == cmin == cmax == 0 else (nrow, ncol), dtype=complex)
        for f in fields:
            d = to_host(f.data)
            if out.ndim == 0:
                out = out + d
            else:
                out[f.extent[0] - rmin:f.extent[1] - rmin + 1, f.extent[2] - cmin:f.extent[3] - cmin + 1] += d
        return Field(data=out, pixelscale=fields[0].pixelscale, offset=offset)
    acc = torch.zeros((nrow, ncol), dtype=torch.complex64, device=_lib.device())
    for f in fields:
        h, w = f.shape
        _device_insert(f.data, 0, 0, h, w, acc, f.extent[0] - rmin, f.extent[2] - cmin, False, 1.0)
    return Field(data=acc, pixelscale=fields[0].pixelscale, offset=offset)


def overlap(fields):
    """field.py:389-410."""
    if len(fields) == 2:
        return _ext.intersect(fields[0].extent, fields[1].extent)
    return len(_reduce(fields)) <= 1


def reduce(fields):
    """Merge transitively overlapping Fields; disjoint ones pass through (field.py:413-439)."""
    return [_merge(g['field']) if len(g['field']) > 1 else g['field'][0] for g in _reduce(fields)]


def _reduce(fields):
    """field.py:442-448."""
    return _disjoint([{'field': [f], 'extent': f.extent} for f in fields])


def _disjoint(groups):
    """Join the first overlapping pair and rescan, until no pair overlaps (field.py:451-461)."""
    while True:
        for m, n in combinations(range(len(groups)), 2):
            if _ext.intersect(groups[m]['extent'], groups[n]['extent']):
                groups[m]['field'].extend(groups[n]['field'])
                groups[m]['extent'] = boundary(groups[m]['field'])
                groups.pop(n)
                break
        else:
            return groups
