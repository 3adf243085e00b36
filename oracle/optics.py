"""Field / Wavefront / Plane bookkeeping and propagate_dft (oracle only).

A complex128 restatement of the reference's host logic, used as the checker
for the GPU path.  Each function cites the reference lines it follows.

Extents are inclusive integer boxes (rmin, rmax, cmin, cmax) relative to the
centre of the parent grid (extent.py:5-40).
"""
from itertools import combinations
import sys

import numpy as np

from . import mft

# --------------------------------------------------------------------------
# extents (extent.py)
# --------------------------------------------------------------------------


def array_extent(shape, shift):
    """extent.py:25-31 (a shape with fewer than 2 dims is treated as 1x1)."""
    if len(shape) < 2:
        shape = (1, 1)
    rmin = int(-(shape[0] // 2) + shift[0])
    cmin = int(-(shape[1] // 2) + shift[1])
    return rmin, int(rmin + shape[0] - 1), cmin, int(cmin + shape[1] - 1)


def array_center(ext):
    """extent.py:56-59."""
    rmin, rmax, cmin, cmax = ext
    return rmin + (rmax - rmin + 1) // 2, cmin + (cmax - cmin + 1) // 2


def intersect(a, b):
    """Inclusive overlap test, extent.py:75-77."""
    return a[0] <= b[1] and a[1] >= b[0] and a[2] <= b[3] and a[3] >= b[2]


def intersection_extent(a, b):
    """extent.py:93-99."""
    return max(a[0], b[0]), min(a[1], b[1]), max(a[2], b[2]), min(a[3], b[3])


def intersection_shape(a, b):
    """extent.py:116-124."""
    rmin, rmax, cmin, cmax = intersection_extent(a, b)
    nr, nc = rmax - rmin + 1, cmax - cmin + 1
    return () if (nr <= 0 or nc <= 0) else (nr, nc)


def intersection_slices(a, b):
    """extent.py:140-150."""
    rmin, rmax, cmin, cmax = intersection_extent(a, b)
    return ((slice(rmin - a[0], rmax - a[0] + 1), slice(cmin - a[2], cmax - a[2] + 1)),
            (slice(rmin - b[0], rmax - b[0] + 1), slice(cmin - b[2], cmax - b[2] + 1)))


def intersection_shift(a, b):
    """extent.py:166-169."""
    return array_center(intersection_extent(a, b))


# --------------------------------------------------------------------------
# tilts (plane.py:895-923, 995-1045 first-order closed form)
# --------------------------------------------------------------------------


class Tilt:
    """Angular tilt; the constructor swaps axes (plane.py:900-901)."""

    def __init__(self, x, y):
        self.x = y
        self.y = x

    def __shift__(self, xs=0, ys=0, z=0, **kwargs):
        # plane.py:921-922
        return xs - z * self.x, ys - z * self.y


class DispersiveTilt:
    """First-order trace/dispersion only (plane.py:995-1004,1031-1034,1043-1045)."""

    def __init__(self, trace, dispersion):
        self.trace = np.asarray(trace, dtype=float)
        self.dispersion = np.asarray(dispersion, dtype=float)
        if len(self.trace) != 2 or len(self.dispersion) != 2:
            raise NotImplementedError('oracle restates first-order DispersiveTilt only')

    def __shift__(self, wavelength, xs=0.0, ys=0.0, **kwargs):
        dist = (wavelength - self.dispersion[1]) / self.dispersion[0]
        x = dist / np.sqrt(1 + self.trace[0] ** 2)
        y = np.polyval(self.trace, x)
        return x + xs, y + ys


# --------------------------------------------------------------------------
# Field (field.py)
# --------------------------------------------------------------------------


class Field:
    """field.py:33-56: complex data, integer offset, tilt list, extent."""

    def __init__(self, data, pixelscale=None, offset=None, tilt=None):
        self.data = np.asarray(data, dtype=complex)
        self.pixelscale = pixelscale
        self.offset = offset if offset is not None else [0, 0]
        self.tilt = tilt if tilt else []
        self.extent = array_extent(self.data.shape, self.offset)

    @property
    def shape(self):
        return self.data.shape

    @property
    def size(self):
        return self.data.size

    def __mul__(self, other):
        """field.py:108-116 with _mul_scalar (118-128) and _mul_array (130-147)."""
        tilt = self.tilt + other.tilt
        if self.size == 1 and other.size == 1:
            if np.array_equal(self.offset, other.offset):
                data, offset = self.data * other.data, self.offset
            else:
                data, offset = [], None
        else:
            ad, ao, bd, bo = _mul_broadcast(self.data, self.offset, other.data, other.offset)
            ae = array_extent(ad.shape, ao)
            be = array_extent(bd.shape, bo)
            if intersect(ae, be):
                sa, sb = intersection_slices(ae, be)
                data = ad[sa] * bd[sb]
                offset = intersection_shift(ae, be)
            else:
                data, offset = [], None
        return Field(data=data, offset=offset, tilt=tilt)

    def shift(self, z, wavelength, pixelscale, oversample, indexing='ij'):
        """field.py:177-194."""
        if indexing not in ('xy', 'ij'):
            raise ValueError("Valid values for `indexing` are 'xy' and 'ij'")
        if pixelscale is None:
            raise ValueError('pixelscale must be defined to compute shift')
        x, y = 0, 0
        for t in self.tilt:
            x, y = t.__shift__(xs=x, ys=y, z=z, wavelength=wavelength)
        ps = np.broadcast_to(pixelscale, (2,))
        out = x / ps[0] * oversample, y / ps[1] * oversample
        if indexing == 'ij':
            out = -out[1], out[0]
        return out


def _mul_broadcast(a_data, a_offset, b_data, b_offset):
    """field.py:477-484."""
    if a_data.shape != b_data.shape:
        if a_data.size == 1:
            a_data = np.broadcast_to(a_data, b_data.shape)
            a_offset = b_offset
        if b_data.size == 1:
            b_data = np.broadcast_to(b_data, a_data.shape)
            b_offset = a_offset
    return a_data, a_offset, b_data, b_offset


def boundary(fields):
    """field.py:220-229 (note the 0 initial value of rmax/cmax)."""
    rmin, rmax, cmin, cmax = sys.maxsize, 0, sys.maxsize, 0
    for f in fields:
        a, b, c, d = f.extent
        rmin = a if a < rmin else rmin
        rmax = b if b > rmax else rmax
        cmin = c if c < cmin else cmin
        cmax = d if d > cmax else cmax
    return rmin, rmax, cmin, cmax


def insert_slices(field_shape, field_offset, out_shape):
    """Clipped (out_slice, field_slice) placement of field.py:254-299."""
    if tuple(field_shape) == tuple(out_shape) and np.array_equal(field_offset, [0, 0]):
        return Ellipsis, Ellipsis
    osh = np.asarray(out_shape)
    fsh = np.asarray(field_shape)
    ul = (osh // 2) - (fsh // 2) + np.asarray(field_offset)
    fr0, fr1, fc0, fc1 = 0, int(fsh[0]), 0, int(fsh[1])
    or0, or1 = int(ul[0]), int(ul[0] + fsh[0])
    oc0, oc1 = int(ul[1]), int(ul[1] + fsh[1])
    if or0 < 0:
        fr0 = -or0
        or0 = 0
    if or1 > osh[0]:
        fr1 -= or1 - osh[0]
        or1 = osh[0]
    if oc0 < 0:
        fc0 = -oc0
        oc0 = 0
    if oc1 > osh[1]:
        fc1 -= oc1 - osh[1]
        oc1 = osh[1]
    return (slice(or0, or1), slice(oc0, oc1)), (slice(fr0, fr1), slice(fc0, fc1))


def insert(field, out, intensity=False, weight=1):
    """field.py:231-305."""
    oslc, fslc = insert_slices(field.shape, field.offset, out.shape)
    if intensity:
        out[oslc] += np.abs(field.data[fslc] ** 2) * weight
    else:
        out[oslc] += field.data[fslc] * weight
    return out


def _merge(fields):
    """Coherent sum into the union box (field.py:331-346, 349-386)."""
    if not np.all([f.pixelscale == fields[0].pixelscale for f in fields]):
        raise ValueError("Can't merge: pixelscales must be equal")
    rmin, rmax, cmin, cmax = boundary(fields)
    if rmin == 0 and rmax == 0 and cmin == 0 and cmax == 0:
        shape, slices = (), [Ellipsis] * len(fields)
    else:
        shape = (rmax - rmin + 1, cmax - cmin + 1)
        slices = [(slice(f.extent[0] - rmin, f.extent[1] - rmin + 1),
                   slice(f.extent[2] - cmin, f.extent[3] - cmin + 1)) for f in fields]
    out = np.zeros(shape, dtype=complex)
    for f, s in zip(fields, slices):
        out[s] += f.data
    nrow, ncol = rmax - rmin + 1, cmax - cmin + 1
    return Field(data=out, pixelscale=fields[0].pixelscale,
                 offset=(rmin + nrow // 2, cmin + ncol // 2))


def merge(a, b, enforce_overlap=True):
    """field.py:324-327."""
    if enforce_overlap and not intersect(a.extent, b.extent):
        raise ValueError("Can't merge non-overlapping fields")
    return _merge((a, b))


def group_fields(fields):
    """Partition into transitively-overlapping groups (field.py:447-461).

    Restarts the pairwise scan after every join, exactly like the recursive
    ``_disjoint``; returns lists of fields in the order the reference builds
    them.
    """
    groups = [{'field': [f], 'extent': f.extent} for f in fields]
    restart = True
    while restart:
        restart = False
        for m, n in combinations(range(len(groups)), 2):
            if intersect(groups[m]['extent'], groups[n]['extent']):
                groups[m]['field'].extend(groups[n]['field'])
                groups[m]['extent'] = boundary(groups[m]['field'])
                groups.pop(n)
                restart = True
                break
    return [g['field'] for g in groups]


def reduce(fields):
    """field.py:429-439."""
    return [_merge(g) if len(g) > 1 else g[0] for g in group_fields(fields)]


# --------------------------------------------------------------------------
# Wavefront (wavefront.py)
# --------------------------------------------------------------------------


class Wavefront:
    """wavefront.py:33-56 (a single scalar Field 1+0j at offset [0, 0])."""

    def __init__(self, wavelength, pixelscale=None, focal_length=None, tilt=None, ptype='none'):
        self.wavelength = wavelength
        self.pixelscale = None if pixelscale is None else np.broadcast_to(pixelscale, (2,))
        self.focal_length = focal_length if focal_length else None
        self.ptype = ptype
        self.shape = ()
        if tilt is not None:
            tilt = [Tilt(x=tilt[0], y=tilt[1])]
        self.data = [Field(data=np.array(1, dtype=complex), offset=None, tilt=tilt)]

    @classmethod
    def empty(cls, wavelength, pixelscale=None, focal_length=None, shape=None, ptype='none'):
        w = cls(wavelength, pixelscale=pixelscale, focal_length=focal_length, ptype=ptype)
        w.data = []
        w.shape = () if shape is None else shape
        return w

    @property
    def field(self):
        """wavefront.py:101-112."""
        out = np.zeros(self.shape, dtype=complex)
        for f in self.data:
            out = insert(f, out)
        return out

    @property
    def intensity(self):
        """wavefront.py:114-125."""
        out = np.zeros(self.shape, dtype=float)
        for f in reduce(self.data):
            out = insert(f, out, intensity=True)
        return out

    def insert(self, out, weight=1):
        """wavefront.py:145-165."""
        for f in reduce(self.data):
            out = insert(f, out, intensity=True, weight=weight)
        return out


# --------------------------------------------------------------------------
# planes (plane.py)
# --------------------------------------------------------------------------


def bbox(x, threshold=0):
    """util.py:190-218."""
    x = np.asarray(x) > threshold
    rows, cols = np.any(x, axis=1), np.any(x, axis=0)
    rmin, rmax = np.where(rows)[0][[0, -1]]
    cmin, cmax = np.where(cols)[0][[0, -1]]
    return rmin, rmax, cmin, cmax


def boundary_slice(x, threshold=0):
    """helper.py:51-62 with zero padding."""
    rmin, rmax, cmin, cmax = bbox(x, threshold)
    return np.s_[max(rmin, 0):min(rmax + 1, x.shape[0]), max(cmin, 0):min(cmax + 1, x.shape[1])]


def slice_offset(slc, shape):
    """helper.py:101-123 (Ellipsis -> (0, 0))."""
    if slc is Ellipsis:
        return (0, 0)
    ss = np.array((slc[0].stop - slc[0].start, slc[1].stop - slc[1].start))
    off = np.array((slc[0].start + ss[0] // 2, slc[1].start + ss[1] // 2)) - np.asarray(shape) // 2
    return (0, 0) if np.all(off == 0) else tuple(off)


class Pupil:
    """Pupil.__mul__ -> Plane.__mul__ restated (plane.py:477-516, 761-767).

    ``mask`` may be 2-D (monolithic) or 3-D (one mask per segment);
    ``tilt`` is an optional list with one tilt object per segment.
    """

    def __init__(self, amplitude, opd, pixelscale, focal_length, mask=None, tilt=None):
        self.amplitude = np.asarray(amplitude)
        self.opd = np.asarray(opd)
        self.pixelscale = np.broadcast_to(pixelscale, (2,))
        self.focal_length = focal_length
        self._mask = None if mask is None else np.asarray(mask)
        self.tilt = list(tilt) if tilt else []

    @property
    def mask(self):
        # plane.py:44-48
        return np.copy(self.amplitude).astype(bool) if self._mask is None else self._mask

    @property
    def size(self):
        return 1 if self.mask.ndim in (0, 1, 2) else self.mask.shape[0]

    @property
    def shape(self):
        return self.mask.shape if self.size == 1 else self.mask.shape[1:]

    def slices(self):
        """plane.py:671-705."""
        mk = self.mask
        if mk.ndim < 2:
            return [Ellipsis]
        if mk.ndim == 2:
            return [boundary_slice(mk)]
        return [boundary_slice(m) for m in mk]

    def phasors(self, wavelength):
        """Per-segment (data, offset, tilt) of plane.py:495-510."""
        out = []
        for n, s in enumerate(self.slices()):
            mask = self.mask if self.size == 1 else self.mask[n]
            amp = self.amplitude if self.amplitude.size == 1 else self.amplitude[s] * mask[s]
            opd = self.opd if self.opd.size == 1 else self.opd[s]
            data = amp * np.exp(2 * np.pi * 1j * opd / wavelength)
            out.append((data, slice_offset(s, self.shape), [self.tilt[n]] if self.tilt else []))
        return out

    def __rmul__(self, w):
        out = Wavefront.empty(w.wavelength, pixelscale=self.pixelscale,
                              focal_length=self.focal_length, shape=self.shape, ptype='pupil')
        for field in w.data:
            for data, off, tilt in self.phasors(w.wavelength):
                ph = Field(data=data, pixelscale=self.pixelscale, offset=off, tilt=tilt)
                res = field * ph
                if res.size > 0:
                    out.data.append(res)
        return out


class TiltPlane:
    """_TiltBase.__mul__ (plane.py:853-876): append the tilt to every Field."""

    def __init__(self, tilt):
        self.tilt = tilt

    def __rmul__(self, w):
        for f in w.data:
            f.tilt.append(self.tilt)
        return w


# --------------------------------------------------------------------------
# propagate_dft (propagate.py:147-260)
# --------------------------------------------------------------------------


def dft_alpha(dx, du, wavelength, z, oversample):
    """propagate.py:121-123."""
    return ((dx[0] * du[0]) / (wavelength * z * oversample),
            (dx[1] * du[1]) / (wavelength * z * oversample))


def mask_extent(mask):
    """_mask_shape/_mask_shift (propagate.py:245-260) as an extent."""
    rmin, rmax, cmin, cmax = bbox(mask, 0)
    shp = (rmax - rmin + 1, cmax - cmin + 1)
    shift = (rmin + shp[0] // 2 - mask.shape[0] // 2, cmin + shp[1] // 2 - mask.shape[1] // 2)
    return array_extent(shp, shift)


def plan_fields(fields, wavelength, dx, du, z, shape, prop_shape=None, oversample=2, mask=None):
    """The per-Field geometry of propagate.py:177-240.

    Returns, per Field that reaches the output, a dict with the dft2
    arguments (alpha, shape, shift, offset) and the output Field offset.
    """
    shape = np.broadcast_to(shape, (2,))
    prop_shape = np.asarray(shape) if prop_shape is None else np.broadcast_to(prop_shape, (2,))
    shape_out = shape * oversample
    prop_out = prop_shape * oversample
    if mask is not None:
        mask = np.asarray(mask)
        if np.all(mask.shape != shape_out):
            raise ValueError(f'shape mismatch: mask shape {mask.shape} != output shape {tuple(shape_out)}')
        out_ext = mask_extent(mask)
    else:
        out_ext = array_extent(shape_out, (0, 0))
    du = np.broadcast_to(du, (2,))
    plans = []
    for f in fields:
        sh = f.shift(z=z, wavelength=wavelength, pixelscale=du, oversample=oversample, indexing='ij')
        fix = np.fix(sh)
        sub = np.asarray(sh) - fix
        pext = array_extent(prop_out, fix)
        if intersect(out_ext, pext):
            ishape = intersection_shape(out_ext, pext)
            ishift = intersection_shift(out_ext, pext)
            iext = array_extent(ishape, ishift)
            pshift = np.array(array_center(pext)) - np.array(array_center(iext))
            plans.append(dict(field=f, alpha=dft_alpha(dx, du, wavelength, z, oversample),
                              shape=ishape, shift=pshift + sub, offset=f.offset,
                              out_offset=ishift))
    return plans, tuple(int(v) for v in shape_out)


def propagate_dft(w, pixelscale, shape=None, prop_shape=None, oversample=2, mask=None):
    """propagate.py:147-242 (complex128)."""
    if w.ptype not in ('pupil', 'image'):
        raise TypeError("Wavefront must have ptype 'pupil' or 'image'")
    ptype_out = 'image' if w.ptype == 'pupil' else 'pupil'
    shape = np.asarray(w.shape) if shape is None else shape
    du = np.broadcast_to(pixelscale, (2,))
    plans, shape_out = plan_fields(w.data, w.wavelength, w.pixelscale, du, w.focal_length,
                                   shape, prop_shape, oversample, mask)
    out = Wavefront.empty(w.wavelength, pixelscale=du / oversample, focal_length=w.focal_length,
                          shape=shape_out, ptype=ptype_out)
    for p in plans:
        data = mft.dft2(p['field'].data, alpha=p['alpha'], shape=p['shape'], shift=p['shift'],
                        offset=p['offset'], unitary=True)
        out.data.append(Field(data=data, pixelscale=du / oversample, offset=p['out_offset']))
    return out


def ptt_vector(shape, pixelscale, mask):
    """plane.py:521-561: rows [mask, mask * r * dx_r, -mask * c * dx_c] per
    segment (3 * nseg rows) over the raveled grid; r, c from the centred mesh
    (helper.py:7-17)."""
    pixelscale = np.broadcast_to(pixelscale, (2,))
    nr, nc = shape
    r, c = np.meshgrid(np.arange(nr) - np.floor(nr / 2.0), np.arange(nc) - np.floor(nc / 2.0), indexing='ij')
    unmasked = np.einsum('ij,i->ij', [np.ones(r.size), r.ravel(), -c.ravel()], [1, pixelscale[0], pixelscale[1]])
    mask = np.asarray(mask)
    if mask.ndim == 2:
        return np.einsum('ij,j->ij', unmasked, mask.ravel())
    out = np.empty((mask.shape[0] * 3, r.size))
    for k in range(mask.shape[0]):
        out[3 * k:3 * k + 3] = unmasked * mask[k].ravel()
    return out


def fit_tilt(opd, mask, pixelscale):
    """plane.py:563-611: least-squares piston/tip/tilt per segment; returns
    (opd with the tilt removed, [(x, y) tilt per segment])."""
    opd = np.asarray(opd, dtype=float)
    mask = np.asarray(mask)
    v = ptt_vector(opd.shape, pixelscale, mask)
    if mask.ndim == 2:
        t = np.linalg.lstsq(v.T, opd.ravel(), rcond=None)[0]
        return opd - np.einsum('ij,i->j', v[1:3], t[1:3]).reshape(opd.shape), [(t[1], t[2])]
    tilts, parts = [], []
    for k in range(mask.shape[0]):
        t = np.linalg.lstsq(v[3 * k:3 * k + 3].T, opd.ravel(), rcond=None)[0]
        seg = np.einsum('ij,i->j', v[3 * k + 1:3 * k + 3], t[1:3]).reshape(opd.shape)
        parts.append((opd - seg) * mask[k])
        tilts.append((t[1], t[2]))
    return np.sum(parts, axis=0), tilts


# --------------------------------------------------------------------------
# FFT propagation (propagate.py:9-135)
# --------------------------------------------------------------------------


def pad(array, shape):
    """util.py:31-90 for 2-D arrays: centred zero pad, or a centred crop along
    an axis where the array is the larger."""
    array = np.asarray(array)
    out = np.zeros((shape[0], shape[1]), dtype=array.dtype)
    src, dst = [], []
    for a, b in zip(array.shape, shape):
        if b - a <= 0:
            lo = (a - b) // 2
            src.append(slice(lo, lo + b))
            dst.append(slice(0, b))
        else:
            lo = (b - a) // 2
            src.append(slice(0, a))
            dst.append(slice(lo, lo + a))
    out[tuple(dst)] = array[tuple(src)]
    return out


def fft_shape(dx, du, z, wavelength, oversample):
    """propagate.py:126-131: the FFT grid is round(1/alpha) per axis and the
    propagation wavelength is recomputed for that integer grid."""
    alpha = np.asarray(dft_alpha(dx, du, wavelength, z, oversample))
    n = np.round(np.reciprocal(alpha)).astype(int)
    return n, np.min((n / oversample * dx * du) / z)


def scratch_shape(wavelength, dx, du, z, oversample):
    """propagate.py:87-118."""
    n, _ = fft_shape(np.broadcast_to(dx, (2,)), np.broadcast_to(du, (2,)), z, np.max(wavelength), oversample)
    return tuple(n)


def fft2_centred(x):
    """propagate.py:134-135: ifftshift(fft2(fftshift(x)), norm='ortho')."""
    return np.fft.ifftshift(np.fft.fft2(np.fft.fftshift(x), norm='ortho'))


def propagate_fft(w, pixelscale, shape=None, oversample=2, scratch=None):
    """propagate.py:9-84 (complex128).  Returns the propagated Wavefront; with
    ``scratch`` the padded input is left in scratch[:N0, :N1] as in the
    reference."""
    if any(f.tilt for f in w.data):
        raise NotImplementedError('propagate_fft does not support Wavefronts with fitted tilt. '
                                  'Use propagate_dft instead.')
    if w.ptype not in ('pupil', 'image'):
        raise TypeError("Wavefront must have ptype 'pupil' or 'image'")
    ptype_out = 'image' if w.ptype == 'pupil' else 'pupil'
    pixelscale = np.broadcast_to(pixelscale, (2,))
    n, prop_wl = fft_shape(w.pixelscale, pixelscale, w.focal_length, w.wavelength, oversample)
    if shape is None:
        shape_out = tuple(int(v) for v in n)
    else:
        shape = tuple(np.broadcast_to(shape, (2,)))
        if np.any(shape > n / oversample):
            raise ValueError(f'requested shape {tuple(shape)} is larger in at least one dimension than '
                             f'maximum propagation shape {tuple(n // oversample)}')
        shape_out = (shape[0] * oversample, shape[1] * oversample)
    out = Wavefront.empty(prop_wl, pixelscale=pixelscale / oversample, focal_length=w.focal_length,
                          shape=shape_out, ptype=ptype_out)
    if scratch is not None:
        if not all(np.asarray(scratch.shape) > n):
            raise ValueError(f'scratch must have shape greater than or equal to {tuple(n)}')
        view = scratch[0:n[0], 0:n[1]]
        view[...] = 0
        for f in w.data:
            view[...] = insert(f, view)
        data = fft2_centred(view)
    else:
        data = fft2_centred(pad(w.field, n))
    out.data.append(Field(data=data, pixelscale=pixelscale / oversample))
    return out


def broadband_image(pupil, wavelengths, weights, pixelscale, shape, oversample=2,
                    prop_shape=None, mask=None, tilts=()):
    """The reference's user loop (docs/user/diffraction.rst:148-160):

    for wl, wt: w = Wavefront(wl) * pupil [* tilt...]; w = propagate_dft(...);
    img = w.insert(img, weight=wt)
    """
    img = None
    for wl, wt in zip(wavelengths, weights):
        w = Wavefront(wl)
        w = pupil.__rmul__(w)
        for t in tilts:
            w = TiltPlane(t).__rmul__(w)
        w = propagate_dft(w, pixelscale, shape=shape, prop_shape=prop_shape,
                          oversample=oversample, mask=mask)
        if img is None:
            img = np.zeros(w.shape)
        img = w.insert(img, weight=wt)
    return img
