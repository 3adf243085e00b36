"""Optical planes (lentil plane.py), the phasor half of the path.

``Plane.__mul__`` turns each segment slice into a device phasor Field
amp*exp(2*pi*i*opd/lambda) through ``mftx_phasor`` (plane.py:495-514).
The plane hooks ``__amp__``/``__opd__``/``__mask__`` and freeze/thaw behave
as in the reference; a frozen plane also keeps its slices and the device
copies of its per-slice amplitude and OPD.
"""
import copy
import ctypes

import numpy as np
import scipy.integrate
import scipy.optimize
import torch

from . import _lib, helper
from .ptype import ptype as as_ptype, none as PT_NONE, pupil as PT_PUPIL, image as PT_IMAGE, tilt as PT_TILT, \
    transform as PT_TRANSFORM
from ._arrays import to_dev
from .field import Field
from .util import boundary


class _PlaneBase:
    """Low-level plane interface (plane.py:19-269)."""

    def __new__(cls, *args, **kwargs):
        self = super().__new__(cls)
        self._amplitude = np.array(1)
        self._opd = np.array(0)
        self._mask = None
        self._pixelscale = None
        self._ptype = as_ptype(None)
        self._diameter = None
        self.tilt = []
        self._frozen = False
        self._cache = None
        self.__freeze_attrs__ = ['amplitude', 'opd']
        return self

    def __repr__(self):
        return f'{self.__class__.__name__}()'

    def __amp__(self):
        return self._amplitude

    def __mask__(self):
        if self._mask is None:
            return np.copy(self.amplitude).astype(bool)
        return self._mask

    def __opd__(self):
        return self._opd

    def __mul__(self, wavefront):
        raise NotImplementedError

    def __rmul__(self, other):
        return self.__mul__(other)

    @property
    def ptype(self):
        return self._ptype

    @ptype.setter
    def ptype(self, value):
        self._ptype = as_ptype(value)

    @property
    def amplitude(self):
        return self._amplitude if self.frozen else self.__amp__()

    @amplitude.setter
    def amplitude(self, value):
        if self.frozen:
            raise RuntimeError("Can't set amplitude while Plane is frozen")
        self._amplitude = np.asarray(value)

    @property
    def opd(self):
        return self._opd if self.frozen else self.__opd__()

    @opd.setter
    def opd(self, value):
        if self.frozen:
            raise RuntimeError("Can't set opd while Plane is frozen")
        self._opd = np.asarray(value)

    @property
    def mask(self):
        # A frozen plane with no mask of its own and no __mask__ override has
        # the mask of its frozen amplitude (plane.py:42-46), so it is derived
        # once per freeze instead of on every read.
        if self._frozen and self._mask is None and type(self).__mask__ is _PlaneBase.__mask__:
            m = self.__dict__.get('_derived_mask')
            if m is None:
                m = self._derived_mask = self.__mask__()
            return m
        return self.__mask__()

    @property
    def global_mask(self):
        return self.mask if self.size < 2 else np.sum(self.mask, axis=0)

    @property
    def pixelscale(self):
        return self._pixelscale

    @property
    def diameter(self):
        """Outer diameter from the mask boundary unless given (plane.py:136-170)."""
        if self._diameter is None:
            rmin, rmax, cmin, cmax = boundary(self.global_mask)
            return np.max(np.max((rmax - rmin, cmax - cmin)) * np.asarray(self.pixelscale))
        return self._diameter

    @property
    def shape(self):
        return self.mask.shape if self.size == 1 else (self.mask.shape[1], self.mask.shape[2])

    @property
    def size(self):
        return 1 if self.mask.ndim in (0, 1, 2) else self.mask.shape[0]

    @property
    def frozen(self):
        return self._frozen

    def freeze(self, inplace=True):
        """Cache ``__freeze_attrs__`` (plane.py:218-255)."""
        if self.frozen:
            raise RuntimeError('Plane is already frozen. Call thaw() to unfreeze.')
        obj = self if inplace else copy.deepcopy(self)
        for attr in obj.__freeze_attrs__:
            setattr(obj, f'_{attr}', getattr(obj, attr))
        obj._frozen = True
        obj._cache = None
        obj._derived_mask = None
        if not inplace:
            return obj

    def thaw(self):
        self._frozen = False
        self._cache = None
        self._derived_mask = None

    def __deepcopy__(self, memo):
        cls = self.__class__
        new = cls.__new__(cls)
        for k, v in self.__dict__.items():
            setattr(new, k, None if k in ('_cache', '_derived_mask') else copy.deepcopy(v, memo))
        return new


def _enforce_ptype(value, valid):
    value = as_ptype(value)
    try:
        ok = value in valid
    except TypeError:
        ok = value == valid
    if not ok:
        raise ValueError('ptype must be one of the following: ' + ', '.join(map(str, valid)))


class Plane(_PlaneBase):
    """A finite geometric plane (plane.py:403-520)."""

    def __init__(self, amplitude=None, opd=None, mask=None, pixelscale=None, diameter=None, ptype=None, **kwargs):
        if 'amp' in kwargs:
            if amplitude is not None:
                raise AttributeError("Got both 'amplitude' and 'amp', which are aliases of one another")
            amplitude = kwargs['amp']
        if amplitude is not None:
            self._amplitude = np.asarray(amplitude)
        if opd is not None:
            self._opd = np.asarray(opd)
        if mask is not None:
            self._mask = np.asarray(mask)
        if pixelscale is not None:
            self._pixelscale = np.broadcast_to(pixelscale, (2,))
        if diameter is not None:
            self._diameter = diameter
        if ptype is not None:
            self._ptype = as_ptype(ptype)

    def _frozen_cache(self, mask):
        """The frozen cache if it was built from ``mask``, else None.  While a
        plane is frozen its amplitude and OPD properties return the stored
        values (plane.py:73-108), but the reference reads the mask on every
        multiply (plane.py:518-520), so the cache is used only for the mask it
        was built from (the same array object, or an equal one: an in-place
        edit of the mask array a frozen plane already holds is not seen until
        ``thaw``)."""
        c = self._cache
        if not self.frozen or c is None:
            return None
        m = c['mask']
        if mask is m or (np.shape(mask) == np.shape(m) and np.array_equal(mask, m)):
            return c
        return None

    @property
    def _slice(self):
        mask = self.mask
        c = self._frozen_cache(mask)
        return c['slices'] if c is not None else _plane_slice(mask)

    def phasor_tiles(self, mask=None):
        """Per-segment (slice, amplitude, opd) with the segment mask applied (plane.py:496-506).

        amplitude/opd are returned as given (NumPy, possibly 0-d scalars);
        ``mask`` is this plane's mask when the caller has already read it.
        """
        mask_all = self.mask if mask is None else mask
        c = self._frozen_cache(mask_all)
        slices = c['slices'] if c is not None else _plane_slice(mask_all)
        amp_all, opd_all = self.amplitude, self.opd
        tiles = []
        single = mask_all.ndim in (0, 1, 2)
        for n, s in enumerate(slices):
            mask = mask_all if single else mask_all[n]
            amp = amp_all if amp_all.size == 1 else amp_all[s] * mask[s]
            opd = opd_all if opd_all.size == 1 else opd_all[s]
            tiles.append((n, s, amp, opd))
        return tiles

    def _device_tiles(self, mask=None):
        """Device copies of per-slice (amp, opd) as float64; cached while
        frozen (see ``_frozen_cache``).  ``mask`` is this plane's mask when
        the caller has already read it."""
        mask = self.mask if mask is None else mask
        c = self._frozen_cache(mask)
        if c is not None:
            return c['tiles']
        out = []
        for n, s, amp, opd in self.phasor_tiles(mask):
            amp = np.asarray(amp)
            opd = np.asarray(opd)
            if amp.size == 1 and opd.size == 1:
                out.append((n, s, amp, opd, None, None))
                continue
            shp = opd.shape if amp.size == 1 else amp.shape
            a = np.broadcast_to(amp, shp)
            o = np.broadcast_to(opd, shp)
            complex_amp = np.iscomplexobj(a)
            ta = to_dev(np.ascontiguousarray(np.abs(a) if complex_amp else a, dtype=np.float64))
            to = to_dev(np.ascontiguousarray(o, dtype=np.float64))
            out.append((n, s, a if complex_amp else None, None, ta, to))
        if self.frozen:
            self._cache = {'mask': mask, 'slices': [t[1] for t in out], 'tiles': out}
        return out

    @property
    def ptt_vector(self):
        """Piston, x-tilt and y-tilt vectors over the raveled plane, three rows
        per segment, each times the segment mask (plane.py:521-561); None for a
        plane without shape."""
        if self.shape == () or self.shape is None:
            return None
        if self.pixelscale is None:
            raise ValueError("can't create ptt_vector with pixelscale = ()")
        pixelscale = np.broadcast_to(self.pixelscale, (2,))
        r, c = helper.mesh(self.shape)
        unmasked = np.einsum('ij,i->ij', [np.ones(r.size), r.ravel(), -c.ravel()], [1, pixelscale[0], pixelscale[1]])
        if self.size == 1:
            return np.einsum('ij,j->ij', unmasked, self.mask.ravel())
        out = np.empty((self.size * 3, np.prod(self.shape)))
        for k in range(self.size):
            out[3 * k:3 * k + 3] = unmasked * self.mask[k].ravel()
        return out

    def fit_tilt(self, inplace=False):
        """Fit and remove tilt from the OPD by least squares, bookkeeping the
        angular tilt in ``tilt`` (plane.py:563-611).

        The normal equations of every segment's piston/tip/tilt fit come from
        one device pass over the OPD and masks (``mftx_ptt_fit``); each 3x3
        system is solved on the host with the minimum-norm least-squares
        solution, and the tilt is removed on the device
        (``mftx_remove_tilt``)."""
        plane = self if inplace else copy.deepcopy(self)
        if plane.shape == () or plane.shape is None:
            return plane
        if plane.pixelscale is None:
            raise ValueError("can't create ptt_vector with pixelscale = ()")
        opd = np.asarray(plane.opd)
        if opd.size == 1:
            return plane
        dx = np.broadcast_to(plane.pixelscale, (2,)).astype(float)
        rows, cols = plane.shape
        nseg = plane.size
        opd_d = to_dev(np.ascontiguousarray(opd, dtype=np.float64))
        mask_d = to_dev(np.ascontiguousarray(np.asarray(plane.mask), dtype=np.float64))
        lib = _lib.load()
        ctx = _lib.context()
        normal = np.zeros((nseg, 9))
        _lib.check(lib.mftx_ptt_fit(ctx, ctypes.c_void_p(opd_d.data_ptr()), _lib.DT_F64, cols,
                                    ctypes.c_void_p(mask_d.data_ptr()), _lib.DT_F64, cols, rows * cols, nseg, rows,
                                    cols, float(dx[0]), float(dx[1]), normal.ctypes.data_as(ctypes.c_void_p)),
                   ctx, 'mftx_ptt_fit')
        t = np.empty((nseg, 3))
        for k, (g00, g01, g02, g11, g12, g22, b0, b1, b2) in enumerate(normal):
            G = np.array([[g00, g01, g02], [g01, g11, g12], [g02, g12, g22]])
            t[k] = np.linalg.lstsq(G, np.array([b0, b1, b2]), rcond=None)[0]
        tilt = np.ascontiguousarray(t[:, 1:3])
        out = torch.empty((rows, cols), dtype=torch.float64, device=opd_d.device)
        _lib.check(lib.mftx_remove_tilt(ctx, ctypes.c_void_p(opd_d.data_ptr()), _lib.DT_F64, cols,
                                        ctypes.c_void_p(mask_d.data_ptr()), _lib.DT_F64, cols, rows * cols, nseg,
                                        1 if nseg > 1 else 0, rows, cols, float(dx[0]), float(dx[1]),
                                        tilt.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(out.data_ptr()), cols),
                   ctx, 'mftx_remove_tilt')
        plane.opd = out.cpu().numpy()
        plane.tilt.extend([Tilt(x=t[k, 1], y=t[k, 2]) for k in range(nseg)])
        return plane

    def __mul__(self, wavefront):
        if not _can_mul_ptype(wavefront.ptype, self.ptype):
            raise TypeError(f"can't multiply Wavefront with ptype '{wavefront.ptype}' by Plane with ptype "
                            f"'{self.ptype}'")
        from .wavefront import Wavefront
        pixelscale = _mul_pixelscale(self.pixelscale, wavefront.pixelscale)
        mask = self.mask  # read once: shape, size and the slices all follow from it
        plane_shape = mask.shape if mask.ndim in (0, 1, 2) else (mask.shape[1], mask.shape[2])
        shape = wavefront.shape if plane_shape == () else plane_shape
        ptype = _mul_ptype_table[wavefront.ptype][self.ptype]
        out = Wavefront.empty(wavelength=wavefront.wavelength, pixelscale=pixelscale,
                              focal_length=wavefront.focal_length, shape=shape, ptype=ptype)
        phasors = []
        for n, s, camp, sopd, ta, to in self._device_tiles(mask):
            if ta is None:  # scalar amplitude and OPD
                data = camp * np.exp(2 * np.pi * 1j * sopd / wavefront.wavelength)
            else:
                data = device_phasor(ta, to, wavefront.wavelength)
                if camp is not None:  # complex amplitude: restore its phase
                    ph = torch.from_numpy(np.ascontiguousarray(np.exp(1j * np.angle(camp))).astype(np.complex64))
                    data = data * ph.to(data.device)
            phasors.append(Field(data=data, pixelscale=self.pixelscale,
                                 offset=helper.slice_offset(s, plane_shape),
                                 tilt=[self.tilt[n]] if self.tilt else []))
        for fld in wavefront.data:
            for ph in phasors:
                res = fld * ph
                if res.size > 0:
                    out.data.append(res)
        return out


def device_phasor(amp, opd, wavelength):
    """amp*exp(2*pi*i*opd/wavelength) on the device (plane.py:507), complex64."""
    m, n = amp.shape
    out = torch.empty((m, n), dtype=torch.complex64, device=amp.device)
    ctx = _lib.context()
    rc = _lib.load().mftx_phasor(ctx, ctypes.c_void_p(amp.data_ptr()), _lib.dtype_code(amp), amp.stride(0),
                                 ctypes.c_void_p(opd.data_ptr()), _lib.dtype_code(opd), opd.stride(0), m, n,
                                 float(wavelength), ctypes.c_void_p(out.data_ptr()))
    _lib.check(rc, ctx, 'mftx_phasor')
    return out


def _mul_pixelscale(a, b):
    """plane.py:614-629."""
    a = None if a is None else np.broadcast_to(a, (2,))
    b = None if b is None else np.broadcast_to(b, (2,))
    if a is None and b is None:
        return None
    if a is None:
        return b
    if b is None:
        return a
    if all(a == b):
        return a
    raise ValueError(f"can't multiply with inconsistent pixelscales: {a} != {b}")


# Wavefront ptype -> {Plane ptype -> result ptype}; a missing entry forbids the product (plane.py:634-652)
_mul_ptype_table = {
    PT_NONE: {PT_NONE: PT_NONE, PT_PUPIL: PT_PUPIL, PT_IMAGE: PT_IMAGE, PT_TILT: PT_NONE,
               PT_TRANSFORM: PT_NONE},
    PT_PUPIL: {PT_PUPIL: PT_PUPIL, PT_TILT: PT_PUPIL, PT_TRANSFORM: PT_PUPIL},
    PT_IMAGE: {PT_IMAGE: PT_IMAGE, PT_TILT: PT_PUPIL, PT_TRANSFORM: PT_PUPIL},
}


def _can_mul_ptype(wavefront_ptype, plane_ptype):
    return plane_ptype in _mul_ptype_table[wavefront_ptype]


def _plane_slice(mask):
    """Bounding-box slice of each segment mask (plane.py:671-705)."""
    if mask is None:
        return [Ellipsis]
    if mask.ndim < 2:
        return [np.s_[...]]
    if mask.ndim == 2:
        return [helper.boundary_slice(mask)]
    if mask.ndim == 3:
        return [helper.boundary_slice(m) for m in mask]
    raise ValueError('mask has invalid dimensions')


class Pupil(Plane):
    """A pupil plane; its focal length becomes the wavefront's (plane.py:708-771)."""

    def __new__(cls, *args, **kwargs):
        self = super().__new__(cls, *args, **kwargs)
        self._focal_length = None
        self._ptype = PT_PUPIL
        return self

    def __init__(self, amplitude=None, opd=None, mask=None, pixelscale=None, focal_length=None, diameter=None,
                 **kwargs):
        super().__init__(amplitude=amplitude, opd=opd, mask=mask, pixelscale=pixelscale, diameter=diameter,
                         ptype=PT_PUPIL, **kwargs)
        if focal_length is not None:
            self._focal_length = focal_length

    def __mul__(self, wavefront):
        wavefront = super().__mul__(wavefront)
        wavefront.focal_length = self.focal_length
        return wavefront

    @property
    def focal_length(self):
        return self._focal_length


class Image(Plane):
    """An image plane (plane.py:774-820)."""

    def __new__(cls, *args, **kwargs):
        self = super().__new__(cls, *args, **kwargs)
        self._ptype = PT_IMAGE
        return self

    def __init__(self, amplitude=None, opd=None, mask=None, pixelscale=None, **kwargs):
        super().__init__(amplitude=amplitude, opd=opd, mask=mask, pixelscale=pixelscale, ptype=PT_IMAGE, **kwargs)

    def __mul__(self, wavefront):
        wavefront = super().__mul__(wavefront)
        wavefront.ptype = PT_IMAGE
        return wavefront


class _TiltBase(Plane):
    """Planes that append themselves to every Field's tilt list (plane.py:823-881)."""

    def __new__(cls, *args, **kwargs):
        self = super().__new__(cls, *args, **kwargs)
        self._ptype = kwargs['ptype'] if 'ptype' in kwargs else PT_TILT
        return self

    def __mul__(self, wavefront):
        wavefront = super().__mul__(wavefront)
        for f in wavefront.data:
            f.tilt.append(self)
        return wavefront

    def __shift__(self, wavelength, x0, y0, **kwargs):
        raise NotImplementedError


class Tilt(_TiltBase):
    """Angular tilt; x tilt is about the y axis and vice versa (plane.py:884-923)."""

    def __init__(self, x, y, **kwargs):
        super().__init__(**kwargs)
        self.x = y
        self.y = x

    def __shift__(self, xs=0, ys=0, z=0, **kwargs):
        return xs - z * self.x, ys - z * self.y


class DispersiveTilt(_TiltBase):
    """Spectral dispersion seen as a wavelength-dependent tilt (plane.py:926-1096).

    First-order trace and dispersion use the closed forms; higher orders are
    solved numerically with the same least-squares/arc-length formulation.
    """

    def __init__(self, trace=None, dispersion=None, **kwargs):
        super().__init__(**kwargs)
        if trace is not None:
            self.trace = np.asarray(trace)
        if dispersion is not None:
            self.dispersion = np.asarray(dispersion)

    def __shift__(self, wavelength, xs=0., ys=0., **kwargs):
        x, y = self._pos(self._dist(wavelength))
        return x + xs, y + ys

    def _dist(self, wavelength):
        if len(self.dispersion) - 1 == 1:
            return (wavelength - self.dispersion[1]) / self.dispersion[0]
        return scipy.optimize.leastsq(lambda d: wavelength - np.polyval(self.dispersion, d), x0=0)[0]

    def _pos(self, dist):
        if len(self.trace) - 1 == 1:
            x = dist / np.sqrt(1 + self.trace[0] ** 2)
        else:
            dtrace = np.polyder(self.trace)
            arc = lambda b: scipy.integrate.quad(lambda t: np.sqrt(1 + np.polyval(dtrace, t) ** 2), 0, b)[0]
            x = scipy.optimize.leastsq(lambda b: dist - arc(b), x0=0)[0]
        return x, np.polyval(self.trace, x)
