"""Monochromatic wavefront (lentil wavefront.py)."""
import numpy as np
import torch

from . import _lib, field as _field
from .ptype import ptype as as_ptype, none as PT_NONE, pupil as PT_PUPIL, image as PT_IMAGE
from ._arrays import is_dev, write_back
from .field import Field


class Wavefront:
    """A monochromatic wavefront made of Fields (wavefront.py:12-56)."""

    def __init__(self, wavelength, pixelscale=None, diameter=None, focal_length=None, tilt=None, ptype=None):
        self.focal_length = focal_length if focal_length else None
        self.diameter = diameter
        self.shape = ()
        self._wavelength = wavelength
        self._pixelscale = None if pixelscale is None else np.broadcast_to(pixelscale, (2,))
        self.ptype = as_ptype(ptype)
        if tilt is not None:
            if len(tilt) != 2:
                raise ValueError('tilt must be specified as [rx, ry]')
            from .plane import Tilt
            tilt = [Tilt(x=tilt[0], y=tilt[1])]
        self.data = [Field(data=np.array(1, dtype=complex), offset=None, tilt=tilt)]

    def __mul__(self, plane):
        return plane.__mul__(self)

    def __rmul__(self, other):
        return self.__mul__(other)

    @property
    def wavelength(self):
        return self._wavelength

    @property
    def pixelscale(self):
        return self._pixelscale

    @property
    def ptype(self):
        return self._ptype

    @ptype.setter
    def ptype(self, value):
        if as_ptype(value) not in (PT_NONE, PT_PUPIL, PT_IMAGE):
            raise TypeError(f"invalid ptype '{value}' for Wavefront")
        self._ptype = as_ptype(value)

    @property
    def field(self):
        """Complex field on the wavefront grid (wavefront.py:101-112), complex128 NumPy."""
        if len(self.shape) < 2:
            out = np.zeros(self.shape, dtype=complex)
            for f in self.data:
                out = _field.insert(f, out)
            return out
        return self.field_tensor().cpu().numpy()

    def field_tensor(self, dtype=torch.complex128):
        """Device-resident complex field on the wavefront grid (no host copy)."""
        acc = torch.zeros(tuple(self.shape), dtype=dtype, device=_lib.device())
        for f in self.data:
            _field.insert(f, acc)
        return acc

    @property
    def intensity(self):
        """|field|^2 with overlapping Fields merged first (wavefront.py:114-125), float64 NumPy."""
        if len(self.shape) < 2:
            out = np.zeros(self.shape, dtype=float)
            for f in _field.reduce(self.data):
                out = _field.insert(f, out, intensity=True)
            return out
        return self.intensity_tensor(torch.float64).cpu().numpy()

    def intensity_tensor(self, dtype=torch.float32):
        """Device-resident intensity (no host copy)."""
        acc = torch.zeros(tuple(self.shape), dtype=dtype, device=_lib.device())
        for f in _field.reduce(self.data):
            _field.insert(f, acc, intensity=True)
        return acc

    @classmethod
    def empty(cls, wavelength, pixelscale=None, diameter=None, focal_length=None, tilt=None, shape=None,
              ptype=None):
        """Wavefront with no Fields (wavefront.py:127-142)."""
        w = cls(wavelength=wavelength, pixelscale=pixelscale, diameter=diameter, focal_length=focal_length,
                tilt=tilt, ptype=ptype)
        w.data = []
        w.shape = () if shape is None else shape
        return w

    def insert(self, out, weight=1):
        """Accumulate weight*intensity into ``out`` (wavefront.py:145-165).

        ``out`` may be a NumPy array or a device tensor; with a device tensor
        the accumulation stays on the device.
        """
        fields = _field.reduce(self.data)
        if is_dev(out) or len(np.shape(out)) < 2:
            for f in fields:
                out = _field.insert(f, out, intensity=True, weight=weight)
            return out
        dev = torch.from_numpy(np.ascontiguousarray(out)).to(_lib.device())
        for f in fields:
            _field.insert(f, dev, intensity=True, weight=weight)
        write_back(out, dev)
        return out
