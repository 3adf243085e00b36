"""lentil_amd — MI355X-native Fraunhofer/MFT path with lentil's Python API.

Drop-in surface of lentil v0.8.8 for the propagation hot path
(lentil/__init__.py:7-79): ``fourier.dft2/idft2``, ``propagate_dft``,
``propagate_fft``/``scratch_shape``, the detector-grid operators
``detector.pixel/pixelate``, ``jitter``, ``smear``, ``rebin``, ``rescale``,
``Wavefront``, ``Field`` and ``field.insert/merge/reduce``, the planes
``Plane/Pupil/Image/Tilt/DispersiveTilt`` and the ptype tokens, plus the
pupil-building helpers.  The array work runs in libmftx (HIP, gfx950).

New beyond lentil: ``PSFPlan`` (batched broadband/segmented PSFs in grouped
launches), ``set_precision`` ('fp32' | 'bf16x3') and ``autograd`` (adjoint
MFT gradients for phase retrieval).
"""
__version__ = '0.8.8+mi355x.1'

from .ptype import ptype, none, pupil, image, tilt, transform  # noqa: F401

from . import extent, field, fourier, helper, util  # noqa: F401
from .fourier import set_precision, get_precision  # noqa: F401
from .field import Field  # noqa: F401
from .plane import Plane, Pupil, Image, Tilt, DispersiveTilt  # noqa: F401
from .propagate import propagate_dft, propagate_fft, scratch_shape  # noqa: F401
from .wavefront import Wavefront  # noqa: F401
from .shape import circle, hexagon, rectangle, spider  # noqa: F401
from .segmented import hex_segments  # noqa: F401
from .util import (boundary, centroid, pad, normalize_power, sanitize_shape, sanitize_bandpass,  # noqa: F401
                   rebin, rescale)
from . import detector  # noqa: F401
from .convolvable import jitter, smear  # noqa: F401
from .zernike import zernike, zernike_compose, zernike_basis, zernike_coordinates  # noqa: F401
from .batched import PSFPlan  # noqa: F401
