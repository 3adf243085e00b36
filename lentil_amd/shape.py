"""Aperture shape generators used to build pupils (lentil shape.py).

Host-side input generation (not part of the propagation hot path).
"""
import numpy as np

from .helper import mesh


def circle(shape, radius, shift=(0, 0), antialias=True):
    """Disc of ``radius`` pixels with a one-pixel linear edge ramp (shape.py:6-30)."""
    rr, cc = mesh(shape)
    rho = np.hypot(rr - shift[0], cc - shift[1])
    out = np.clip(radius + 0.5 - rho, 0.0, 1.0)
    if not antialias:
        out[out > 0] = 1
    return out


def hexagon(shape, radius, shift=(0, 0), rotate=False, antialias=True):
    """Regular hexagon of circumradius ``radius`` (shape.py:33-75)."""
    shape = np.broadcast_to(shape, (2,))
    r, c = mesh(shape, shift)
    out = np.ones(shape)
    apothem = radius * np.sqrt(3) / 2
    for k in range(6):
        theta = k * np.pi / 3 + (0 if rotate else np.pi / 6)
        d = r * np.sin(theta) + c * np.cos(theta)
        if antialias:
            edge = np.clip(apothem + 0.5 - d, 0.0, 1.0)
        else:
            edge = np.where(d > apothem, 0.0, 1.0)
        out = np.minimum(out, edge)
    return out


def rectangle(shape, width, height, shift=(0, 0), angle=0, antialias=True):
    """Rectangle rotated by ``angle`` degrees (shape.py:78-114)."""
    shape = np.broadcast_to(shape, (2,))
    rr, cc = mesh(shape, shift, angle)
    out = np.minimum(np.minimum(np.ones(shape), np.clip(0.5 + width / 2 - np.abs(cc), 0, 1)),
                     np.clip(0.5 + height / 2 - np.abs(rr), 0, 1))
    if not antialias:
        out[out > 0] = 1
    return out


def spider(shape, width, angle=0, shift=(0, 0), antialias=True):
    """One spider vane from the centre outward (shape.py:117-145)."""
    shape = np.broadcast_to(shape, (2,))
    length = np.sqrt(2) * np.max(shape) / 2
    a = np.deg2rad(angle)
    shift = (shift[0] - length / 2 * np.sin(a), shift[1] + length / 2 * np.cos(a))
    return 1 - rectangle(shape, length, width, shift, angle, antialias)
