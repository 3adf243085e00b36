"""Hexagonal segmented apertures (lentil segmented.py).

Host-side input generation (not part of the propagation hot path).
"""
import numpy as np

from .shape import hexagon

# cube-coordinate unit steps walked around a hex ring (segmented.py:90-91)
_DIRS = ((1, 0, -1), (1, -1, 0), (0, -1, 1), (-1, 0, 1), (-1, 1, 0), (0, 1, -1))


def _ring(radius):
    q, r, s = -radius, radius, 0
    cells = []
    for d in _DIRS:
        for _ in range(radius):
            cells.append((q, r, s))
            q, r, s = q + d[0], r + d[1], s + d[2]
    return cells


def _center_rc(cell, pitch, rotate):
    q, r = cell[0], cell[1]
    if rotate:
        x, y = pitch * (np.sqrt(3) * q + np.sqrt(3) / 2 * r), pitch * 1.5 * r
    else:
        x, y = pitch * 1.5 * q, pitch * (np.sqrt(3) / 2 * q + np.sqrt(3) * r)
    return -y, x


def hex_segments(rings, seg_radius, seg_gap, rotate=False, antialias=True, flatten=False, pad=2, drop=(0,)):
    """Stack of per-segment hexagon masks, centre segment index 0 (segmented.py:8-80)."""
    apothem = seg_radius * np.sqrt(3) / 2
    size = int(np.ceil((rings * 2 + 1) * apothem * 2 + rings * 2 * seg_gap + pad * 2))
    shape = (size, size)
    masks = []
    if 0 not in drop:
        masks.append(hexagon(shape, seg_radius, shift=(0, 0), antialias=antialias, rotate=rotate))
    index = 1
    for ring in range(1, rings + 1):
        for cell in _ring(ring):
            if index not in drop:
                masks.append(hexagon(shape, seg_radius, shift=_center_rc(cell, seg_radius + seg_gap / 2, rotate),
                                     antialias=antialias, rotate=rotate))
            index += 1
    masks = np.asarray(masks)
    return np.sum(masks, axis=0) if flatten else masks
