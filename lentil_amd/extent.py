"""Inclusive integer extents (rmin, rmax, cmin, cmax) (lentil extent.py:5-169).

An extent is measured from the centre of its parent grid; an array of shape
(nr, nc) shifted by (r, c) covers rows -(nr//2)+r .. -(nr//2)+r+nr-1.
"""
import numpy as np


def array_extent(shape, shift, parent_shape=None):
    """extent.py:5-40; fewer than two dimensions count as (1, 1)."""
    if len(shape) < 2:
        shape = (1, 1)
    rmin = int(-(shape[0] // 2) + shift[0])
    cmin = int(-(shape[1] // 2) + shift[1])
    rmax = int(rmin + shape[0] - 1)
    cmax = int(cmin + shape[1] - 1)
    if parent_shape is not None:
        pr, pc = (np.asarray(parent_shape) // 2)
        rmin, rmax, cmin, cmax = rmin + pr, rmax + pr, cmin + pc, cmax + pc
    return rmin, rmax, cmin, cmax


def array_center(extent):
    """extent.py:43-59."""
    rmin, rmax, cmin, cmax = extent
    return rmin + (rmax - rmin + 1) // 2, cmin + (cmax - cmin + 1) // 2


def intersect(a, b):
    """True when two inclusive extents share at least one pixel (extent.py:62-77)."""
    return a[0] <= b[1] and a[1] >= b[0] and a[2] <= b[3] and a[3] >= b[2]


def intersection_extent(a, b):
    """extent.py:80-99."""
    return max(a[0], b[0]), min(a[1], b[1]), max(a[2], b[2]), min(a[3], b[3])


def intersection_shape(a, b):
    """extent.py:102-124; () when the extents do not overlap."""
    rmin, rmax, cmin, cmax = intersection_extent(a, b)
    nr, nc = rmax - rmin + 1, cmax - cmin + 1
    return () if nr <= 0 or nc <= 0 else (nr, nc)


def intersection_slices(a, b):
    """extent.py:127-150."""
    rmin, rmax, cmin, cmax = intersection_extent(a, b)
    return ((slice(rmin - a[0], rmax - a[0] + 1), slice(cmin - a[2], cmax - a[2] + 1)),
            (slice(rmin - b[0], rmax - b[0] + 1), slice(cmin - b[2], cmax - b[2] + 1)))


def intersection_shift(a, b):
    """extent.py:153-169."""
    return array_center(intersection_extent(a, b))
