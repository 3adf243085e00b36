"""Field-tile geometry helpers (lentil helper.py:7-123)."""
import numpy as np

from .util import boundary


def mesh(shape, shift=(0, 0), angle=0):
    """Centred (row, col) coordinate grids, optionally rotated (helper.py:7-17)."""
    nr, nc = shape[0], shape[1]
    rr, cc = np.meshgrid(np.arange(nr) - np.floor(nr / 2.0) - shift[0],
                         np.arange(nc) - np.floor(nc / 2.0) - shift[1], indexing='ij')
    a = np.deg2rad(angle)
    return rr * np.cos(a) + cc * np.sin(a), rr * -np.sin(a) + cc * np.cos(a)


def boundary_slice(x, threshold=0, pad=(0, 0)):
    """Bounding box of x > threshold as a pair of slices (helper.py:27-62)."""
    pad = np.asarray(pad)
    if pad.shape == ():
        pad = np.append(pad, pad)
    rmin, rmax, cmin, cmax = boundary(x, threshold)
    return np.s_[max(rmin - pad[0], 0):min(rmax + pad[0] + 1, x.shape[0]),
                 max(cmin - pad[1], 0):min(cmax + pad[1] + 1, x.shape[1])]


def slice_offset(slc, shape):
    """Offset of a slice's centre from the centre of its parent (helper.py:65-123)."""
    if slc is Ellipsis:
        return (0, 0)
    if Ellipsis in slc:
        if slice(None, None, None) in slc:
            return (0, 0)
        raise ValueError(f"Can't compute offset from slice {slc}")
    size = np.array((slc[0].stop - slc[0].start, slc[1].stop - slc[1].start))
    off = np.array((slc[0].start + size[0] // 2, slc[1].start + size[1] // 2)) - np.asarray(shape) // 2
    return (0, 0) if np.all(off == 0) else tuple(off)
