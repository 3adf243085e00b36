"""Array utilities used on the propagation path (lentil util.py)."""
import numpy as np


def boundary(x, threshold=0):
    """Bounding (rmin, rmax, cmin, cmax) of entries > threshold (util.py:190-218)."""
    x = np.asarray(x) > threshold
    rows = np.flatnonzero(np.any(x, axis=1))
    cols = np.flatnonzero(np.any(x, axis=0))
    return rows[0], rows[-1], cols[0], cols[-1]


def centroid(img):
    """(r, c) intensity centroid (util.py:6-28)."""
    img = np.asarray(img)
    img = img / np.sum(img)
    nr, nc = img.shape
    rr, cc = np.mgrid[0:nr, 0:nc]
    return np.dot(rr.ravel(), img.ravel()), np.dot(cc.ravel(), img.ravel())


def pad(array, shape):
    """Centre-pad or centre-crop a 2-D (or 3-D stacked) array (util.py:31-90)."""
    array = np.asarray(array)
    lead = 1 if array.ndim == 3 else 0
    src_r, src_c = array.shape[lead], array.shape[lead + 1]

    def spans(src, dst):
        if dst - src <= 0:
            a = (src - dst) // 2
            return (a, a + dst), (0, dst)
        b = (dst - src) // 2
        return (0, src), (b, b + src)

    (r0, r1), (R0, R1) = spans(src_r, shape[0])
    (c0, c1), (C0, C1) = spans(src_c, shape[1])
    if array.ndim < 3:
        out = np.zeros((shape[0], shape[1]), dtype=array.dtype)
        out[R0:R1, C0:C1] = array[r0:r1, c0:c1]
    else:
        out = np.zeros((array.shape[0], shape[0], shape[1]), dtype=array.dtype)
        out[:, R0:R1, C0:C1] = array[:, r0:r1, c0:c1]
    return out


def normalize_power(array, power=1):
    """Scale so that sum(|array|^2) == power (util.py:378-411)."""
    array = np.asarray(array)
    return array * np.sqrt(power / np.sum(np.abs(array) ** 2))


def sanitize_shape(shape):
    """util.py:414-421."""
    shape = np.asarray(shape)
    if shape.size == 0:
        return ()
    if shape.shape == ():
        shape = np.append(shape, shape)
    return tuple(shape)


def sanitize_bandpass(vec):
    """util.py:424-428."""
    vec = np.asarray(vec)
    return vec[np.newaxis, ...] if vec.shape == () else vec
