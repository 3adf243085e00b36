"""Detector-grid post-processing (oracle only, test infrastructure).

A restatement of the reference's image operators on the accumulated PSF,
with numpy's FFT and scipy.ndimage.map_coordinates as the reference itself
calls them:

  pixel      detector.py:167-225     |ifft2(fft2(img) * sinc MTF)|
  pixelate   detector.py:228-251     pixel, then rescale(1/oversample)
  jitter     convolvable.py:4-82     Gaussian MTF, weight preserved
  smear      convolvable.py:85-169   directional sinc MTF, weight preserved
  rebin      util.py:221-259         integer block sums
  rescale    util.py:262-349         cubic-spline resampling (+ mask, unitary)

Below them, the explicit algorithm the device kernels follow for rescale's
map_coordinates(order=3, mode='nearest'): edge padding, the cubic B-spline
prefilter as its two-sided impulse response, and the 4x4 spline taps.  These
are checked against scipy in tests/test_oracle_detector.py.
"""
import numpy as np
from scipy.ndimage import map_coordinates


def mtf_pixel(shape, oversample):
    """The pixel MTF kernel exactly as detector.py:216-222 builds it:
    outer(sinc(fftfreq(cols) * os), sinc(fftfreq(rows) * os)), a (cols, rows)
    array, so it broadcasts against the image only when it is square."""
    x = np.fft.fftfreq(shape[1])
    y = np.fft.fftfreq(shape[0])
    return np.dot(np.sinc(x * oversample)[:, np.newaxis], np.sinc(y * oversample)[np.newaxis, :])


def pixel(img, oversample=1):
    """detector.py:167-225."""
    img = np.asarray(img)
    return np.abs(np.fft.ifft2(np.fft.fft2(img) * mtf_pixel(img.shape, oversample)))


def pixelate(img, oversample):
    """detector.py:228-251."""
    return rescale(pixel(img, oversample), 1 / oversample, order=3, mode='nearest', unitary=True)


def mtf_jitter(shape, scale, pixelscale=1, oversample=1):
    """convolvable.py:72-77."""
    x = np.fft.fftfreq(shape[1])
    y = np.fft.fftfreq(shape[0])
    xx, yy = np.meshgrid(x, y)
    rho = np.sqrt(xx ** 2 + yy ** 2)
    return np.exp(-2 * (np.pi * (scale / pixelscale) * oversample * rho) ** 2)


def jitter(img, scale, pixelscale=1, oversample=1):
    """convolvable.py:4-82."""
    img = np.asarray(img)
    out = np.abs(np.fft.ifft2(np.fft.fft2(img) * mtf_jitter(img.shape, scale, pixelscale, oversample)))
    return out * np.sum(img) / np.sum(out)


def mtf_smear(shape, distance, angle_rad, pixelscale=1, oversample=1):
    """convolvable.py:156-166 for a given angle in radians."""
    x = np.fft.fftfreq(shape[1])
    y = np.fft.fftfreq(shape[0])
    xx, yy = np.meshgrid(x, y)
    yy_rot = np.sin(angle_rad) * yy + np.cos(angle_rad) * xx
    return np.sinc(yy_rot * (distance / pixelscale) * oversample)


def smear(img, distance, angle=None, pixelscale=1, oversample=1, rng=None):
    """convolvable.py:85-169 (a random angle is drawn from ``rng`` when angle is None)."""
    img = np.asarray(img)
    if angle is None:
        angle = (rng or np.random).uniform(0, 2 * np.pi)
    else:
        angle = np.radians(angle)
    out = np.abs(np.fft.ifft2(np.fft.fft2(img) * mtf_smear(img.shape, distance, angle, pixelscale, oversample)))
    return out * np.sum(img) / np.sum(out)


def rebin(img, factor):
    """util.py:221-259."""
    img = np.asarray(img)
    if np.iscomplexobj(img):
        raise ValueError('rebin is not defined for complex data')
    if img.ndim == 3:
        shp = (img.shape[0], img.shape[1] // factor, img.shape[2] // factor)
        out = np.zeros(shp, dtype=img.dtype)
        for i in range(img.shape[0]):
            out[i] = img[i].reshape(shp[1], factor, shp[2], factor).sum(-1).sum(1)
        return out
    return img.reshape(img.shape[0] // factor, factor, img.shape[1] // factor, factor).sum(-1).sum(1)


def rescale_shape(img_shape, scale, shape=None):
    """Output shape of util.py:325-331."""
    if shape is None:
        return np.ceil((img_shape[0] * scale, img_shape[1] * scale)).astype(int)
    if np.isscalar(shape):
        return np.ceil((shape * scale, shape * scale)).astype(int)
    return np.ceil((shape[0] * scale, shape[1] * scale)).astype(int)


def rescale_coords(img_shape, scale, shape):
    """Sample coordinates of util.py:333-336."""
    x = (np.arange(shape[1], dtype=np.float64) - shape[1] / 2.) / scale + img_shape[1] / 2.
    y = (np.arange(shape[0], dtype=np.float64) - shape[0] / 2.) / scale + img_shape[0] / 2.
    return np.meshgrid(x, y)


def rescale(img, scale, shape=None, mask=None, order=3, mode='nearest', unitary=True):
    """util.py:262-349."""
    img = np.asarray(img)
    if mask is None:
        mask = np.zeros_like(img).real
        mask[img != 0] = 1
    shape = rescale_shape(img.shape, scale, shape)
    xx, yy = rescale_coords(img.shape, scale, shape)
    mask = map_coordinates(mask, [yy, xx], order=1, mode='nearest')
    mask[mask < np.finfo(mask.dtype).eps] = 0
    if np.iscomplexobj(img):
        out = np.zeros(shape, dtype=np.complex128)
        out.real = map_coordinates(img.real, [yy, xx], order=order, mode=mode)
        out.imag = map_coordinates(img.imag, [yy, xx], order=order, mode=mode)
    else:
        out = map_coordinates(img, [yy, xx], order=order, mode=mode)
    if unitary:
        out *= np.sum(img) / np.sum(out)
    out *= mask
    return out


# --------------------------------------------------------------------------
# the explicit algorithm of map_coordinates(order=3, mode='nearest')
# --------------------------------------------------------------------------

SPLINE_POLE = np.sqrt(3.0) - 2.0   # the pole of the cubic B-spline prefilter
SPLINE_PAD = 12                    # edge samples added before prefiltering
SPLINE_TAPS = 32                   # |k| range of the truncated impulse response


def _fold(i, n):
    """Half-sample symmetric (reflect) index of i on an n-sample line."""
    i = i % (2 * n)
    return 2 * n - 1 - i if i >= n else i


def spline_coefficients(img, taps=SPLINE_TAPS):
    """Cubic B-spline coefficients of the edge-padded image: along each axis
    c[i] = sqrt(3) * sum_{|k| <= taps} z^|k| x[fold(i + k)], z = sqrt(3) - 2,
    the two-sided impulse response of the inverse of [1, 4, 1] / 6 applied to
    the reflect-extended line.  Equal to the recursive prefilter with the
    reflect boundary to rounding for lines of 12 or more samples."""
    p = np.pad(np.asarray(img, dtype=float), SPLINE_PAD, mode='edge')
    k = np.arange(-taps, taps + 1)
    w = np.sqrt(3.0) * SPLINE_POLE ** np.abs(k)
    for axis in (0, 1):
        n = p.shape[axis]
        idx = np.array([[_fold(i + kk, n) for kk in k] for i in range(n)])   # (n, 2 taps + 1)
        p = np.tensordot(np.take(p, idx, axis=axis), w, axes=([axis + 1], [0]))
    return p


def bspline3_weights(t):
    """Cubic B-spline weights of taps floor(c) - 1 .. floor(c) + 2 at fraction t."""
    return np.stack([(1 - t) ** 3 / 6, (4 - 6 * t * t + 3 * t ** 3) / 6,
                     (1 + 3 * t + 3 * t * t - 3 * t ** 3) / 6, t ** 3 / 6])


def spline_sample(coef, yy, xx):
    """Cubic interpolation of prefiltered, padded coefficients at image
    coordinates (yy, xx), mode 'nearest' (coordinates and taps clamped)."""
    P0, P1 = coef.shape
    cy = np.clip(yy + SPLINE_PAD, 0.0, P0 - 1.0)
    cx = np.clip(xx + SPLINE_PAD, 0.0, P1 - 1.0)
    fy, fx = np.floor(cy), np.floor(cx)
    wy, wx = bspline3_weights(cy - fy), bspline3_weights(cx - fx)
    out = np.zeros(yy.shape)
    for a in range(4):
        r = np.clip(fy.astype(int) - 1 + a, 0, P0 - 1)
        row = np.zeros(yy.shape)
        for b in range(4):
            c = np.clip(fx.astype(int) - 1 + b, 0, P1 - 1)
            row += wx[b] * coef[r, c]
        out += wy[a] * row
    return out


def linear_sample(m, yy, xx):
    """map_coordinates(order=1, mode='nearest'): bilinear, clamped."""
    P0, P1 = m.shape
    cy = np.clip(yy, 0.0, P0 - 1.0)
    cx = np.clip(xx, 0.0, P1 - 1.0)
    fy, fx = np.floor(cy).astype(int), np.floor(cx).astype(int)
    ty, tx = cy - fy, cx - fx
    r1, c1 = np.minimum(fy + 1, P0 - 1), np.minimum(fx + 1, P1 - 1)
    return ((1 - ty) * ((1 - tx) * m[fy, fx] + tx * m[fy, c1]) +
            ty * ((1 - tx) * m[r1, fx] + tx * m[r1, c1]))


# --------------------------------------------------------------------------
# the explicit algorithm of map_coordinates(order 0..5, every boundary mode)
# --------------------------------------------------------------------------
# Restated from scipy.ndimage's documented behaviour (scipy >= 1.6, the
# version range lentil's rescale runs on) and pinned against the library in
# tests/test_oracle_detector.py:
#   * orders >= 2 prefilter the input with the B-spline's inverse filter,
#     whose boundary rule depends on the mode; 'nearest' and 'grid-constant'
#     first pad 12 samples (edge values, or cval);
#   * the spline of order p is sampled with p + 1 taps per axis, the taps
#     outside the (padded) line extended by the mode's rule;
#   * 'constant' gives cval outside [0, n - 1] and 'wrap' first maps the
#     coordinate into [0, n - 1] with period n - 1.
# cval is 0 throughout: util.rescale never passes one.

MODES = ('nearest', 'reflect', 'grid-mirror', 'mirror', 'grid-wrap', 'wrap', 'constant', 'grid-constant')
# poles of the B-spline prefilter of each order (Unser's values)
SPLINE_POLES = {2: (np.sqrt(8.0) - 3.0,),
                3: (np.sqrt(3.0) - 2.0,),
                4: (np.sqrt(664.0 - np.sqrt(438976.0)) + np.sqrt(304.0) - 19.0,
                    np.sqrt(664.0 + np.sqrt(438976.0)) - np.sqrt(304.0) - 19.0),
                5: (np.sqrt(67.5 - np.sqrt(4436.25)) + np.sqrt(26.25) - 6.5,
                    np.sqrt(67.5 + np.sqrt(4436.25)) - np.sqrt(26.25) - 6.5)}
FILTER_RULE = {'mirror': 'mirror', 'wrap': 'mirror', 'constant': 'mirror', 'grid-constant': 'mirror',
               'reflect': 'reflect', 'grid-mirror': 'reflect', 'nearest': 'reflect', 'grid-wrap': 'periodic'}
TAP_RULE = {'nearest': 'clamp', 'reflect': 'reflect', 'grid-mirror': 'reflect', 'mirror': 'mirror',
            'wrap': 'mirror', 'constant': 'mirror', 'grid-wrap': 'periodic', 'grid-constant': 'zero'}
PREFILTER_TAPS = 48   # |k| range of the truncated impulse response (|z|^48 < 3e-18 for order 5)


def extend_index(i, n, rule):
    """Index of sample i of an n-sample line under an extension rule; None
    for a tap that reads the constant 0 ('zero')."""
    if 0 <= i < n:
        return i
    if rule == 'clamp':
        return min(max(i, 0), n - 1)
    if rule == 'reflect':            # half-sample symmetric, period 2n
        i %= 2 * n
        return 2 * n - 1 - i if i >= n else i
    if rule == 'mirror':             # whole-sample symmetric, period 2n - 2
        if n == 1:
            return 0
        i %= 2 * n - 2
        return 2 * n - 2 - i if i >= n else i
    if rule == 'periodic':
        return i % n
    if rule == 'zero':
        return None
    raise ValueError(rule)


def prefilter_response(order, taps=PREFILTER_TAPS):
    """Two-sided impulse response h[-taps..taps] of the order's B-spline
    prefilter: the product over its poles z of (1 - z) / (1 + z) z^|k|."""
    k = np.arange(-4 * taps, 4 * taps + 1)
    h = np.zeros(k.size)
    h[4 * taps] = 1.0
    for z in SPLINE_POLES.get(order, ()):
        h = np.convolve(h, (1 - z) / (1 + z) * z ** np.abs(k), mode='same')
    return h[3 * taps:5 * taps + 1]


def prefilter(img, order, rule, taps=PREFILTER_TAPS):
    """The prefilter along both axes with the boundary rule of the mode."""
    p = np.asarray(img, dtype=float)
    if order < 2:
        return p.copy()
    h = prefilter_response(order, taps)
    k = np.arange(-taps, taps + 1)
    for axis in (0, 1):
        n = p.shape[axis]
        idx = np.array([[extend_index(i + kk, n, rule) for kk in k] for i in range(n)])
        p = np.tensordot(np.take(p, idx, axis=axis), h, axes=([axis + 1], [0]))
    return p


def bspline(order, t):
    """The centred B-spline of degree ``order`` at t."""
    t = np.abs(np.asarray(t, dtype=float))
    if order == 0:
        return np.where(t <= 0.5, 1.0, 0.0)
    if order == 1:
        return np.where(t < 1, 1 - t, 0.0)
    if order == 2:
        return np.where(t < 0.5, 0.75 - t * t, np.where(t < 1.5, 0.5 * (1.5 - t) ** 2, 0.0))
    if order == 3:
        return np.where(t < 1, 2 / 3 - t * t + t ** 3 / 2, np.where(t < 2, (2 - t) ** 3 / 6, 0.0))
    if order == 4:
        return np.where(t < 0.5, 115 / 192 - 5 * t * t / 8 + t ** 4 / 4,
                        np.where(t < 1.5, 55 / 96 + 5 * t / 24 - 5 * t * t / 4 + 5 * t ** 3 / 6 - t ** 4 / 6,
                                 np.where(t < 2.5, (2.5 - t) ** 4 / 24, 0.0)))
    if order == 5:
        return np.where(t < 1, 11 / 20 - t * t / 2 + t ** 4 / 4 - t ** 5 / 12,
                        np.where(t < 2, 17 / 40 + 5 * t / 8 - 7 * t * t / 4 + 5 * t ** 3 / 4 - 3 * t ** 4 / 8
                                 + t ** 5 / 24, np.where(t < 3, (3 - t) ** 5 / 120, 0.0)))
    raise ValueError(order)


def map_coordinate(x, n, mode):
    """A coordinate outside [0, n - 1] moved inside by the mode's symmetry
    ('mirror': whole-sample, 'reflect'/'grid-mirror': half-sample, 'wrap':
    period n - 1, 'grid-wrap': period n, 'nearest': clamp); the other modes
    keep it.  Taps then follow the same rule, so for orders >= 1 the mapping
    changes nothing; for order 0 it decides which neighbour a half-integer
    coordinate rounds to."""
    x = np.asarray(x, dtype=float)
    if n <= 1 and mode in ('mirror', 'reflect', 'grid-mirror', 'wrap', 'grid-wrap', 'nearest'):
        return np.zeros_like(x)
    if mode == 'mirror':
        p = 2 * n - 2
        lo = x + p * np.trunc(-x / p)
        lo = np.where(lo <= 1 - n, lo + p, -lo)
        hi = x - p * np.trunc(x / p)
        hi = np.where(hi >= n, p - hi, hi)
        return np.where(x < 0, lo, np.where(x > n - 1, hi, x))
    if mode in ('reflect', 'grid-mirror'):
        p = 2 * n
        lo = np.where(x < -p, x + p * np.trunc(-x / p), x)
        lo = np.where(lo < -n, lo + p, np.where(lo > -1e-15, 1e-15, -lo) - 1)
        hi = x - p * np.trunc(x / p)
        hi = np.where(hi >= n, p - hi - 1, hi)
        return np.where(x < 0, lo, np.where(x > n - 1, hi, x))
    if mode == 'wrap':
        p = n - 1
        lo = x + p * (np.trunc(-x / p) + 1)
        hi = x - p * np.trunc(x / p)
        return np.where(x < 0, lo, np.where(x > n - 1, hi, x))
    if mode == 'grid-wrap':
        lo = x + n * (np.trunc(-x / n) + 1)
        hi = x - n * np.trunc(x / n)
        return np.where(x < 0, lo, np.where(x > n - 1, hi, x))
    if mode == 'nearest':
        return np.clip(x, 0, n - 1)
    return x


def _first_tap(x, order):
    return np.floor(x).astype(int) - (order - 1) // 2 if order % 2 else np.floor(x + 0.5).astype(int) - order // 2


def map_coordinates_explicit(img, yy, xx, order, mode):
    """map_coordinates(img, [yy, xx], order, mode) for a real 2-D image,
    restated: pad, prefilter, then p + 1 taps per axis."""
    img = np.asarray(img, dtype=float)
    rows, cols = img.shape
    npad = 12 if order > 1 and mode in ('nearest', 'grid-constant') else 0
    if npad:
        img = np.pad(img, npad, mode='edge' if mode == 'nearest' else 'constant')
    coef = prefilter(img, order, FILTER_RULE[mode])
    yy = np.asarray(yy, dtype=float)
    xx = np.asarray(xx, dtype=float)
    outside = np.zeros(yy.shape, dtype=bool)
    if mode == 'constant':
        outside = (yy < 0) | (yy > rows - 1) | (xx < 0) | (xx > cols - 1)
    if not npad:   # a padded line already holds the extension near its edges
        yy, xx = map_coordinate(yy, rows, mode), map_coordinate(xx, cols, mode)
    cy, cx = yy + npad, xx + npad
    P0, P1 = coef.shape
    rule = TAP_RULE[mode]
    sy, sx = _first_tap(cy, order), _first_tap(cx, order)
    out = np.zeros(yy.shape)
    for idx in np.ndindex(yy.shape):
        if outside[idx]:
            continue
        v = 0.0
        for a in range(order + 1):
            r = extend_index(int(sy[idx]) + a, P0, rule)
            if r is None:
                continue
            wy = float(bspline(order, cy[idx] - (sy[idx] + a)))
            row = 0.0
            for b in range(order + 1):
                c = extend_index(int(sx[idx]) + b, P1, rule)
                if c is not None:
                    row += float(bspline(order, cx[idx] - (sx[idx] + b))) * coef[r, c]
            v += wy * row
        out[idx] = v
    return out
