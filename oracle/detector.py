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
