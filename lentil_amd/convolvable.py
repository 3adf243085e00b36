"""Image jitter and smear on the device (lentil convolvable.py:4-169).

Both are an MTF convolution that preserves the image weight:
|ifft2(fft2(img) * K)| * sum(img) / sum(out), run by ``mftx_fft_filter``
with K built by the reference's numpy expressions.
"""
import numpy as np

from .detector import real_image, finish, device_kernel, fft_filter


def mtf_jitter(shape, scale, pixelscale=1, oversample=1):
    """Gaussian jitter MTF (convolvable.py:72-77)."""
    x = np.fft.fftfreq(shape[1])
    y = np.fft.fftfreq(shape[0])
    xx, yy = np.meshgrid(x, y)
    rho = np.sqrt(xx ** 2 + yy ** 2)
    return np.exp(-2 * (np.pi * (scale / pixelscale) * oversample * rho) ** 2)


def mtf_smear(shape, distance, angle, pixelscale=1, oversample=1):
    """Directional smear MTF (convolvable.py:156-166), angle in radians."""
    x = np.fft.fftfreq(shape[1])
    y = np.fft.fftfreq(shape[0])
    xx, yy = np.meshgrid(x, y)
    yy_rot = np.sin(angle) * yy + np.cos(angle) * xx
    return np.sinc(yy_rot * (distance / pixelscale) * oversample)


def jitter(img, scale, pixelscale=1, oversample=1, *, precision=None):
    """Apply image jitter via convolution (convolvable.py:4-82)."""
    t, host = real_image(img)
    k = device_kernel('jitter', t.shape, scale, pixelscale, oversample)
    return finish(fft_filter(t, k, True, precision), host)


def smear(img, distance, angle=None, pixelscale=1, oversample=1, *, precision=None):
    """Apply image smear via convolution (convolvable.py:85-169).  With
    angle None a direction is drawn from numpy's global generator, as the
    reference does."""
    t, host = real_image(img)
    if angle is None:
        angle = np.random.uniform(0, 2 * np.pi)
    else:
        angle = np.radians(angle)
    k = device_kernel('smear', t.shape, distance, angle, pixelscale, oversample)
    return finish(fft_filter(t, k, True, precision), host)
