"""Seeded inputs of the five named workloads of the MFT path.

Each builder returns a dict with the pupil arrays and the propagation
arguments of the reference user loop:

    amp, opd, mask (2-D, or 3-D per-segment), dx, focal_length,
    wavelengths, weights, pixelscale, shape, oversample, tilts [(x, y) angles]

Physical parameters the workload descriptions leave open are fixed here and
recorded in the dict ('notes').  Inputs are host NumPy arrays; nothing here
touches the device.
"""
import numpy as np

from .segmented import hex_segments
from .shape import circle
from .util import normalize_power
from .zernike import zernike_compose

DU = 5e-6          # detector pixel pitch [m]
FOCAL = 10.0       # focal length [m]


def _coeffs(rng, nmax, sigma, rms=None):
    c = np.zeros(nmax)
    c[1:] = rng.normal(0, sigma, size=nmax - 1)
    if rms is not None:
        c *= rms / np.sqrt(np.sum(c ** 2))
    return c


def mono_256(seed=1):
    """256^2 circular pupil (r = 128 px, 1 m), Noll 2-11 ~ N(0, 20 nm), 650 nm,
    5 um pixels, npix 128, oversample 2 -> one 256^2 PSF."""
    rng = np.random.default_rng(seed)
    amp = normalize_power(circle((256, 256), 128))
    opd = zernike_compose(amp, _coeffs(rng, 11, 20e-9))
    return dict(name='mono_256', amp=amp, opd=opd, mask=None, dx=1 / 256, focal_length=FOCAL,
                wavelengths=np.array([650e-9]), weights=np.array([1.0]), pixelscale=DU, shape=128,
                oversample=2, tilts=[], notes='1 m aperture over 256 px, f = 10 m')


def broadband_512(seed=2, nwave=64):
    """512^2 pupil with 12% central obscuration, Noll 2-22 ~ N(0, 30 nm),
    64 wavelengths 500-900 nm with flat weights, npix 256, oversample 4 ->
    1024^2 image."""
    rng = np.random.default_rng(seed)
    n = 512
    amp = normalize_power(circle((n, n), n // 2) - circle((n, n), 0.12 * n / 2))
    opd = zernike_compose(amp, _coeffs(rng, 22, 30e-9))
    wl = np.linspace(500e-9, 900e-9, nwave)
    return dict(name='broadband_512', amp=amp, opd=opd, mask=None, dx=1 / n, focal_length=FOCAL,
                wavelengths=wl, weights=np.full(nwave, 1 / nwave), pixelscale=DU, shape=256, oversample=4,
                tilts=[], notes='1 m aperture over 512 px, f = 10 m, 5 um pixels')


def segmented_18(seed=3, nwave=32):
    """18 hexagonal segments (flat-to-flat 160 px, ring 2 with the centre
    dropped) in a 1024^2 pupil, clipped by a circular stop so the tile crops
    are uneven; per-segment piston N(0, 50 nm) and tip/tilt N(0, 0.1 urad),
    global Noll 2-11 at 20 nm rms; 32 wavelengths 600-1000 nm; npix 512,
    oversample 2; a global sub-pixel Tilt."""
    rng = np.random.default_rng(seed)
    seg_radius = 160 / np.sqrt(3)           # circumradius for a 160 px flat-to-flat
    segs = hex_segments(rings=2, seg_radius=seg_radius, seg_gap=4, drop=(0,))
    size = segs.shape[-1]
    n = 1024
    pad = (n - size) // 2
    masks = np.zeros((segs.shape[0], n, n))
    masks[:, pad:pad + size, pad:pad + size] = segs
    stop = circle((n, n), 372)
    masks = masks * stop
    keep = [i for i in range(masks.shape[0]) if masks[i].sum() > 0]
    masks = masks[keep]
    dx = 1.0 / 800                          # 1.28 m grid
    glob = np.sum(masks, axis=0)
    amp = normalize_power(glob)
    opd = zernike_compose(glob > 0, _coeffs(rng, 11, 1.0, rms=20e-9))
    rr, cc = np.mgrid[0:n, 0:n] - n // 2
    for m in masks:
        piston = rng.normal(0, 50e-9)
        tx, ty = rng.normal(0, 0.1e-6, size=2)
        opd = opd + (m > 0) * (piston + tx * cc * dx - ty * rr * dx)
    wl = np.linspace(600e-9, 1000e-9, nwave)
    # a global tilt giving ~(0.37, -0.21) output pixels at oversample 2
    tilt = (-0.21 * DU / (FOCAL * 2), -0.37 * DU / (FOCAL * 2))
    return dict(name='segmented_18', amp=amp, opd=opd, mask=masks, dx=dx, focal_length=FOCAL,
                wavelengths=wl, weights=np.full(nwave, 1 / nwave), pixelscale=DU, shape=512, oversample=2,
                tilts=[tilt], notes='hex ring 2 (18 segments), 4 px gaps, circular stop r=372 px, dx=1/800 m')


def phase_retrieval_512(seed=4, nwave=16):
    """512^2 circular pupil, truth OPD Noll 2-37 ~ N(0, 40 nm), 4 defocus
    diversities (-2, -1, +1, +2 waves at 650 nm), 16 wavelengths 550-750 nm,
    npix 256, oversample 2."""
    rng = np.random.default_rng(seed)
    n = 512
    amp = normalize_power(circle((n, n), n // 2))
    opd = zernike_compose(amp, _coeffs(rng, 37, 40e-9))
    defocus = zernike_compose(amp, np.r_[0, 0, 0, 1.0])
    lam_c = 650e-9
    # Z4 coefficient giving d waves peak-to-valley of defocus: PV(Z4) = 2*sqrt(3)
    div = [d * lam_c / (2 * np.sqrt(3)) * defocus for d in (-2, -1, 1, 2)]
    wl = np.linspace(550e-9, 750e-9, nwave)
    return dict(name='phase_retrieval_512', amp=amp, opd=opd, mask=None, dx=1 / n, focal_length=FOCAL,
                wavelengths=wl, weights=np.full(nwave, 1 / nwave), pixelscale=DU, shape=256, oversample=2,
                tilts=[], diversity=div, notes='defocus diversity in waves PV at 650 nm')


def stress_2048(seed=5, nwave=128):
    """2048^2 pupil with 6 hexagonal segments (ring 1, centre dropped), Noll
    2-22 at 25 nm rms, 128 wavelengths 450-950 nm, npix 512, oversample 8 ->
    4096^2, off-axis shift (37.3, -12.6) output pixels via a Tilt."""
    rng = np.random.default_rng(seed)
    seg_radius = 560.0
    segs = hex_segments(rings=1, seg_radius=seg_radius, seg_gap=8, drop=(0,))
    size = segs.shape[-1]
    n = 2048
    masks = np.zeros((segs.shape[0], n, n))
    a = (n - size) // 2
    if a >= 0:
        masks[:, a:a + size, a:a + size] = segs
    else:
        masks[:] = segs[:, -a:-a + n, -a:-a + n]
    glob = np.sum(masks, axis=0)
    amp = normalize_power(glob)
    opd = zernike_compose(glob > 0, _coeffs(rng, 22, 1.0, rms=25e-9))
    dx = 2.0 / n
    os_ = 8
    # Field.shift 'ij': (r, c) = (-y', x') with x' = -z*y_tilt_ctor/du*os ...
    # Tilt(x, y) stores self.x = y, self.y = x; shift x = -z*self.x = -z*y -> r = -(-z*x)/du*os
    r_px, c_px = 37.3, -12.6
    tilt = (r_px * DU / (FOCAL * os_), -c_px * DU / (FOCAL * os_))
    wl = np.linspace(450e-9, 950e-9, nwave)
    return dict(name='stress_2048', amp=amp, opd=opd, mask=masks, dx=dx, focal_length=FOCAL,
                wavelengths=wl, weights=np.full(nwave, 1 / nwave), pixelscale=DU, shape=512, oversample=os_,
                tilts=[tilt], notes='hex ring 1 (6 segments), seg radius 560 px, 8 px gaps, dx=2/2048 m')


WORKLOADS = {
    'mono_256': mono_256,
    'broadband_512': broadband_512,
    'segmented_18': segmented_18,
    'phase_retrieval_512': phase_retrieval_512,
    'stress_2048': stress_2048,
}
