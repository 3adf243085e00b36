"""Detector post-processing: the oracle against the reference's own tests and
closed forms, and the explicit spline algorithm (the one the device kernels
follow) against scipy.ndimage.

Reference known answers: tests/test_util.py:13-36 (rebin tables, rescale
unitary), tests/test_detector.py:86-100 (pixel unitary, pixelate),
tests/test_convolvable.py:5-38 (jitter scale and unitary, smear unitary).
"""
import numpy as np
import pytest
import scipy.ndimage as nd

from oracle import detector as D


def test_rebin_known_answers():
    img = np.array([[1, 1, 2, 2], [1, 1, 2, 2], [3, 3, 4, 4], [3, 3, 4, 4]])
    assert np.array_equal(D.rebin(img, 2), np.array([[4, 8], [12, 16]]))
    cube = np.stack([img, 2 * img])
    assert np.array_equal(D.rebin(cube, 2), np.array([[[4, 8], [12, 16]], [[8, 16], [24, 32]]]))
    with pytest.raises(ValueError):
        D.rebin(img * 1j, 2)


def test_rescale_unitary_known_answer():
    a = np.random.default_rng(0).uniform(low=0, high=1, size=(100, 100))
    b = D.rescale(a, scale=0.5, order=3, mode='nearest', unitary=True)
    assert np.isclose(np.sum(a), np.sum(b))


def test_pixel_and_convolvers_preserve_weight():
    a = np.random.default_rng(1).uniform(size=(64, 64))
    assert np.isclose(np.sum(D.pixel(a, oversample=1)), np.sum(a))
    assert np.isclose(np.sum(D.jitter(a, 2)), np.sum(a))
    assert np.isclose(np.sum(D.smear(a, 10, angle=30)), np.sum(a))
    out = D.pixelate(np.random.default_rng(2).uniform(size=(20, 20)), 2)
    assert out.shape == (10, 10)


def test_pixel_closed_form():
    """At oversample 1 the MTF zeroes only the Nyquist bin of an even grid;
    at oversample 2 on a 2-periodic image it leaves the mean alone."""
    a = np.zeros((8, 8))
    a[::2, ::2] = 1.0
    assert np.allclose(D.pixel(a, 2), 0.25)


def test_jitter_scale_known_answer():
    """tests/test_convolvable.py:5-18: the Gaussian kernel's 1-sigma width."""
    a = np.zeros((101, 101))
    a[50, 50] = 1
    b = D.jitter(a, 3)
    assert np.isclose(b[50, 47] / b[50, 50], np.exp(-0.5), atol=1e-3)


@pytest.mark.parametrize('n', [12, 25, 40, 97])
def test_spline_coefficients_match_scipy(n):
    rng = np.random.default_rng(n)
    img = rng.standard_normal((n, n + 3))
    p = np.pad(img, D.SPLINE_PAD, mode='edge')
    ref = nd.spline_filter(p, 3, output=np.float64, mode='nearest')
    got = D.spline_coefficients(img)
    assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()


@pytest.mark.parametrize('scale', [0.5, 1 / 3, 0.25, 1.7, 2.0])
def test_spline_sample_matches_map_coordinates(scale):
    rng = np.random.default_rng(5)
    img = rng.random((31, 26))
    img[:4] = 0
    shape = D.rescale_shape(img.shape, scale)
    xx, yy = D.rescale_coords(img.shape, scale, shape)
    ref = nd.map_coordinates(img, [yy, xx], order=3, mode='nearest')
    got = D.spline_sample(D.spline_coefficients(img), yy, xx)
    assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()
    mask = (img != 0).astype(float)
    assert np.abs(D.linear_sample(mask, yy, xx) - nd.map_coordinates(mask, [yy, xx], order=1,
                                                                     mode='nearest')).max() < 1e-15


@pytest.mark.parametrize('bad', ['pole_sign', 'no_edge_pad', 'tap_offset'])
def test_planted_spline_mistakes_are_caught(monkeypatch, bad):
    """A flipped pole, a dropped edge padding, or taps one sample off, each
    fail the map_coordinates comparison.  (A mirror boundary in place of
    reflect would not: after 12 edge samples the two differ by z^12 ~ 1e-7.)"""
    rng = np.random.default_rng(6)
    img = rng.random((30, 30))
    shape = D.rescale_shape(img.shape, 2.0)
    xx, yy = D.rescale_coords(img.shape, 2.0, shape)   # samples reach past the edges
    good = nd.map_coordinates(img, [yy, xx], order=3, mode='nearest')
    if bad == 'pole_sign':
        monkeypatch.setattr(D, 'SPLINE_POLE', -D.SPLINE_POLE)
    elif bad == 'no_edge_pad':
        monkeypatch.setattr(D, 'SPLINE_PAD', 0)
    else:
        yy = yy + 1
    got = D.spline_sample(D.spline_coefficients(img), yy, xx)
    assert np.abs(got - good).max() > 1e-3


def test_jitter_fwhm_known_answer():
    """tests/test_convolvable.py:5-18 verbatim in behaviour."""
    a = np.zeros((11, 11))
    a[5, 5] = 1
    b = D.jitter(a, 2)
    arr = b / np.max(b)
    arr[arr < 0.5] = 0
    rows, cols = np.flatnonzero(arr.any(1)), np.flatnonzero(arr.any(0))
    assert max(rows[-1] - rows[0], cols[-1] - cols[0]) // 2 == 2


def test_host_errors_before_device_work():
    """The drop-in layer raises the reference's exception types before any
    device work: util.py:243-244 for complex rebin data, and scipy's own
    RuntimeError for a spline order or boundary mode map_coordinates does not
    have (util.rescale passes both straight to it)."""
    from lentil_amd import util
    with pytest.raises(ValueError):
        util.rebin(np.ones((4, 4)) * 1j, 2)
    with pytest.raises(RuntimeError, match='spline order'):
        util.rescale(np.ones((4, 4)), 0.5, order=6)
    with pytest.raises(RuntimeError, match='boundary mode'):
        util.rescale(np.ones((4, 4)), 0.5, mode='bogus')
    with pytest.raises(RuntimeError, match='spline order'):
        nd.map_coordinates(np.ones((4, 4)), [[0.5], [0.5]], order=6)
    with pytest.raises(RuntimeError, match='boundary mode'):
        nd.map_coordinates(np.ones((4, 4)), [[0.5], [0.5]], mode='bogus')
    from lentil_amd import detector
    with pytest.raises(TypeError):
        detector.pixel(np.ones((4, 4)) * 1j)


@pytest.mark.parametrize('order', range(6))
@pytest.mark.parametrize('mode', D.MODES)
def test_map_coordinates_explicit_matches_scipy(order, mode):
    """The explicit restatement of map_coordinates (prefilter boundary rule,
    padding, tap extension, 'constant' cut-off and 'wrap' coordinate map)
    equals scipy's for every mode and spline order, at coordinates inside,
    on and far outside the image, on an odd, non-square image.  Lines are at
    least 24 samples: scipy starts its recursive prefilter from a truncated
    boundary sum, which on shorter lines differs from the exact filter by
    about |z|^(2n) (4e-8 for order 5 on 9 samples)."""
    rng = np.random.default_rng(order * 11 + D.MODES.index(mode))
    img = rng.standard_normal((25, 30))
    ys = np.concatenate([np.linspace(-49.3, 77.1, 19), [0.0, 24.0, -0.5, 24.5, 3.5, -1e-12, -1.5, 25.5]])
    xs = np.concatenate([np.linspace(-55.2, 89.6, 23), [0.0, 29.0, -0.5, 29.5, 6.5, 29 + 1e-12, -1.5, 30.5]])
    xx, yy = np.meshgrid(xs, ys)
    ref = nd.map_coordinates(img, [yy, xx], order=order, mode=mode)
    got = D.map_coordinates_explicit(img, yy, xx, order, mode)
    assert np.max(np.abs(got - ref)) < 1e-11 * max(1.0, np.max(np.abs(ref)))


def test_map_coordinates_explicit_planted_mistakes():
    """Mistakes a restatement could make are caught by scipy: the reflect
    and mirror rules swapped, and legacy 'wrap' taken as periodic."""
    rng = np.random.default_rng(5)
    img = rng.standard_normal((9, 14))
    xx, yy = np.meshgrid(np.linspace(-3.3, 16.1, 17), np.linspace(-2.2, 11.4, 13))
    saved = dict(D.TAP_RULE), dict(D.FILTER_RULE)
    try:
        D.TAP_RULE['reflect'] = 'mirror'
        bad = D.map_coordinates_explicit(img, yy, xx, 3, 'reflect')
        assert np.max(np.abs(bad - nd.map_coordinates(img, [yy, xx], order=3, mode='reflect'))) > 1e-3
        D.TAP_RULE.update(saved[0])
        D.FILTER_RULE['wrap'] = 'periodic'
        bad = D.map_coordinates_explicit(img, yy, xx, 3, 'wrap')
        assert np.max(np.abs(bad - nd.map_coordinates(img, [yy, xx], order=3, mode='wrap'))) > 1e-3
    finally:
        D.TAP_RULE.update(saved[0])
        D.FILTER_RULE.update(saved[1])
