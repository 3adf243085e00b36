"""propagate_fft / scratch_shape: the oracle restatement and the host logic
of the device path.

The oracle's FFT propagation (propagate.py:9-135) is numpy's FFT, a library
routine; what needs pinning is the product's mapping of that zero-padded,
shifted FFT onto an MFT with alpha = 1/N (``centred_fft_geometry``), checked
here on even and odd grids, padded and cropped arrays, against the oracle's
FFT in complex128.  The Airy known answer is tests/test_propagate_fft.py:5-29
(whose comparison lacks an assert; it is asserted here).
"""
import numpy as np
import pytest
import scipy.special

from oracle import inputs, mft, optics
from tests.conftest import rel_l2


def _cplx(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _mft_route(x, n, geometry):
    """The device path's arithmetic on the CPU: crop like lentil.pad, then one
    MFT with alpha = 1/N and the mapped offsets."""
    (s0, s1), (n0, n1) = x.shape, n
    a0, a1 = max(s0 - n0, 0) // 2, max(s1 - n1, 0) // 2
    v = x[a0:a0 + min(s0, n0), a1:a1 + min(s1, n1)]
    pad = (max(n0 - s0, 0) // 2, max(n1 - s1, 0) // 2)
    (o0, h0), (o1, h1) = geometry(n0, v.shape[0], pad[0]), geometry(n1, v.shape[1], pad[1])
    return mft.dft2(v, (1 / n0, 1 / n1), (n0, n1), (h0, h1), (o0, o1), unitary=True)


@pytest.mark.parametrize('n', [(16, 16), (15, 15), (16, 15), (9, 12)])
@pytest.mark.parametrize('m', [(8, 8), (7, 9), (16, 15), (20, 21), (1, 1)])
def test_fft_is_mft(n, m):
    from lentil_amd.propagate import centred_fft_geometry
    x = _cplx(np.random.default_rng(sum(n) * 100 + sum(m)), *m)
    ref = optics.fft2_centred(optics.pad(x, n))
    got = _mft_route(x, n, centred_fft_geometry)
    assert rel_l2(got, ref) < 1e-12


@pytest.mark.parametrize('bad', ['no_odd_shift', 'floor_pad'])
def test_planted_mapping_mistakes_are_caught(bad):
    """A mapping that ignores the odd-grid shift, or centres the pad the other
    way, fails the same comparison."""
    from lentil_amd.propagate import _wrap

    def geometry(n, m, pad):
        h = n // 2
        if bad == 'no_odd_shift':
            return _wrap(pad + h + m // 2, n), 0
        return _wrap(pad + 1 + h + m // 2, n), _wrap(-2 * h, n)

    x = _cplx(np.random.default_rng(3), 7, 9)
    n = (15, 15)
    ref = optics.fft2_centred(optics.pad(x, n))
    assert rel_l2(_mft_route(x, n, geometry), ref) > 1e-3


def test_pad_known_answers():
    """util.py:31-90: centred pad and centred crop."""
    a = np.arange(6).reshape(2, 3)
    p = optics.pad(a, (4, 5))
    assert p.shape == (4, 5) and np.array_equal(p[1:3, 1:4], a) and p.sum() == a.sum()
    c = optics.pad(np.arange(25).reshape(5, 5), (3, 2))
    assert np.array_equal(c, np.arange(25).reshape(5, 5)[1:4, 1:3])
    from lentil_amd import util
    for shape in [(4, 5), (3, 2), (2, 7)]:
        x = np.arange(30).reshape(5, 6)
        assert np.array_equal(util.pad(x, shape), optics.pad(x, shape))


def test_fft_shape_known_answers():
    """propagate.py:126-131, 87-118: the reference test's geometry gives a 512
    grid, and the propagation wavelength is recomputed from the integer grid."""
    dx, du = np.array([1 / 512] * 2), np.array([5e-6] * 2)
    n, wl = optics.fft_shape(dx, du, 10, 500e-9, 1)
    assert tuple(n) == (512, 512) and wl == pytest.approx(500e-9, rel=1e-12)
    n, wl = optics.fft_shape(dx, du, 10, 503e-9, 2)
    assert tuple(n) == (1030, 1030)    # round(1 / alpha) = round(1030.14)
    assert wl == pytest.approx(1030 / 2 * dx[0] * du[0] / 10, rel=1e-15)
    assert optics.scratch_shape([400e-9, 503e-9], 1 / 512, 5e-6, 10, 2) == (1030, 1030)
    from lentil_amd import propagate
    assert propagate.scratch_shape([400e-9, 503e-9], 1 / 512, 5e-6, 10, 2) == (1030, 1030)
    pn, pwl = propagate._fft_shape(dx, du, 10, 503e-9, 2)
    assert tuple(pn) == (1030, 1030) and pwl == wl


def test_oracle_airy():
    """tests/test_propagate_fft.py:5-29 with tests/fixtures/airy2.py."""
    amp = inputs.normalize_power(inputs.circle((512, 512), 256))
    w = optics.Pupil(amp, 0, 1 / 512, 10).__rmul__(optics.Wavefront(500e-9))
    w = optics.propagate_fft(w, 5e-6, shape=(511, 511), oversample=1)
    assert w.shape == (511, 511) and w.data[0].data.shape == (512, 512)
    psf = w.intensity
    psf /= psf.max()
    c = (511 - 1) / 2
    y, x = np.indices((511, 511), dtype=float)
    X = np.pi * np.hypot((x - c) * 5e-6, (y - c) * 5e-6) / (500e-9 * 10)
    X[X == 0] = np.finfo(float).eps
    airy = (2 * scipy.special.jn(1, X) / X) ** 2
    assert np.all(np.isclose(psf, airy / airy.max(), atol=1e-3))


def test_oracle_scratch_equals_pad():
    """With a scratch array the FFT input is every Field inserted into the
    zeroed scratch[:N0, :N1]; for one Field filling the wavefront it is the
    padded field."""
    amp = inputs.normalize_power(inputs.circle((64, 64), 30))
    p = optics.Pupil(amp, 1e-7 * np.random.default_rng(4).standard_normal((64, 64)), 1 / 64, 10)
    w = p.__rmul__(optics.Wavefront(600e-9))
    a = optics.propagate_fft(w, 2e-5, oversample=2)
    scratch = np.full((200, 200), 7 + 1j)
    b = optics.propagate_fft(w, 2e-5, oversample=2, scratch=scratch)
    n = a.data[0].data.shape
    assert rel_l2(b.data[0].data, a.data[0].data) < 1e-14
    assert np.array_equal(scratch[:n[0], :n[1]], optics.pad(w.field, n))
    assert np.all(scratch[n[0]:] == 7 + 1j)


def _pupil_wavefront(L, n=32, tilt=None):
    amp = inputs.circle((n, n), n // 2 - 1)
    p = L.Pupil(amplitude=amp, opd=np.zeros((n, n)), pixelscale=1 / n, focal_length=10)
    w = L.Wavefront.empty(650e-9, pixelscale=[1 / n, 1 / n], focal_length=10, shape=(n, n), ptype=L.pupil)
    w.data = [L.Field(data=np.ones((2, 2), dtype=complex), tilt=tilt)]
    return p, w


def test_errors_before_device_work():
    """Reference error behaviour (propagate.py:36-38, 56-58, 70-72, 263-272),
    raised before any device work."""
    import lentil_amd as L
    _, w = _pupil_wavefront(L, tilt=[L.Tilt(1e-6, 0)])
    with pytest.raises(NotImplementedError):
        L.propagate_fft(w, 5e-6)
    _, w = _pupil_wavefront(L)
    nmax = L.scratch_shape(650e-9, 1 / 32, 5e-6, 10, 2)
    with pytest.raises(ValueError):
        L.propagate_fft(w, 5e-6, shape=nmax[0] // 2 + 1, oversample=2)
    with pytest.raises(ValueError):
        L.propagate_fft(w, 5e-6, oversample=2, scratch=np.zeros(nmax, dtype=complex))
    w.ptype = L.none
    with pytest.raises(TypeError):
        L.propagate_fft(w, 5e-6)
