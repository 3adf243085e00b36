"""GPU parity of the detector-grid operators against the oracle
(numpy FFT and scipy.ndimage as the reference calls them).

Tolerances: the FFT filters run two FP32 MFTs: rel-L2 <= 1e-5 (the workload
gate; measured about 1e-6).  rebin and rescale compute in FP64: rebin equals
the oracle to rounding, rescale within 1e-11 relative (float64 images, every
boundary mode and spline order 0-5, real and complex) and
FP32 rounding (float32 images).  The reference's own assertions
(tests/test_detector.py:86-97, test_convolvable.py:5-36, test_util.py:13-36)
are repeated on the device path.
"""
import numpy as np
import pytest
import torch

from oracle import detector as D, inputs, optics
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def L():
    import lentil_amd
    assert torch.cuda.is_available()
    return lentil_amd


def _psf(n, seed=0):
    """A PSF-like image (bright core, faint wings) plus a little background."""
    amp = inputs.circle((n // 2, n // 2), n // 5)
    f = np.fft.fftshift(np.fft.fft2(optics.pad(amp, (n, n))))
    img = np.abs(f) ** 2
    return img / img.sum() + 1e-7 * np.random.default_rng(seed).random((n, n))


@pytest.mark.parametrize('n', [64, 65, 128])
@pytest.mark.parametrize('oversample', [1, 2, 3])
def test_pixel(L, n, oversample):
    img = _psf(n, n + oversample)
    ref = D.pixel(img, oversample)
    got = L.detector.pixel(img, oversample)
    assert isinstance(got, np.ndarray) and got.dtype == np.float64 and got.shape == ref.shape
    assert rel_l2(got, ref) < 1e-5
    dev = L.detector.pixel(torch.tensor(img, dtype=torch.float32, device='cuda'), oversample)
    assert dev.dtype == torch.float32 and dev.is_cuda
    assert rel_l2(dev.double().cpu().numpy(), ref) < 1e-5


def test_pixel_reference_assertions(L):
    a = np.random.default_rng(1).uniform(low=0, high=1, size=(100, 100))
    b = L.detector.pixel(a, oversample=1)
    assert np.isclose(np.sum(a), np.sum(b))
    assert L.detector.pixelate(np.ones((10, 10)), 2).shape == (5, 5)
    with pytest.raises(ValueError):
        L.detector.pixel(np.ones((10, 12)), 2)      # the reference's MTF only broadcasts on square images


def test_detector_at_workload_grids(L):
    """The operators at the detector grids the workloads produce: the
    1024^2 broadband_512 image (pixelate and jitter at oversample 2) and the
    4096^2 stress_2048 image (pixel MTF and rebin at oversample 8), from a
    device tensor as a plan returns it, against the oracle."""
    img = _psf(1024, 11)
    dev = torch.tensor(img, dtype=torch.float32, device='cuda')
    print('detector 1024^2: pixelate, jitter', flush=True)
    got = L.detector.pixelate(dev, 2)
    assert got.shape == (512, 512)
    err = rel_l2(got.double().cpu().numpy(), D.pixelate(img, 2))
    assert err < 1e-5, err
    err = rel_l2(L.jitter(dev, 2.0, pixelscale=1, oversample=2).double().cpu().numpy(),
                 D.jitter(img, 2.0, oversample=2))
    assert err < 1e-5, err
    img = _psf(4096, 12)
    dev = torch.tensor(img, dtype=torch.float32, device='cuda')
    print('detector 4096^2: pixel, rebin', flush=True)
    err = rel_l2(L.detector.pixel(dev, 8).double().cpu().numpy(), D.pixel(img, 8))
    assert err < 1e-5, err
    got = L.rebin(dev, 8)
    assert tuple(got.shape) == (512, 512)
    err = rel_l2(got.double().cpu().numpy(), D.rebin(img, 8))
    assert err < 1e-6, err


@pytest.mark.parametrize('n', [64, 96])
def test_pixelate(L, n):
    img = _psf(n, 7)
    ref = D.pixelate(img, 2)
    got = L.detector.pixelate(img, 2)
    assert got.shape == ref.shape
    assert rel_l2(got, ref) < 1e-5


@pytest.mark.parametrize('shape', [(64, 64), (50, 70)])
def test_jitter_and_smear(L, shape):
    img = np.random.default_rng(2).random(shape) ** 4
    assert rel_l2(L.jitter(img, 1.3, pixelscale=1, oversample=2), D.jitter(img, 1.3, oversample=2)) < 1e-5
    assert rel_l2(L.smear(img, 3.0, angle=30, oversample=1), D.smear(img, 3.0, angle=30)) < 1e-5
    np.random.seed(5)
    a = L.smear(img, 2.5)
    np.random.seed(5)
    b = D.smear(img, 2.5)
    assert rel_l2(a, b) < 1e-5


def test_convolvable_reference_assertions(L):
    a = np.zeros((11, 11))
    a[5, 5] = 1
    b = L.jitter(a, 2)
    assert np.isclose(np.sum(a), np.sum(b))
    arr = b / np.max(b)
    arr[arr < 0.5] = 0
    ext = L.util.boundary(arr)
    assert np.max([ext[1] - ext[0], ext[3] - ext[2]]) // 2 == 2
    assert np.isclose(np.sum(a), np.sum(L.smear(a, 2)))


def test_rebin(L):
    img = np.array([[1, 1, 2, 2], [1, 1, 2, 2], [3, 3, 4, 4], [3, 3, 4, 4]])
    got = L.rebin(img, 2)
    assert np.array_equal(got, np.array([[4, 8], [12, 16]])) and got.dtype == np.zeros(1, img.dtype).sum().dtype
    cube = np.stack([img, 2 * img])
    assert np.array_equal(L.rebin(cube, 2), [[[4, 8], [12, 16]], [[8, 16], [24, 32]]])
    x = np.random.default_rng(3).random((3, 120, 96))
    assert np.abs(L.rebin(x, 4) - D.rebin(x, 4)).max() < 1e-13
    t = torch.tensor(x[0], dtype=torch.float32, device='cuda')
    assert np.abs(L.rebin(t, 3).double().cpu().numpy() - D.rebin(x[0].astype(np.float32), 3)).max() < 1e-5
    with pytest.raises(ValueError):
        L.rebin(img * 1j, 2)
    with pytest.raises(ValueError):
        L.rebin(np.ones((5, 4)), 2)


@pytest.mark.parametrize('scale', [0.5, 1 / 3, 0.25, 1.7, 2.0])
@pytest.mark.parametrize('order', [3, 1, 0])
def test_rescale(L, scale, order):
    rng = np.random.default_rng(int(scale * 100) + order)
    img = rng.random((61, 48))
    img[:5] = 0
    ref = D.rescale(img, scale, order=order)
    got = L.rescale(img, scale, order=order)
    assert got.shape == ref.shape and got.dtype == np.float64
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


def test_rescale_options(L):
    rng = np.random.default_rng(9)
    img = rng.random((40, 40))
    mask = (rng.random((40, 40)) > 0.3).astype(float)
    for kw in [dict(mask=mask), dict(unitary=False), dict(shape=30), dict(shape=(36, 44))]:
        ref = D.rescale(img, 0.5, **kw)
        got = L.rescale(img, 0.5, **kw)
        assert got.shape == ref.shape
        assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max(), kw
    f32 = img.astype(np.float32)
    ref = D.rescale(f32, 0.5)
    got = L.rescale(torch.tensor(f32, device='cuda'), 0.5)
    assert got.dtype == torch.float32
    assert rel_l2(got.double().cpu().numpy(), ref.astype(float)) < 1e-6
    a = np.random.default_rng(0).uniform(low=0, high=1, size=(100, 100))     # tests/test_util.py:32-36
    assert np.isclose(np.sum(a), np.sum(L.rescale(a, scale=0.5, order=3, mode='nearest', unitary=True)))
    # an order or mode scipy does not have raises scipy's RuntimeError
    with pytest.raises(RuntimeError):
        L.rescale(img, 0.5, order=6)
    with pytest.raises(RuntimeError):
        L.rescale(img, 0.5, mode='bogus')


@pytest.mark.parametrize('order', range(6))
@pytest.mark.parametrize('mode', ['nearest', 'reflect', 'grid-mirror', 'mirror', 'grid-wrap', 'wrap', 'constant',
                                  'grid-constant'])
def test_rescale_modes(L, mode, order):
    """Every boundary mode and spline order against the oracle (scipy's
    map_coordinates as util.rescale calls it), shrinking, growing, and with an
    output larger than the image so coordinates fall outside it."""
    rng = np.random.default_rng(17 * order + len(mode))
    img = rng.random((37, 52))
    img[:4] = 0
    for kw in (dict(scale=0.5), dict(scale=1.7), dict(scale=1.0, shape=(60, 75))):
        ref = D.rescale(img, order=order, mode=mode, **kw)
        got = L.rescale(img, order=order, mode=mode, **kw)
        assert got.shape == ref.shape and got.dtype == np.float64
        assert np.abs(got - ref).max() <= 1e-11 * np.abs(ref).max(), (mode, order, kw)


@pytest.mark.parametrize('order,mode', [(3, 'nearest'), (1, 'constant'), (5, 'reflect'), (2, 'grid-wrap')])
def test_rescale_complex(L, order, mode):
    """Complex images: real and imaginary parts resampled separately, the
    complex ratio of the sums applied, the mask taken from nonzero values
    (util.py:328-347); complex128 out for complex128 and complex64 in."""
    rng = np.random.default_rng(order)
    img = rng.random((44, 39)) + 1j * rng.standard_normal((44, 39))
    img[:3] = 0
    img[10, :] = 1j * rng.random(39)          # zero real part, nonzero values
    for kw in (dict(), dict(unitary=False), dict(shape=(50, 30))):
        ref = D.rescale(img, 0.6, order=order, mode=mode, **kw)
        got = L.rescale(img, 0.6, order=order, mode=mode, **kw)
        assert got.dtype == np.complex128 and got.shape == ref.shape
        assert np.abs(got - ref).max() <= 1e-11 * np.abs(ref).max(), kw
    c64 = torch.tensor(img.astype(np.complex64), device='cuda')
    got = L.rescale(c64, 0.6, order=order, mode=mode)
    assert got.dtype == torch.complex128
    ref = D.rescale(img.astype(np.complex64), 0.6, order=order, mode=mode)
    assert rel_l2(got.cpu().numpy(), ref) < 1e-6
