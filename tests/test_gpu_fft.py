"""GPU parity of propagate_fft (propagate.py:9-84) against the oracle's
complex128 FFT propagation.

Tolerances: complex fields rel-L2 <= 2e-6 (FP32) / 2e-5 (BF16x3); PSF
intensities rel-L2 <= 1e-5.
"""
import numpy as np
import pytest
import scipy.special
import torch

from oracle import inputs, optics
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu

TOL_FIELD = {'fp32': 2e-6, 'bf16x3': 2e-5}


@pytest.fixture(scope='module')
def L():
    import lentil_amd
    assert torch.cuda.is_available()
    return lentil_amd


def _pupils(L, n, radius, dx, f=10.0, seed=0):
    amp = inputs.normalize_power(inputs.circle((n, n), radius))
    rng = np.random.default_rng(seed)
    opd = inputs.zernike_compose(amp, np.r_[0, rng.normal(0, 30e-9, 10)])
    return (L.Pupil(amplitude=amp, opd=opd, pixelscale=dx, focal_length=f),
            optics.Pupil(amp, opd, dx, f))


# (pupil n, radius, dx, du, wavelength, oversample, shape): FFT grids
# N = round(1/alpha) even and odd, larger than the pupil grid (centred pad)
# or smaller (centred crop), default and explicit output shapes
CASES = [
    (128, 60, 1 / 128, 6.5e-6, 650e-9, 2, None),      # N = 256, even pupil in an even grid
    (128, 60, 1 / 128, 5e-6, 650e-9, 2, (100, 90)),   # N = 333 (odd), explicit shape
    (199, 95, 1 / 199, 9.25e-6, 600e-9, 2, None),     # N = 258, odd pupil in an even grid
    (300, 140, 1 / 300, 20e-6, 500e-9, 1, None),      # N = 75 < 300: centred crop, odd grid
]


@pytest.mark.parametrize('precision', ['fp32', 'bf16x3'])
@pytest.mark.parametrize('case', CASES)
def test_propagate_fft_vs_oracle(L, precision, case):
    n, radius, dx, du, wl, os_, shape = case
    print(f'propagate_fft {precision} {case}', flush=True)
    p, op = _pupils(L, n, radius, dx)
    w = L.propagate_fft(L.Wavefront(wl) * p, du, shape=shape, oversample=os_, precision=precision)
    ow = optics.propagate_fft(op.__rmul__(optics.Wavefront(wl)), du, shape=shape, oversample=os_)
    assert tuple(w.shape) == tuple(ow.shape)
    assert w.wavelength == ow.wavelength
    assert np.allclose(w.pixelscale, ow.pixelscale)
    got = w.data[0].data
    got = got.cpu().numpy() if hasattr(got, 'cpu') else np.asarray(got)
    ref = ow.data[0].data
    assert got.shape == ref.shape
    assert rel_l2(got, ref) < TOL_FIELD[precision]
    assert rel_l2(w.intensity, ow.intensity) < 1e-5


def test_propagate_fft_airy(L):
    """tests/test_propagate_fft.py:5-29, asserted."""
    amp = inputs.normalize_power(inputs.circle((512, 512), 256))
    p = L.Pupil(amplitude=amp, opd=np.zeros_like(amp), pixelscale=1 / 512, focal_length=10)
    w = L.Wavefront(wavelength=500e-9)
    w *= p
    w = L.propagate_fft(w, shape=(511, 511), pixelscale=5e-6, oversample=1)
    psf = w.intensity
    psf = psf / np.max(psf)
    c = (511 - 1) / 2
    y, x = np.indices((511, 511), dtype=float)
    X = np.pi * np.hypot((x - c) * 5e-6, (y - c) * 5e-6) / (500e-9 * 10)
    X[X == 0] = np.finfo(float).eps
    airy = (2 * scipy.special.jn(1, X) / X) ** 2
    assert np.all(np.isclose(psf, airy / airy.max(), atol=1e-3))


@pytest.mark.parametrize('device_scratch', [False, True])
def test_propagate_fft_scratch(L, device_scratch):
    """With scratch the Fields are inserted into the zeroed scratch[:N0, :N1]
    (left holding the padded input, the rest untouched) and the result equals
    the padded path."""
    p, op = _pupils(L, 96, 44, 1 / 96, seed=3)
    wl, du = 550e-9, 9e-6
    nshape = L.scratch_shape(wl, 1 / 96, du, 10, 2)
    assert nshape == optics.scratch_shape(wl, 1 / 96, du, 10, 2)
    host = np.full((nshape[0] + 3, nshape[1] + 5), 2 - 1j)
    scratch = torch.from_numpy(host.copy()).cuda() if device_scratch else host
    w = L.propagate_fft(L.Wavefront(wl) * p, du, oversample=2, scratch=scratch)
    oscr = host.copy()
    ow = optics.propagate_fft(op.__rmul__(optics.Wavefront(wl)), du, oversample=2, scratch=oscr)
    got = w.data[0].data.cpu().numpy()
    assert rel_l2(got, ow.data[0].data) < TOL_FIELD['fp32']
    s = scratch.cpu().numpy() if device_scratch else scratch
    # the padded input went through a complex64 insert: equal to FP32 rounding
    assert rel_l2(s[:nshape[0], :nshape[1]], oscr[:nshape[0], :nshape[1]]) < 1e-7
    assert np.all(s[nshape[0]:] == 2 - 1j) and np.all(s[:, nshape[1]:] == 2 - 1j)
