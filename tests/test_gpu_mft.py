"""GPU parity of the MFT kernels against the complex128 oracle.

Tolerances: the FP32 form is held to rel-L2 <= 2e-6 on complex fields; the
split-BF16x3 form to rel-L2 <= 2e-5 on complex fields; PSF intensities of
both forms to rel-L2 <= 1e-5 (the workload gate).
"""
import numpy as np
import pytest
import torch

from oracle import mft
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu

TOL_FIELD = {'fp32': 2e-6, 'bf16x3': 2e-5}


@pytest.fixture(scope='module')
def L():
    import lentil_amd
    assert torch.cuda.is_available()
    return lentil_amd


def _cplx(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


CASES = [
    # (m, n, M, N, alpha, shift, offset)
    (10, 10, 10, 10, (0.1, 0.1), (0, 0), (0, 0)),
    (11, 11, 11, 11, (1 / 11, 1 / 11), (0, 0), (0, 0)),
    (10, 11, 10, 11, (0.1, 1 / 11), (0, 0), (0, 0)),
    (7, 5, 9, 13, (0.071, -0.033), (0.3, -1.7), (2.0, -0.5)),
    (130, 77, 200, 333, (0.0021, 0.0043), (3.25, -7.5), (-4.0, 6.0)),
    (256, 256, 256, 256, (1 / 512, 1 / 512), (0.0, 0.0), (0.0, 0.0)),
    (161, 186, 257, 129, (0.0017, 0.0011), (-0.37, 0.21), (-210.0, 155.0)),
]


@pytest.mark.parametrize('precision', ['fp32', 'bf16x3'])
@pytest.mark.parametrize('case', CASES)
def test_dft2_vs_oracle(L, precision, case):
    m, n, M, N, alpha, shift, offset = case
    print(f'dft2 {precision} {case}', flush=True)
    rng = np.random.default_rng(m * 1000 + n)
    f = _cplx(rng, m, n)
    ref = mft.dft2(f, alpha, (M, N), shift, offset, unitary=True)
    got = L.fourier.dft2(f, alpha, (M, N), shift, offset, unitary=True, precision=precision)
    assert isinstance(got, np.ndarray) and got.dtype == np.complex128 and got.shape == (M, N)
    assert rel_l2(got, ref) < TOL_FIELD[precision]


@pytest.mark.parametrize('precision', ['fp32', 'bf16x3'])
def test_dft2_fft_equivalence(L, precision):
    # tests/test_fourier.py:7-14 on the device
    rng = np.random.default_rng(3)
    for shape in [(10, 10), (11, 11), (10, 11), (64, 48)]:
        f = _cplx(rng, *shape)
        F = L.fourier.dft2(f, [1 / shape[0], 1 / shape[1]], unitary=False, precision=precision)
        ref = np.fft.fftshift(np.fft.fft2(np.fft.ifftshift(f)))
        assert rel_l2(F, ref) < TOL_FIELD[precision]


def test_idft2_golden(L, golden_small):
    g = golden_small
    for i in range(4):
        k = f'idft2_{i}_'
        out = L.fourier.idft2(g[k + 'in'], g[k + 'alpha'], shape=tuple(g[k + 'shape']), shift=g[k + 'shift'],
                              unitary=bool(g[k + 'unitary']))
        assert rel_l2(out, g[k + 'out']) < TOL_FIELD['fp32']


def test_round_trip_parseval_out(L):
    rng = np.random.default_rng(5)
    f = _cplx(rng, 10, 10)
    F = L.fourier.dft2(f, 0.1, unitary=False)
    assert np.allclose(L.fourier.idft2(F, 0.1, unitary=False), f, rtol=1e-5, atol=1e-5)
    F = L.fourier.dft2(f, 0.1, unitary=True)
    assert np.isclose(np.sum(np.abs(F) ** 2), np.sum(np.abs(f) ** 2), rtol=1e-5)
    g = f.copy()
    out = L.fourier.dft2(g, 0.1, out=g)
    assert out is g
    assert rel_l2(g, mft.dft2(f, 0.1)) < TOL_FIELD['fp32']
    with pytest.raises(TypeError):
        L.fourier.dft2(f, 0.1, out=np.zeros((10, 10)))


def test_dft2_shift_centroid(L):
    # tests/test_fourier.py:68-78
    n = 100
    F = L.fourier.dft2(np.ones((n, n)), 1 / n, shift=(7, -12))
    r, c = L.util.centroid(np.abs(F) ** 2)
    assert abs(r - 57) < 1e-3 and abs(c - 38) < 1e-3


def test_dft2_device_tensors(L):
    rng = np.random.default_rng(6)
    f = _cplx(rng, 40, 30)
    tf = torch.from_numpy(f.astype(np.complex64)).cuda()
    out = torch.empty((50, 60), dtype=torch.complex64, device='cuda')
    res = L.fourier.dft2(tf, (0.013, 0.02), (50, 60), (1.5, -2.0), (3, -1), out=out)
    assert res is out
    ref = mft.dft2(f.astype(np.complex64).astype(complex), (0.013, 0.02), (50, 60), (1.5, -2.0), (3, -1))
    assert rel_l2(out.cpu().numpy(), ref) < TOL_FIELD['fp32']
    real = L.fourier.dft2(torch.from_numpy(f.real.copy()).cuda(), 0.02, (20, 20))
    assert rel_l2(real.cpu().numpy(), mft.dft2(f.real, 0.02, (20, 20))) < TOL_FIELD['fp32']


@pytest.mark.parametrize('precision', ['fp32', 'bf16x3'])
def test_adjoint(L, precision):
    rng = np.random.default_rng(7)
    f = _cplx(rng, 24, 31)
    G = _cplx(rng, 40, 36)
    alpha, shift, offset = (0.011, 0.017), (0.4, -2.2), (1.0, -3.0)
    g = L.fourier.dft2_adjoint(G, alpha, f.shape, shift, offset, precision=precision)
    ref = mft.dft2_adjoint(G, alpha, f.shape, shift, offset)
    assert rel_l2(g, ref) < TOL_FIELD[precision]
    Ff = L.fourier.dft2(f, alpha, G.shape, shift, offset, precision=precision)
    assert abs(np.vdot(G, Ff) - np.vdot(g, f)) / abs(np.vdot(G, Ff)) < 5 * TOL_FIELD[precision]


def test_large_phase_arguments(L):
    # |alpha * x * y| of hundreds of turns: the FP64 range reduction keeps the
    # twiddles accurate where an FP32 phase would not
    rng = np.random.default_rng(8)
    f = _cplx(rng, 64, 64)
    alpha, shift, offset = (0.013, 0.009), (1500.3, -2200.6), (700.0, -900.0)
    ref = mft.dft2(f, alpha, (96, 80), shift, offset)
    got = L.fourier.dft2(f, alpha, (96, 80), shift, offset)
    assert rel_l2(got, ref) < 5 * TOL_FIELD['fp32']
