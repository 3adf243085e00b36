"""GPU parity of the MFT kernels against the complex128 oracle.

Tolerances: the FP32 form is held to rel-L2 <= 2e-6 on complex fields; the
split-BF16x3 form to rel-L2 <= 2e-5 on complex fields; PSF intensities of
both forms to rel-L2 <= 1e-5 (the workload gate).
"""
import ctypes

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


def _many(L, entries, precision):
    lib, ctx = L._lib.load(), L._lib.context()
    arr = (L._lib.MftxMft * len(entries))(*entries)
    L._lib.check(lib.mftx_dft2_many(ctx, arr, len(entries), L._lib.precision_code(precision)), ctx,
                 'mftx_dft2_many')
    torch.cuda.synchronize()


def _mft_entry(L, F, alpha, shift, offset, scale, conj=0, f=None, amp=None, opd=None, inv_wl=0.0, w=None):
    q = L._lib.MftxMft()
    if f is not None:
        q.f, q.f_dtype, q.m, q.n, q.ldf = f.data_ptr(), L._lib.dtype_code(f), f.shape[0], f.shape[1], f.stride(0)
    else:
        q.amp, q.amp_dtype, q.ld_amp = amp.data_ptr(), L._lib.dtype_code(amp), amp.stride(0)
        q.opd, q.opd_dtype, q.ld_opd = opd.data_ptr(), L._lib.dtype_code(opd), opd.stride(0)
        q.m, q.n, q.inv_wl = amp.shape[0], amp.shape[1], inv_wl
    q.F, q.M, q.N, q.ldF = F.data_ptr(), F.shape[0], F.shape[1], F.stride(0)
    q.alpha_r, q.alpha_c = alpha
    q.shift_r, q.shift_c = shift
    q.offset_r, q.offset_c = offset
    q.scale, q.conj = scale, conj
    if w is not None:
        q.w, q.w_dtype, q.ldw = w.data_ptr(), L._lib.dtype_code(w), w.stride(0)
    return q


@pytest.mark.parametrize('precision', ['fp32', 'bf16x3'])
def test_dft2_many_vs_oracle(L, precision):
    """One grouped call over a mixed batch: every CASES geometry as complex128
    input with forward and conjugate twiddles, then phasor inputs of one pupil
    at four wavelengths with a float32 weight map (the entries of the
    phase-retrieval backward), each output against the complex128 oracle."""
    rng = np.random.default_rng(11)
    entries, refs, keep = [], [], []
    for i, (m, n, M, N, alpha, shift, offset) in enumerate(CASES):
        f = _cplx(rng, m, n)
        conj = i % 2
        fd = torch.from_numpy(f).cuda()
        F = torch.empty((M, N), dtype=torch.complex64, device='cuda')
        unit = float(np.sqrt(abs(alpha[0] * alpha[1])))
        entries.append(_mft_entry(L, F, alpha, shift, offset, unit, conj, f=fd))
        if conj:
            refs.append(np.conj(mft.dft2(np.conj(f), alpha, (M, N), shift, offset, unitary=True)))
        else:
            refs.append(mft.dft2(f, alpha, (M, N), shift, offset, unitary=True))
        keep += [fd, F]
    n = 64
    yy, xx = np.mgrid[:n, :n] - n / 2
    amp = (np.hypot(yy, xx) < n / 2 - 1).astype(float)
    opd = 3e-7 * (yy / n) ** 2 - 2e-7 * (xx / n) * (yy / n) + 1e-7 * rng.standard_normal((n, n)) * amp
    ad, od = torch.from_numpy(amp).cuda(), torch.from_numpy(opd).cuda()
    M = N = 96
    alpha = (0.0123, 0.0117)
    w = rng.uniform(0.2, 1.5, (M, N)).astype(np.float32)
    wd = torch.from_numpy(w).cuda()
    for wl in (500e-9, 600e-9, 700e-9, 800e-9):
        F = torch.empty((M, N), dtype=torch.complex64, device='cuda')
        entries.append(_mft_entry(L, F, alpha, (0.7, -1.3), (0.0, 0.0), 0.37, amp=ad, opd=od, inv_wl=1 / wl, w=wd))
        f = amp * np.exp(2j * np.pi * opd / wl)
        refs.append(0.37 * w * mft.dft2(f, alpha, (M, N), (0.7, -1.3), (0, 0), unitary=False))
        keep.append(F)
    # float32 amplitude and OPD (the phasor kernel's other load paths; a
    # phasor op of their own, since the dtypes differ)
    amp32, opd32 = amp.astype(np.float32), opd.astype(np.float32)
    ad32, od32 = torch.from_numpy(amp32).cuda(), torch.from_numpy(opd32).cuda()
    keep += [ad32, od32]
    for wl in (550e-9, 650e-9):
        F = torch.empty((M, N), dtype=torch.complex64, device='cuda')
        entries.append(_mft_entry(L, F, alpha, (0.7, -1.3), (0.0, 0.0), 0.37, amp=ad32, opd=od32, inv_wl=1 / wl,
                                  w=wd))
        f = amp32.astype(float) * np.exp(2j * np.pi * opd32.astype(float) / wl)
        refs.append(0.37 * w * mft.dft2(f, alpha, (M, N), (0.7, -1.3), (0, 0), unitary=False))
        keep.append(F)
    _many(L, entries, precision)
    outs = [t for t in keep if t.dtype == torch.complex64]
    for i, (F, ref) in enumerate(zip(outs, refs)):
        err = rel_l2(F.cpu().numpy(), ref)
        print(f'dft2_many {precision} entry {i}: rel-L2 {err:.2e}', flush=True)
        assert err < TOL_FIELD[precision]


def test_dft2_many_split_batch(L):
    """2000 small MFTs of assorted shapes: more descriptors than one upload
    slot holds, so the call splits the batch; every output against the oracle."""
    rng = np.random.default_rng(12)
    entries, data = [], []
    for i in range(2000):
        m, n, M, N = (int(v) for v in rng.integers(3, 13, 4))
        alpha = tuple(rng.uniform(0.02, 0.2, 2))
        shift = tuple(rng.uniform(-2, 2, 2))
        f = _cplx(rng, m, n)
        fd = torch.from_numpy(f).cuda()
        F = torch.empty((M, N), dtype=torch.complex64, device='cuda')
        entries.append(_mft_entry(L, F, alpha, shift, (0.0, 0.0), 1.0, f=fd))
        data.append((f, alpha, shift, fd, F))
    _many(L, entries, 'fp32')
    worst = 0.0
    for f, alpha, shift, _, F in data:
        ref = mft.dft2(f, alpha, F.shape, shift, unitary=False)
        worst = max(worst, rel_l2(F.cpu().numpy(), ref))
    print(f'dft2_many split batch: worst rel-L2 {worst:.2e}', flush=True)
    assert worst < TOL_FIELD['fp32']


def test_phase_grad_many(L):
    """The multi-wavelength phase gradient equals the sum of the complex128
    per-wavelength formula."""
    rng = np.random.default_rng(13)
    m, n, nw = 40, 56, 5
    amp = rng.uniform(0, 1, (m, n))
    opd = rng.normal(0, 2e-7, (m, n))
    wls = np.linspace(500e-9, 800e-9, nw)
    D = (rng.standard_normal((nw, m, n)) + 1j * rng.standard_normal((nw, m, n))).astype(np.complex64)
    g0 = rng.standard_normal((m, n))
    ga0 = rng.standard_normal((m, n))
    lib, ctx = L._lib.load(), L._lib.context()
    ad, od, Dd = torch.from_numpy(amp).cuda(), torch.from_numpy(opd).cuda(), torch.from_numpy(D).cuda()
    g, ga = torch.from_numpy(g0.copy()).cuda(), torch.from_numpy(ga0.copy()).cuda()
    wl_arr = (ctypes.c_double * nw)(*wls)
    L._lib.check(lib.mftx_phase_grad_many(ctx, ad.data_ptr(), 3, n, od.data_ptr(), 3, n, m, n, wl_arr, nw,
                                          Dd.data_ptr(), n, m * n, g.data_ptr(), n, ga.data_ptr(), n),
                 ctx, 'mftx_phase_grad_many')
    torch.cuda.synchronize()
    e = np.exp(2j * np.pi * opd[None] / wls[:, None, None])
    cd = np.conj(D.astype(complex)) * e
    ref_g = g0 + np.sum(-4 * np.pi / wls[:, None, None] * amp[None] * cd.imag, axis=0)
    ref_ga = ga0 + np.sum(2 * cd.real, axis=0)
    assert rel_l2(g.cpu().numpy() - g0, ref_g - g0) < 1e-6
    assert rel_l2(ga.cpu().numpy() - ga0, ref_ga - ga0) < 1e-6


def test_mft_plan_repeats(L):
    """A persistent MFT plan (mftx_mft_plan_*) over a mixed list equals the
    oracle on every execute, with the inputs' contents changed in between."""
    lib, ctx = L._lib.load(), L._lib.context()
    rng = np.random.default_rng(14)
    cases = [(40, 33, 50, 47, (0.011, 0.013), (0.5, -1.5), (1.0, 2.0)), (64, 64, 96, 96, (1 / 96, 1 / 96), (0, 0), (0, 0))]
    fs, Fs, entries = [], [], []
    for i, (m, n, M, N, alpha, shift, offset) in enumerate(cases):
        fd = torch.empty((m, n), dtype=torch.complex128, device='cuda')
        F = torch.empty((M, N), dtype=torch.complex64, device='cuda')
        entries.append(_mft_entry(L, F, alpha, shift, offset, 1.0, i % 2, f=fd))
        fs.append(fd)
        Fs.append(F)
    arr = (L._lib.MftxMft * len(entries))(*entries)
    h = ctypes.c_void_p()
    L._lib.check(lib.mftx_mft_plan_create(ctx, arr, len(entries), L._lib.precision_code('fp32'), ctypes.byref(h)),
                 ctx, 'mftx_mft_plan_create')
    try:
        for rep in range(3):
            hosts = []
            for fd in fs:
                x = _cplx(rng, *fd.shape)
                fd.copy_(torch.from_numpy(x))
                hosts.append(x)
            L._lib.check(lib.mftx_mft_plan_execute(h), ctx, 'mftx_mft_plan_execute')
            torch.cuda.synchronize()
            for i, ((m, n, M, N, alpha, shift, offset), x, F) in enumerate(zip(cases, hosts, Fs)):
                if i % 2:
                    ref = np.conj(mft.dft2(np.conj(x), alpha, (M, N), shift, offset, unitary=False))
                else:
                    ref = mft.dft2(x, alpha, (M, N), shift, offset, unitary=False)
                assert rel_l2(F.cpu().numpy(), ref) < TOL_FIELD['fp32'], (rep, i)
    finally:
        lib.mftx_mft_plan_destroy(h)
