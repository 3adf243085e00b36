"""Pins the CPU oracle before it is trusted as the checker.

Each check is chosen so that a plausible mistake (a sign flip, a dropped
offset, a transposed operand, a wrong centre) fails at least one of them:
brute-force double sums, FFT equivalence at alpha = 1/n, Parseval, shift and
crop theorems, the Airy closed form, and the golden vectors written from the
reference for the behaviours nothing else pins.
"""
import numpy as np
import pytest
import scipy.special

from oracle import mft, optics, inputs
from tests.conftest import rel_l2


RNG = np.random.default_rng(1234)


def _cplx(*shape):
    return RNG.standard_normal(shape) + 1j * RNG.standard_normal(shape)


# ---------------------------------------------------------------- dft2/idft2

@pytest.mark.parametrize('m,n,M,N', [(3, 4, 5, 2), (4, 4, 4, 4), (5, 3, 2, 6), (1, 3, 3, 1)])
def test_dft2_matches_brute_force(m, n, M, N):
    f = _cplx(m, n)
    alpha = (0.137, -0.061)
    shift, offset = (0.3, -1.7), (2.0, -0.5)
    for unitary in (True, False):
        a = mft.dft2(f, alpha, (M, N), shift, offset, unitary)
        b = mft.dft2_direct(f, alpha, (M, N), shift, offset, unitary)
        assert np.allclose(a, b, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize('shape', [(10, 10), (11, 11), (10, 11)])
def test_dft2_fft_equivalence(shape):
    # tests/test_fourier.py:7-14, 27-34, 104-111
    m, n = shape
    f = _cplx(m, n)
    F = mft.dft2(f, [1 / m, 1 / n], unitary=False)
    assert np.allclose(F, np.fft.fftshift(np.fft.fft2(np.fft.ifftshift(f))), rtol=1e-10, atol=1e-10)


@pytest.mark.parametrize('n', [10, 11])
def test_idft2_ifft_equivalence(n):
    # tests/test_fourier.py:17-24, 37-44
    F = _cplx(n, n)
    f = mft.idft2(F, 1 / n, unitary=False)
    assert np.allclose(f, np.fft.fftshift(np.fft.ifft2(np.fft.ifftshift(F))), rtol=1e-10, atol=1e-10)


def test_round_trip_and_parseval():
    # tests/test_fourier.py:47-65
    f = _cplx(10, 10)
    assert np.allclose(mft.idft2(mft.dft2(f, 0.1, unitary=False), 0.1, unitary=False), f)
    F = mft.dft2(f, 0.1, unitary=True)
    assert np.isclose(np.sum(np.abs(f) ** 2), np.sum(np.abs(F) ** 2))


def test_integer_shift_is_roll():
    n = 16
    f = _cplx(n, n)
    F0 = mft.dft2(f, 1 / n, unitary=False)
    F1 = mft.dft2(f, 1 / n, shift=(3, -5), unitary=False)
    assert np.allclose(F1, np.roll(F0, (3, -5), axis=(0, 1)))


def test_shift_moves_centroid():
    # tests/test_fourier.py:68-78
    n = 100
    F = mft.dft2(np.ones((n, n)), 1 / n, shift=(7, -12))
    r, c = inputs.centroid(np.abs(F) ** 2)
    assert np.allclose((r - 50, c - 50), (7, -12))


def test_crop_offset_invariance():
    # tests/test_fourier.py:81-94
    n, m = 10, 3
    f = np.zeros((n, n), dtype=complex)
    f[2:2 + m, 5:5 + m] = _cplx(m, m)
    slc = optics.boundary_slice(np.abs(f))
    off = optics.slice_offset(slc, f.shape)
    assert np.allclose(mft.dft2(f, 1 / m, shape=10), mft.dft2(f[slc], 1 / m, shape=10, offset=off))


def test_separable_input_reduces_to_1d():
    a, b = _cplx(7), _cplx(5)
    f = np.outer(a, b)
    alpha = (0.07, 0.11)
    F = mft.dft2(f, alpha, (6, 9), (0.25, -1.0), (1.0, 0.5), unitary=False)
    u = np.arange(6) - 3 - 0.25
    v = np.arange(9) - 4 + 1.0
    r = np.arange(7) - 3 + 1.0
    s = np.arange(5) - 2 + 0.5
    Fa = np.exp(-2j * np.pi * alpha[0] * np.outer(u, r)) @ a
    Fb = np.exp(-2j * np.pi * alpha[1] * np.outer(v, s)) @ b
    assert np.allclose(F, np.outer(Fa, Fb))


def test_out_alias():
    # tests/test_fourier.py:97-101
    f = RNG.standard_normal((10, 10)).astype(complex)
    F = mft.dft2(f, 0.1, out=f)
    assert F is f
    with pytest.raises(TypeError):
        mft.dft2(np.ones((4, 4)), 0.25, out=np.zeros((4, 4)))


def test_idft2_golden(golden_small):
    g = golden_small
    for i in range(4):
        k = f'idft2_{i}_'
        out = mft.idft2(g[k + 'in'], g[k + 'alpha'], shape=tuple(g[k + 'shape']),
                        shift=g[k + 'shift'], unitary=bool(g[k + 'unitary']))
        assert rel_l2(out, g[k + 'out']) < 1e-12


def test_adjoint_dot_product():
    f = _cplx(6, 7)
    G = _cplx(9, 8)
    alpha, shift, offset = (0.05, 0.09), (0.4, -2.2), (1.0, -3.0)
    Ff = mft.dft2(f, alpha, G.shape, shift, offset)
    g = mft.dft2_adjoint(G, alpha, f.shape, shift, offset)
    assert np.isclose(np.vdot(G, Ff), np.vdot(g, f), rtol=1e-12)


# -------------------------------------------------- planted mistakes fail pins

def _dft2_wrong_sign(f, alpha, shape=None, shift=(0, 0), offset=(0, 0), unitary=True):
    return np.conj(mft.dft2(np.conj(f), alpha, shape, shift, offset, unitary))


def _dft2_dropped_offset(f, alpha, shape=None, shift=(0, 0), offset=(0, 0), unitary=True):
    return mft.dft2(f, alpha, shape, shift, (0, 0), unitary)


def _dft2_transposed(f, alpha, shape=None, shift=(0, 0), offset=(0, 0), unitary=True):
    return mft.dft2(np.asarray(f).T, alpha, shape, shift, offset, unitary)


@pytest.mark.parametrize('bad', [_dft2_wrong_sign, _dft2_dropped_offset, _dft2_transposed])
def test_planted_mistakes_are_caught(bad):
    f = _cplx(4, 4)
    args = ((0.137, 0.061), (4, 4), (0.3, -1.7), (2.0, -0.5))
    good = mft.dft2_direct(f, *args)
    assert not np.allclose(bad(f, *args), good, rtol=1e-6, atol=1e-9)


# ---------------------------------------------------------------- fields

@pytest.mark.parametrize('a_data, a_offset, b_data, b_offset, data, offset', [
    (1, [0, 0], 1, [0, 0], 1 + 0j, [0, 0]),
    (1, [0, 0], 1, [1, 0], [], [0, 0]),
    (np.ones((2, 2)), [0, 0], np.ones((2, 2)), [0, 0], np.ones((2, 2)), [0, 0]),
    (np.ones((2, 2)), [0, 0], np.ones((2, 2)), [4, 4], [], [0, 0]),
    (1, [0, 0], np.ones((3, 3)), [-5, -5], np.ones((3, 3)), [-5, -5]),
    (np.ones((5, 4)), [-2, -2], np.ones((3, 3)), [0, -1], np.ones((2, 2)), [0, -1]),
])
def test_field_multiply_table(a_data, a_offset, b_data, b_offset, data, offset):
    # tests/test_field.py:7-21
    c = optics.Field(a_data, 1, a_offset) * optics.Field(b_data, 1, b_offset)
    assert np.array_equal(c.data, data)
    assert np.array_equal(c.offset, offset)


def test_merge_known_answer():
    # tests/test_field.py:30-40
    c = optics.merge(optics.Field(np.ones((5, 4)), 1, [-2, -2]), optics.Field(np.ones((3, 3)), 1, [0, -1]))
    exp = np.zeros((6, 5))
    exp[0:5, 0:4] += 1
    exp[3:6, 2:5] += 1
    assert np.array_equal(c.data, exp)
    assert np.array_equal(c.offset, [-1, -2])


def test_reduce_golden(golden_small):
    g = golden_small
    for i in range(4):
        k = f'red_{i}_'
        fields = [optics.Field(g[k + f'in{j}_data'], 1, list(g[k + f'in{j}_offset']))
                  for j in range(int(g[k + 'nin']))]
        red = optics.reduce(fields)
        assert len(red) == int(g[k + 'nout'])
        for j, f in enumerate(red):
            assert np.array_equal(f.offset, g[k + f'out{j}_offset'])
            assert np.allclose(f.data, g[k + f'out{j}_data'], rtol=0, atol=1e-14)
        img = np.zeros((48, 48))
        for f in red:
            img = optics.insert(f, img, intensity=True, weight=0.5)
        assert np.allclose(img, g[k + 'intensity'], rtol=0, atol=1e-13)


# ---------------------------------------------------------------- propagate

def _golden_prop_case(g, i):
    k = f'prop_{i}_'
    dx, z, du, wl, os_ = g[k + 'params']
    pupil = optics.Pupil(g[k + 'amp'], g[k + 'opd'], dx, z)
    w = pupil.__rmul__(optics.Wavefront(wl))
    tilt = g[k + 'tilt']
    if not np.any(np.isnan(tilt)):
        w = optics.TiltPlane(optics.Tilt(x=tilt[0], y=tilt[1])).__rmul__(w)
    if (k + 'disp_trace') in g:
        w = optics.TiltPlane(optics.DispersiveTilt(g[k + 'disp_trace'], g[k + 'disp_dispersion'])).__rmul__(w)
    ps = tuple(g[k + 'prop_shape'])
    mask = g[k + 'mask'] if (k + 'mask') in g else None
    return w, dict(pixelscale=du, shape=tuple(g[k + 'shape']),
                   prop_shape=None if ps[0] < 0 else ps, oversample=int(os_), mask=mask)


@pytest.mark.parametrize('i', range(9))
def test_propagate_golden(golden_small, i):
    g = golden_small
    k = f'prop_{i}_'
    w, kw = _golden_prop_case(g, i)
    out = optics.propagate_dft(w, **kw)
    assert len(out.data) == int(g[k + 'nfields'])
    for j, f in enumerate(out.data):
        assert np.array_equal(np.asarray(f.offset), g[k + f'field{j}_offset'])
        ref = g[k + f'field{j}_data']
        assert f.data.shape == ref.shape
        assert rel_l2(f.data, ref) < 1e-10
    assert np.allclose(out.intensity, g[k + 'intensity'], rtol=1e-9, atol=1e-18)


def test_segmented_golden(golden_small):
    g = golden_small
    masks = g['seg_masks']
    tilts = [optics.Tilt(x=t[0], y=t[1]) for t in g['seg_tilts']]
    p = optics.Pupil(g['seg_amp'], g['seg_opd'], 1 / masks.shape[-1], 10, mask=masks, tilt=tilts)
    w = p.__rmul__(optics.Wavefront(700e-9))
    assert len(w.data) == int(g['seg_nphasor'])
    for j, f in enumerate(w.data):
        assert np.array_equal(np.asarray(f.offset), g[f'seg_phasor{j}_offset'])
        assert np.array_equal(f.data, g[f'seg_phasor{j}_data'])
    w2 = optics.propagate_dft(w, pixelscale=5e-6, shape=20, oversample=2)
    for j, f in enumerate(w2.data):
        assert np.array_equal(np.asarray(f.offset), g[f'seg_field{j}_offset'])
        assert rel_l2(f.data, g[f'seg_field{j}_data']) < 1e-10
    assert rel_l2(w2.intensity, g['seg_intensity']) < 1e-10
    img = np.zeros((40, 40))
    for k, wl in enumerate([650e-9, 720e-9, 810e-9]):
        w = p.__rmul__(optics.Wavefront(wl))
        w = optics.propagate_dft(w, pixelscale=5e-6, shape=20, oversample=2)
        img = w.insert(img, weight=0.2 + 0.1 * k)
    assert rel_l2(img, g['seg_broadband']) < 1e-10


def test_airy_closed_form():
    # tests/test_propagate.py:26-42 with tests/fixtures/airy2.py
    amp = inputs.normalize_power(inputs.circle((512, 512), 250))
    p = optics.Pupil(amp, 0, 1 / 500, 10)
    w = optics.propagate_dft(p.__rmul__(optics.Wavefront(650e-9)), 5e-6, shape=511, oversample=1)
    psf = w.intensity
    psf /= psf.max()
    c = (511 - 1) / 2
    y, x = np.indices((511, 511), dtype=float)
    q = np.hypot((x - c) * 5e-6, (y - c) * 5e-6)
    X = np.pi * q / (650e-9 * 10 / (500 / 500))
    X[X == 0] = np.finfo(float).eps
    airy = (2 * scipy.special.jn(1, X) / X) ** 2
    assert np.all(np.isclose(psf, airy / airy.max(), atol=1e-3))


def test_mask_equals_masked_psf():
    # tests/test_propagate_mask.py:4-37
    amp = inputs.normalize_power(inputs.circle((256, 256), 120))
    p = optics.Pupil(amp, 0, 1 / 240, 10)
    w = p.__rmul__(optics.Wavefront(650e-9))
    psf = optics.propagate_dft(w, 5e-6, shape=128, oversample=2).intensity
    mask = inputs.rectangle((256, 256), 128, 128, shift=(20, -30), antialias=False)
    w = p.__rmul__(optics.Wavefront(650e-9))
    psf_m = optics.propagate_dft(w, 5e-6, shape=128, oversample=2, mask=mask).intensity
    assert np.allclose(psf_m, psf * mask)


# ---------------------------------------------------------------- inputs

def test_inputs_golden(golden_small):
    g = golden_small
    assert np.array_equal(inputs.circle((64, 61), 23.3, shift=(1.5, -2.25)), g['in_circle'])
    assert np.array_equal(inputs.hexagon((64, 64), 25, shift=(-3, 2), rotate=True), g['in_hexagon'])
    assert np.array_equal(inputs.hexagon((63, 64), 21, shift=(1, 0.5)), g['in_hexagon_norot'])
    assert np.array_equal(inputs.hex_segments(rings=2, seg_radius=8, seg_gap=1, drop=(0, 5)), g['in_hexseg'])
    assert np.allclose(inputs.zernike_compose(g['in_zern_amp'], g['in_zern_coeffs']), g['in_zern_opd'],
                       rtol=0, atol=1e-20)
    assert np.allclose(inputs.normalize_power(g['in_zern_amp']), g['in_normpow'], rtol=1e-15)


def test_hexagon_boundary_known_answer():
    # tests/test_util.py:7-10
    assert list(optics.bbox(inputs.hexagon((256, 256), 100, rotate=True))) == [28, 228, 41, 215]


def test_c1_golden(golden_c1):
    g = golden_c1
    amp = inputs.normalize_power(inputs.circle((256, 256), 128))
    opd = inputs.zernike_compose(amp, g['c1_coeffs'])
    assert np.allclose(opd, g['c1_opd'], rtol=0, atol=1e-22)
    p = optics.Pupil(amp, opd, 1 / 256, 10)
    psf = optics.propagate_dft(p.__rmul__(optics.Wavefront(650e-9)), 5e-6, shape=128, oversample=2).intensity
    assert rel_l2(psf, g['c1_psf']) < 1e-12
