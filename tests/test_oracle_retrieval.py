"""Pins the oracle's phase-retrieval gradient by complex128 central finite
differences (the reference has no gradient to compare with)."""
import numpy as np

from oracle import inputs, retrieval


def _case():
    rng = np.random.default_rng(40)
    n = 24
    amp = inputs.normalize_power(inputs.circle((n, n), 10))
    opd = inputs.zernike_compose(amp, np.r_[0, rng.normal(0, 40e-9, 9)])
    defocus = inputs.zernike_compose(amp, [0, 0, 0, 1.0])
    div = [d * 650e-9 / (2 * np.sqrt(3)) * defocus for d in (-1, 1)]
    wl = np.array([600e-9, 700e-9])
    wt = np.array([0.4, 0.6])
    args = (div, wl, wt, 1 / 20, 5e-6, 10.0, 12, 2)
    target = retrieval.psf_stack(amp, opd + rng.normal(0, 5e-9, opd.shape) * (amp > 0), *args)
    return amp, opd, args, target


def _loss(amp, opd, args, target):
    I = retrieval.psf_stack(amp, opd, *args)
    return np.sum((I - target) ** 2)


def test_gradient_matches_finite_differences():
    amp, opd, args, target = _case()
    I = retrieval.psf_stack(amp, opd, *args)
    g_opd, g_amp = retrieval.psf_stack_grad(amp, opd, *args, 2 * (I - target))
    rng = np.random.default_rng(41)
    mask = amp > 0
    for _ in range(3):
        v = rng.standard_normal(opd.shape) * mask
        h = 1e-12
        fd = (_loss(amp, opd + h * v, args, target) - _loss(amp, opd - h * v, args, target)) / (2 * h)
        assert np.isclose(np.sum(g_opd * v), fd, rtol=1e-5, atol=0)
        va = rng.standard_normal(amp.shape) * mask
        ha = 1e-6
        fda = (_loss(amp + ha * va, opd, args, target) - _loss(amp - ha * va, opd, args, target)) / (2 * ha)
        assert np.isclose(np.sum(g_amp * va), fda, rtol=1e-5, atol=0)


def test_single_pixel_derivatives():
    amp, opd, args, target = _case()
    I = retrieval.psf_stack(amp, opd, *args)
    g_opd, _ = retrieval.psf_stack_grad(amp, opd, *args, 2 * (I - target))
    for (r, c) in [(12, 12), (8, 15), (14, 6)]:
        e = np.zeros(opd.shape)
        e[r, c] = 1.0
        h = 1e-11
        fd = (_loss(amp, opd + h * e, args, target) - _loss(amp, opd - h * e, args, target)) / (2 * h)
        assert np.isclose(g_opd[r, c], fd, rtol=1e-4, atol=1e-6 * abs(g_opd).max())
