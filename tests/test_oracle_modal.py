"""Plane.fit_tilt / ptt_vector restatement in the oracle, pinned by closed
forms: an OPD built from known piston/tip/tilt per segment is fitted back
exactly, and a column sign error is caught."""
import numpy as np
import pytest

from oracle import inputs, optics


def _segments(n=64):
    m = np.zeros((3, n, n))
    m[0, 4:30, 4:30] = 1
    m[1, 34:60, 6:28] = 1
    m[2, 10:50, 36:60] = 1
    return m


def test_fit_tilt_recovers_known_tilts():
    n, dx = 64, 1 / 64
    r, c = np.meshgrid(np.arange(n) - n // 2, np.arange(n) - n // 2, indexing='ij')
    masks = _segments(n)
    truth = [(3e-7, 1e-6, -2e-6), (-1e-7, -3e-6, 5e-7), (2e-7, 4e-7, 1.5e-6)]
    opd = sum(m * (p + a * r * dx + b * (-c) * dx) for m, (p, a, b) in zip(masks, truth))
    new, tilts = optics.fit_tilt(opd, masks, dx)
    for (p, a, b), (ta, tb) in zip(truth, tilts):
        assert np.isclose(ta, a, rtol=1e-9) and np.isclose(tb, b, rtol=1e-9)
    piston = sum(m * p for m, (p, _, _) in zip(masks, truth))
    assert np.abs(new - piston).max() < 1e-15
    # one segment, circular mask
    amp = inputs.circle((n, n), 30)
    opd1 = amp * (2e-7 * r * dx - 7e-7 * (-c) * dx + 1e-8)
    new1, [(a1, b1)] = optics.fit_tilt(opd1, amp, dx)
    assert np.isclose(a1, 2e-7, rtol=1e-9) and np.isclose(b1, -7e-7, rtol=1e-9)


def test_planted_column_sign_is_caught():
    n, dx = 32, 1 / 32
    r, c = np.meshgrid(np.arange(n) - n // 2, np.arange(n) - n // 2, indexing='ij')
    mask = inputs.circle((n, n), 14)
    opd = mask * (1e-6 * (-c) * dx)
    v = optics.ptt_vector((n, n), dx, mask)
    v[2] *= -1                                    # the planted mistake: +c instead of -c
    t = np.linalg.lstsq(v.T, opd.ravel(), rcond=None)[0]
    assert not np.isclose(t[2], 1e-6)
    _, [(a, b)] = optics.fit_tilt(opd, mask, dx)
    assert np.isclose(b, 1e-6, rtol=1e-9)
