"""Complex128 phase-retrieval forward model and its exact gradient (oracle only).

No counterpart exists in the reference (lentil has no gradients).  The model
is the reference's broadband loop (docs/user/diffraction.rst:148-160) for a
monolithic pupil, one image per diversity OPD:

    f_{d,l} = amp * exp(2 pi i (opd + div_d) / l)             (plane.py:507)
    F_{d,l} = dft2(f_{d,l}, alpha_l, shape*os, unitary=True)   (fourier.py:79-103)
    I_d     = sum_l w_l |F_{d,l}|^2                            (field.py:302)

For an upstream gradient g_d = dL/dI_d (real), Wirtinger calculus gives
D = dL/d conj(f) = dft2_adjoint(w g F) and, with theta = 2 pi (opd+div)/l,

    dL/d opd = sum_{d,l} (2 pi / l) * (-2) Im(conj(D) f)
    dL/d amp = sum_{d,l} 2 Re(conj(D) exp(i theta))

``psf_stack_grad`` is pinned by complex128 central finite differences in
tests/test_oracle_retrieval.py.
"""
import numpy as np

from . import mft


def _alpha(dx, du, wl, z, os_):
    return (dx * du) / (wl * z * os_)


def psf_stack(amp, opd, diversity, wavelengths, weights, dx, du, z, npix, oversample):
    """Images I[d] (complex128 arithmetic), shape (len(diversity), npix*os, npix*os)."""
    M = npix * oversample
    out = np.zeros((len(diversity), M, M))
    for d, div in enumerate(diversity):
        for wl, wt in zip(wavelengths, weights):
            f = amp * np.exp(2j * np.pi * (opd + div) / wl)
            F = mft.dft2(f, _alpha(dx, du, wl, z, oversample), (M, M), unitary=True)
            out[d] += wt * np.abs(F) ** 2
    return out


def psf_stack_grad(amp, opd, diversity, wavelengths, weights, dx, du, z, npix, oversample, grad_images):
    """(dL/d opd, dL/d amp) for upstream dL/dI = grad_images (recomputing F)."""
    M = npix * oversample
    g_opd = np.zeros(opd.shape)
    g_amp = np.zeros(amp.shape)
    for d, div in enumerate(diversity):
        for wl, wt in zip(wavelengths, weights):
            a = _alpha(dx, du, wl, z, oversample)
            e = np.exp(2j * np.pi * (opd + div) / wl)
            f = amp * e
            F = mft.dft2(f, a, (M, M), unitary=True)
            D = mft.dft2_adjoint(wt * grad_images[d] * F, a, f.shape, unitary=True)
            g_opd += (2 * np.pi / wl) * (-2.0) * np.imag(np.conj(D) * f)
            g_amp += 2.0 * np.real(np.conj(D) * e)
    return g_opd, g_amp
