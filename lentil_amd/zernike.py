"""Zernike polynomials (lentil zernike.py, Noll order).

``zernike``, ``zernike_compose`` and ``zernike_coordinates`` build OPD inputs
on the host; ``zernike_basis`` evaluates a mode stack on the device.
"""
from math import factorial

import numpy as np

from .helper import mesh
from .util import centroid


def noll_to_nm(j):
    """Noll index -> (radial degree n, azimuthal degree m) (zernike.py:348-391)."""
    if j < 1:
        raise ValueError('Noll index must be greater than zero')
    n = int(np.ceil((-1 + np.sqrt(1 + 8 * j)) / 2) - 1)
    if n == 0:
        return 0, 0
    pos = int(j - (n + 1) * (n + 2) / 2 - 1)
    ms = [1, 1] if n % 2 else [0]
    while len(ms) < n + 1:
        ms += [ms[-1] + 2, ms[-1] + 2]
    m = ms[pos]
    return n, (-m if j % 2 else m)


def radial(n, m, rho):
    """Radial polynomial R_n^m (zernike.py:96-109)."""
    n, m = abs(int(n)), abs(int(m))
    if (n - m) % 2:
        return 0
    out = np.zeros(np.shape(rho))
    for k in range((n - m) // 2 + 1):
        out += ((-1) ** k * factorial(n - k)
                / (factorial(k) * factorial((n + m) // 2 - k) * factorial((n - m) // 2 - k))) * rho ** (n - 2 * k)
    return out


def zernike_coordinates(mask, shift=None, rotate=0):
    """(rho, theta) centred on the mask centroid, rho = 1 at the edge (zernike.py:444-480)."""
    mask = np.asarray(mask, dtype=bool)
    if shift is None:
        cen = centroid(mask)
        shift = (cen[0] - mask.shape[0] / 2, cen[1] - mask.shape[1] / 2)
    rr, cc = mesh(mask.shape, shift)
    r = np.abs(rr + 1j * cc)
    rho = r / np.max(r * mask)
    rot = np.exp(1j * (90 - rotate) * np.pi / 180)
    theta = np.angle(-rr * rot + 1j * cc * rot)
    return rho, theta


def zernike(mask, index, normalize=True, rho=None, theta=None):
    """Noll-ordered Zernike polynomial over a mask (zernike.py:11-93)."""
    mask = np.asarray(mask, dtype=bool)
    n, m = noll_to_nm(index)
    if rho is None:
        rho, theta = zernike_coordinates(mask)
    elif theta is None:
        raise ValueError('Both rho and theta must be specified')
    if m == 0:
        if n == 0:
            return mask
        return (np.sqrt(n + 1) if normalize else 1) * radial(n, m, rho) * mask
    c = np.sqrt(2) * np.sqrt(n + 1) if normalize else 1
    return c * radial(n, m, rho) * (np.cos(m * theta) if m > 0 else np.sin(m * theta)) * mask


def zernike_compose(mask, coeffs, normalize=True, rho=None, theta=None):
    """OPD = sum_j coeffs[j-1] * Z_j (zernike.py:112-180)."""
    mask = np.asarray(mask)
    opd = np.zeros(mask.shape)
    for i, c in enumerate(np.asarray(coeffs).ravel()):
        opd += c * zernike(mask, i + 1, normalize, rho, theta)
    return opd


def zernike_basis(mask, modes, vectorize=False, normalize=True, rho=None, theta=None, *, device=False):
    """Stack of Noll-ordered Zernike modes over a mask (zernike.py:183-234),
    evaluated on the device by ``mftx_zernike_basis`` (one thread per pixel
    and mode, the radial polynomial summed in the reference's order).

    NumPy float64 out by default; ``device=True`` returns the float64 device
    tensor (no host copy), e.g. as the modal basis of a phase retrieval."""
    import ctypes
    import torch
    from . import _lib
    from ._arrays import to_dev
    modes = np.atleast_1d(np.asarray(modes))
    mask = np.asarray(mask)
    if rho is None:
        rho, theta = zernike_coordinates(mask)
    elif theta is None:
        raise ValueError('Both rho and theta must be specified')
    rho = np.broadcast_to(np.asarray(rho, dtype=float), mask.shape)
    theta = np.broadcast_to(np.asarray(theta, dtype=float), mask.shape)
    nm = np.array([noll_to_nm(int(j)) for j in modes.ravel()], dtype=np.int32).reshape(-1, 2)
    npix = int(np.prod(mask.shape))
    rd = to_dev(np.ascontiguousarray(rho))
    td = to_dev(np.ascontiguousarray(theta))
    md = to_dev(np.ascontiguousarray(mask, dtype=np.float64))
    out = torch.empty((nm.shape[0], npix), dtype=torch.float64, device=rd.device)
    lib = _lib.load()
    ctx = _lib.context()
    _lib.check(lib.mftx_zernike_basis(ctx, ctypes.c_void_p(rd.data_ptr()), ctypes.c_void_p(td.data_ptr()),
                                      ctypes.c_void_p(md.data_ptr()), _lib.DT_F64, npix,
                                      nm.ctypes.data_as(ctypes.c_void_p), nm.shape[0], 1 if normalize else 0,
                                      ctypes.c_void_p(out.data_ptr())), ctx, 'mftx_zernike_basis')
    shape = (modes.shape[0], npix) if vectorize else modes.shape + mask.shape
    out = out.reshape(shape)
    return out if device else out.cpu().numpy()
