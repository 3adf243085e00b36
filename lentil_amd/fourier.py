"""Matrix Fourier transforms on the HIP device (lentil fourier.py).

``dft2`` and ``idft2`` keep lentil's signatures and semantics; the work runs
in ``mftx_dft2`` (two complex GEMMs with on-the-fly twiddles on MFMA).

Returned types follow the input: a NumPy input gives a complex128 NumPy
result (the drop-in contract), a device tensor gives a complex64 device
tensor.  ``out`` may be either kind and is returned.

The extra keyword-only ``precision`` selects 'fp32' (default) or 'bf16x3'.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._arrays import is_dev, to_dev, write_back

_default_precision = 'fp32'


def set_precision(precision):
    """Set the default precision of dft2/idft2/propagate_dft ('fp32' or 'bf16x3')."""
    global _default_precision
    _lib.precision_code(precision)
    _default_precision = precision


def get_precision():
    return _default_precision


def _mft(f, alpha, shape, shift, offset, scale, conj, out, precision):
    alpha_r, alpha_c = (float(v) for v in np.broadcast_to(alpha, (2,)))
    host_input = not is_dev(f)
    if host_input:
        f = np.asarray(f)
    if f.ndim != 2:
        raise ValueError(f'f must be 2-dimensional, got shape {tuple(f.shape)}')
    m, n = (int(v) for v in f.shape)
    if shape is None:
        shape = (m, n)
    M, N = (int(v) for v in np.broadcast_to(shape, (2,)))
    shift_r, shift_c = (float(v) for v in np.broadcast_to(shift, (2,)))
    off_r, off_c = (float(v) for v in np.broadcast_to(offset, (2,)))
    if out is not None:
        odt = out.dtype
        ok = odt.is_complex if isinstance(out, torch.Tensor) else np.can_cast(complex, odt)
        if not ok:
            raise TypeError(f"Cannot cast complex output to dtype('{odt}')")
    prec = _lib.precision_code(precision or _default_precision)
    src = to_dev(f)
    if src.dtype not in (torch.complex64, torch.float32, torch.complex128, torch.float64):
        src = src.to(torch.complex128 if src.is_complex() else torch.float64)
    src = src.contiguous()
    if is_dev(out) and out.dtype == torch.complex64 and out.dim() == 2 and out.stride(1) == 1 \
            and tuple(out.shape) == (M, N) and out.data_ptr() != src.data_ptr():
        F = out
    else:
        F = torch.empty((M, N), dtype=torch.complex64, device=src.device)
    ctx = _lib.context()
    rc = _lib.load().mftx_dft2(ctx, ctypes.c_void_p(src.data_ptr()), _lib.dtype_code(src), m, n, src.stride(0),
                               ctypes.c_void_p(F.data_ptr()), M, N, F.stride(0), alpha_r, alpha_c,
                               shift_r, shift_c, off_r, off_c, float(scale), 1 if conj else 0, prec)
    _lib.check(rc, ctx, 'mftx_dft2')
    if out is None:
        return F.cpu().numpy().astype(complex) if host_input else F
    if F is out:
        return out
    if is_dev(out):
        out.copy_(F)
    else:
        write_back(out, F)
    return out


def dft2(f, alpha, shape=None, shift=(0, 0), offset=(0, 0), unitary=True, out=None, *, precision=None):
    """2-D DFT by matrix triple product (fourier.py:5-103).

    F[u, v] = c * sum_{r, s} f[r, s] exp(-2 pi i (ar (R_r + or)(U_u - sr) + ac (S_s + oc)(V_v - sc)))
    with centred coordinates R = arange(m) - floor(m/2) (fourier.py:113-121)
    and c = sqrt(|ar ac|) when ``unitary``.
    """
    ar, ac = np.broadcast_to(alpha, (2,))
    scale = np.sqrt(np.abs(ar * ac)) if unitary else 1.0
    return _mft(f, alpha, shape, shift, offset, scale, False, out, precision)


def idft2(F, alpha, shape=None, shift=(0, 0), unitary=True, out=None, *, precision=None):
    """Inverse DFT (fourier.py:124-198): conj(dft2(conj(F))) / F.size.

    The division is by the size of the *input* ``F`` whatever ``shape`` is,
    as in the reference.  Implemented as conjugated twiddles on F directly.
    """
    size = int(F.numel()) if is_dev(F) else int(np.size(F))
    ar, ac = np.broadcast_to(alpha, (2,))
    scale = (np.sqrt(np.abs(ar * ac)) if unitary else 1.0) / size
    return _mft(F, alpha, shape, shift, (0, 0), scale, True, out, precision)


def dft2_adjoint(G, alpha, in_shape, shift=(0, 0), offset=(0, 0), unitary=True, *, precision=None):
    """Adjoint of f -> dft2(f, alpha, G.shape, shift, offset, unitary).

    E1^H . G . E2^H is an MFT from the output grid back to the input grid with
    conjugated twiddles, the output coordinates' shift acting as the input
    offset and vice versa.  No counterpart in the reference (it has no
    gradients); used by the autograd path.
    """
    ar, ac = np.broadcast_to(alpha, (2,))
    scale = np.sqrt(np.abs(ar * ac)) if unitary else 1.0
    sr, sc = np.broadcast_to(shift, (2,))
    orr, oc = np.broadcast_to(offset, (2,))
    m, n = in_shape
    M, N = tuple(G.shape)
    # input coords of the adjoint: U - s = (u - floor(M/2)) - s  -> offset -s
    # output coords:               R + o = (r - floor(m/2)) + o  -> shift  -o
    return _mft(G, alpha, (m, n), (-orr, -oc), (-sr, -sc), scale, True, None, precision)
