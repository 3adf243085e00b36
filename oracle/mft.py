"""Complex128 matrix Fourier transform (oracle; test infrastructure only).

Restates lentil/fourier.py of the reference:

* ``dft2``   follows fourier.py:79-103  (F = E1 . f . E2, unitary scale)
* twiddles   follow fourier.py:106-110 (phase = -2*pi*alpha*outer(x, u))
* coords     follow fourier.py:113-121 (arange(k) - floor(k/2))
* ``idft2``  follows fourier.py:193-198 (conj -> dft2 -> conj -> / F.size)

``dft2_direct`` is an independent brute-force double sum used to pin the
matrix form on small shapes; ``dft2_adjoint`` is the exact adjoint of the
linear map f -> dft2(f) used to check gradients (no counterpart in the
reference, which has no adjoint).
"""
import numpy as np


def coords(k):
    """Centred sample coordinates ``arange(k) - floor(k/2)`` (fourier.py:117-120)."""
    return np.arange(k) - np.floor(k / 2.0)


def _pair(v):
    a, b = np.broadcast_to(v, (2,))
    return a, b


def twiddles(m, n, M, N, alpha, shift=(0, 0), offset=(0, 0)):
    """Return the twiddle matrices (E1, E2) of shapes (M x m) and (n x N).

    E1[u, r] = exp(-2*pi*i*alpha_r*(R_r + off_r)*(U_u - shift_r))
    E2[s, v] = exp(-2*pi*i*alpha_c*(S_s + off_c)*(V_v - shift_c))
    (fourier.py:108-109).
    """
    ar, ac = _pair(alpha)
    sr, sc = _pair(shift)
    orr, oc = _pair(offset)
    R, S, U, V = coords(m), coords(n), coords(M), coords(N)
    E1 = np.exp(-2.0j * np.pi * ar * np.outer(R + orr, U - sr)).T
    E2 = np.exp(-2.0j * np.pi * ac * np.outer(S + oc, V - sc))
    return E1, E2


def dft2(f, alpha, shape=None, shift=(0, 0), offset=(0, 0), unitary=True, out=None):
    """Matrix triple product DFT (fourier.py:79-103)."""
    ar, ac = _pair(alpha)
    f = np.asarray(f)
    m, n = f.shape
    if shape is None:
        shape = (m, n)
    M, N = np.broadcast_to(shape, (2,))
    if out is not None and not np.can_cast(complex, out.dtype):
        raise TypeError(f"Cannot cast complex output to dtype('{out.dtype}')")
    E1, E2 = twiddles(m, n, M, N, (ar, ac), shift, offset)
    F = np.dot(E1.dot(f), E2, out=out)
    if unitary:
        np.multiply(F, np.sqrt(np.abs(ar * ac)), out=F)
    return F


def idft2(F, alpha, shape=None, shift=(0, 0), unitary=True, out=None):
    """Inverse via conjugation; divides by the INPUT size (fourier.py:193-198)."""
    F = np.asarray(F)
    size = F.size
    G = dft2(np.conj(F), alpha, shape, shift, unitary=unitary, out=out)
    np.conj(G, out=G)
    return np.divide(G, size, out=G)


def dft2_direct(f, alpha, shape=None, shift=(0, 0), offset=(0, 0), unitary=True):
    """Brute-force double sum over every (input, output) pair, for tiny shapes.

    F[u, v] = c * sum_r sum_s f[r, s] *
              exp(-2*pi*i*(ar*(R_r+or)*(U_u-sr) + ac*(S_s+oc)*(V_v-sc)))
    """
    ar, ac = _pair(alpha)
    sr, sc = _pair(shift)
    orr, oc = _pair(offset)
    f = np.asarray(f, dtype=complex)
    m, n = f.shape
    if shape is None:
        shape = (m, n)
    M, N = (int(v) for v in np.broadcast_to(shape, (2,)))
    out = np.zeros((M, N), dtype=complex)
    for u in range(M):
        yu = (u - (M // 2)) - sr
        for v in range(N):
            yv = (v - (N // 2)) - sc
            acc = 0j
            for r in range(m):
                xr = (r - (m // 2)) + orr
                for s in range(n):
                    xs = (s - (n // 2)) + oc
                    acc += f[r, s] * np.exp(-2j * np.pi * (ar * xr * yu + ac * xs * yv))
            out[u, v] = acc
    if unitary:
        out *= np.sqrt(abs(ar * ac))
    return out


def dft2_adjoint(G, alpha, in_shape, shift=(0, 0), offset=(0, 0), unitary=True):
    """Adjoint of f -> dft2(f, alpha, G.shape, shift, offset, unitary).

    Returns E1^H . G . E2^H (times the unitary scale), an m x n array with
    <dft2(f), G> == <f, dft2_adjoint(G)> for the Hermitian inner product.
    """
    ar, ac = _pair(alpha)
    m, n = in_shape
    M, N = G.shape
    E1, E2 = twiddles(m, n, M, N, (ar, ac), shift, offset)
    g = E1.conj().T.dot(G).dot(E2.conj().T)
    if unitary:
        g = g * np.sqrt(abs(ar * ac))
    return g
