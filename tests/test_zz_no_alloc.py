"""mftx_dft2_noalloc against mftx_dft2 through the C ABI: the same transform
on the library's fixed scratch array gives the same bits as on the growable
context workspace, for two calls in a row (the second smaller, so it runs on
what the first left in the array), and sizes beyond the array are refused
with a clear error before anything is launched."""
import ctypes

import numpy as np
import pytest
import torch

from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 32 << 20


def _call(lib, name, ctx, f, F, shape_in, shape_out, alpha, shift, offset, scale, conj, prec):
    m, n = shape_in
    M, N = shape_out
    return getattr(lib, name)(ctx, ctypes.c_void_p(f.data_ptr()), 0, m, n, f.stride(0), ctypes.c_void_p(F.data_ptr()),
                              M, N, F.stride(0), alpha[0], alpha[1], shift[0], shift[1], offset[0], offset[1], scale,
                              conj, prec)


@pytest.mark.parametrize('prec', [0, 1])
def test_noalloc_matches_dft2(prec):
    from lentil_amd import _lib
    from oracle import mft
    assert torch.cuda.is_available()
    lib = _lib.load()
    ctx = _lib.context()
    rng = np.random.default_rng(7 + prec)
    cases = [((200, 168), (256, 240), (1.3e-3, 1.1e-3), (3.25, -7.5), (1.0, -2.0), 0.5, 0),
             ((96, 72), (100, 90), (2.1e-3, 1.7e-3), (-1.5, 0.5), (0.0, 3.0), 2.0, 1)]
    for shape_in, shape_out, alpha, shift, offset, scale, conj in cases:
        print(f'noalloc prec {prec}: {shape_in} -> {shape_out}', flush=True)
        x = rng.standard_normal(shape_in) + 1j * rng.standard_normal(shape_in)
        f = torch.from_numpy(x.astype(np.complex64)).cuda()
        a = torch.full(shape_out, float('nan'), dtype=torch.complex64, device='cuda')
        b = torch.full(shape_out, float('nan'), dtype=torch.complex64, device='cuda')
        _lib.check(_call(lib, 'mftx_dft2', ctx, f, a, shape_in, shape_out, alpha, shift, offset, scale, conj, prec),
                   ctx, 'mftx_dft2')
        _lib.check(_call(lib, 'mftx_dft2_noalloc', ctx, f, b, shape_in, shape_out, alpha, shift, offset, scale, conj,
                         prec), ctx, 'mftx_dft2_noalloc')
        torch.cuda.synchronize()
        assert torch.equal(a, b)
        ref = mft.dft2(f.cpu().numpy().astype(np.complex128), alpha, shape_out, shift, offset, unitary=False)
        ref = (np.conj(mft.dft2(np.conj(f.cpu().numpy().astype(np.complex128)), alpha, shape_out, shift, offset,
                                unitary=False)) if conj else ref) * scale
        assert rel_l2(b.cpu().numpy(), ref) < (2e-6 if prec == 0 else 2e-5)


def test_noalloc_refuses_oversize():
    from lentil_amd import _lib
    lib = _lib.load()
    ctx = _lib.context()
    f = torch.zeros((8, 8), dtype=torch.complex64, device='cuda')
    F = torch.zeros((8, 8), dtype=torch.complex64, device='cuda')
    # 2048 x 2048 -> 2048 x 2048 needs 2 x 32 MiB of packed operands: more than the array
    rc = lib.mftx_dft2_noalloc(ctx, ctypes.c_void_p(f.data_ptr()), 0, 2048, 2048, 2048, ctypes.c_void_p(F.data_ptr()),
                               2048, 2048, 2048, 1e-3, 1e-3, 0.0, 0.0, 0.0, 0.0, 1.0, 0, 1)
    assert rc == _lib.EINVAL
    msg = lib.mftx_last_error(ctx).decode()
    assert 'workspace' in msg and str(SCRATCH_BYTES) in msg
    assert lib.mftx_dft2_noalloc(None, None, 0, 1, 1, 1, None, 1, 1, 1, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0,
                                 0) == _lib.EHANDLE
