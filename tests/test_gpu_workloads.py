"""GPU parity of the batched plan on the segmented and large-sweep workloads
at their real sizes, against the complex128 oracle of the reference loop
(docs/user/diffraction.rst:148-160) on wavelength subsets the host finishes in
seconds, and through size-independent checks on the full sweeps.

Gate: rel-L2 <= 1e-5 on the PSF intensity (the workload gate)."""
import numpy as np
import pytest
import torch

from oracle import optics
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu

TOL_I = 1e-5


@pytest.fixture(scope='module')
def L():
    import lentil_amd
    assert torch.cuda.is_available()
    return lentil_amd


def _plan(L, cfg, wl, wt, precision, **kw):
    p = L.Pupil(amplitude=cfg['amp'], opd=cfg['opd'], mask=cfg['mask'], pixelscale=cfg['dx'],
                focal_length=cfg['focal_length'])
    return L.PSFPlan(p, wl, wt, pixelscale=cfg['pixelscale'], shape=cfg['shape'], oversample=cfg['oversample'],
                     tilts=[L.Tilt(*t) for t in cfg['tilts']], precision=precision, **kw)


def _oracle(cfg, wl, wt):
    p = optics.Pupil(cfg['amp'], cfg['opd'], cfg['dx'], cfg['focal_length'], mask=cfg['mask'])
    return optics.broadband_image(p, wl, wt, cfg['pixelscale'], cfg['shape'], cfg['oversample'],
                                  tilts=[optics.Tilt(*t) for t in cfg['tilts']])


@pytest.fixture(scope='module')
def seg18():
    from lentil_amd import workloads
    cfg = workloads.segmented_18()
    idx = np.arange(0, len(cfg['wavelengths']), 4)          # 8 of the 32 wavelengths
    wl, wt = cfg['wavelengths'][idx], cfg['weights'][idx]
    return cfg, wl, wt, _oracle(cfg, wl, wt)


@pytest.mark.parametrize('precision', ['fp32', 'bf16x3'])
@pytest.mark.parametrize('tile_n', [64, 128])
def test_segmented_18(L, seg18, precision, tile_n):
    """18 uneven hex tiles (one Field each) x 8 wavelengths in one coherent
    K-concatenated launch per stage."""
    cfg, wl, wt, ref = seg18
    plan = _plan(L, cfg, wl, wt, precision, tile_n=tile_n, lambda_group=4)
    assert plan.info['coherent_fast'] == 1
    assert plan.n_problems == 18 * len(wl)
    img = plan().double().cpu().numpy()
    err = rel_l2(img, ref)
    print(f'segmented_18 {precision} tile_n={tile_n}: rel-L2 {err:.2e}', flush=True)
    assert err < TOL_I


def test_segmented_18_full_sweep(L, seg18):
    """All 32 wavelengths (576 tile x wavelength MFTs in one launch per stage):
    the autotuned bf16x3 plan equals the fp32 plan within the gate."""
    from lentil_amd import workloads
    cfg = workloads.segmented_18()
    wl, wt = cfg['wavelengths'], cfg['weights']
    a = _plan(L, cfg, wl, wt, 'bf16x3', autotune=True)
    assert a.n_problems == 576
    b = _plan(L, cfg, wl, wt, 'fp32')
    ia, ib = a().double().cpu().numpy(), b().double().cpu().numpy()
    err = rel_l2(ia, ib)
    print(f'segmented_18 full: bf16x3 vs fp32 rel-L2 {err:.2e}; tuned {a.lambda_group}, {a.tile_n}', flush=True)
    assert err < TOL_I
    assert np.all(ia >= 0)


@pytest.fixture(scope='module')
def stress():
    from lentil_amd import workloads
    cfg = workloads.stress_2048()
    idx = np.array([0, len(cfg['wavelengths']) - 1])        # 450 and 950 nm
    wl, wt = cfg['wavelengths'][idx], cfg['weights'][idx]
    return cfg, wl, wt, _oracle(cfg, wl, wt)


@pytest.mark.parametrize('precision', ['fp32', 'bf16x3'])
def test_stress_2048_vs_oracle(L, stress, precision):
    """2048^2 pupil, 6 segment tiles, 4096^2 output with the off-axis
    (37.3, -12.6) px shift, at the two extreme wavelengths: phases of hundreds
    of turns, every tile merged coherently."""
    cfg, wl, wt, ref = stress
    plan = _plan(L, cfg, wl, wt, precision)
    img = plan().double().cpu().numpy()
    err = rel_l2(img, ref)
    print(f'stress_2048 {precision}: rel-L2 {err:.2e}', flush=True)
    assert err < TOL_I
    # the core sits at the commanded off-axis position
    r, c = np.unravel_index(np.argmax(img), img.shape)
    rr, cc = np.unravel_index(np.argmax(ref), ref.shape)
    assert (r, c) == (rr, cc)


def test_stress_2048_full_sweep(L):
    """All 128 wavelengths from a hipGraph: replays equal the eager result and
    the bf16x3 image equals the fp32 image within the gate."""
    from lentil_amd import workloads
    cfg = workloads.stress_2048()
    wl, wt = cfg['wavelengths'], cfg['weights']
    a = _plan(L, cfg, wl, wt, 'bf16x3', use_graph=True)
    out = torch.empty(a.shape_out, dtype=torch.float32, device='cuda')
    first = a(out).clone()          # captures the graph for this image
    for _ in range(2):
        out.fill_(-1.0)
        a(out)                      # replays
        torch.cuda.synchronize()
        assert torch.equal(out, first)
    b = _plan(L, cfg, wl, wt, 'fp32', use_graph=False)
    err = rel_l2(first.double().cpu().numpy(), b().double().cpu().numpy())
    print(f'stress_2048 full sweep: bf16x3 vs fp32 rel-L2 {err:.2e}', flush=True)
    assert err < TOL_I
    a.close()
    b.close()
