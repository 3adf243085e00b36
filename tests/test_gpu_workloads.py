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
    wl, wt = cfg['wavelengths'], cfg['weights']             # all 32 wavelengths
    print('segmented_18: oracle image of 32 wavelengths', flush=True)
    return cfg, wl, wt, _oracle(cfg, wl, wt)


@pytest.mark.parametrize('precision', ['fp32', 'bf16x3'])
@pytest.mark.parametrize('tile_n', [64, 128])
@pytest.mark.parametrize('contraction', ['per_tile', 'merged'])
def test_segmented_18(L, seg18, precision, tile_n, contraction):
    """The full 32-wavelength image against the oracle, both ways of
    contracting the 18 coherent segments: 18 uneven hex tiles (one Field
    each) x 32 wavelengths = 576 MFTs in one K-concatenated launch per stage,
    or the tiles summed into their union box, one MFT per wavelength."""
    cfg, wl, wt, ref = seg18
    plan = _plan(L, cfg, wl, wt, precision, tile_n=tile_n, lambda_group=8, contraction=contraction)
    assert plan.info['coherent_fast'] == 1
    assert plan.merged == (contraction == 'merged')
    assert plan.n_problems == (1 if plan.merged else 18) * len(wl)
    img = plan().double().cpu().numpy()
    err = rel_l2(img, ref)
    print(f'segmented_18 {precision} tile_n={tile_n} {contraction}: rel-L2 {err:.2e}', flush=True)
    assert err < TOL_I


def test_segmented_18_autotuned(L, seg18):
    """The default plan (contraction chosen by cost, schedule by the
    autotuner) merges the segments and meets the gate on all 32 wavelengths;
    its cost model counts fewer MACs than the per-tile form."""
    cfg, wl, wt, ref = seg18
    a = _plan(L, cfg, wl, wt, 'bf16x3', autotune=True)
    assert a.merged and a.contraction_cost['merged'] < a.contraction_cost['per_tile']
    ia = a().double().cpu().numpy()
    err = rel_l2(ia, ref)
    print(f'segmented_18 autotuned: rel-L2 {err:.2e}; tuned {a.lambda_group}, {a.tile_n}', flush=True)
    assert err < TOL_I
    assert np.all(ia >= 0)


@pytest.fixture(scope='module')
def stress():
    from lentil_amd import workloads
    cfg = workloads.stress_2048()
    idx = np.array([0, 42, 85, len(cfg['wavelengths']) - 1])   # 450 nm to 950 nm
    wl, wt = cfg['wavelengths'][idx], cfg['weights'][idx]
    print('stress_2048: oracle image of 4 wavelengths', flush=True)
    return cfg, wl, wt, _oracle(cfg, wl, wt)


@pytest.mark.parametrize('precision', ['fp32', 'bf16x3'])
@pytest.mark.parametrize('contraction', ['per_tile', 'merged'])
def test_stress_2048_vs_oracle(L, stress, precision, contraction):
    """2048^2 pupil, 6 segment tiles, 4096^2 output with the off-axis
    (37.3, -12.6) px shift, at four wavelengths from 450 to 950 nm: phases of
    hundreds of turns, every tile merged coherently (per tile, or as one
    union-box MFT)."""
    cfg, wl, wt, ref = stress
    plan = _plan(L, cfg, wl, wt, precision, contraction=contraction)
    assert plan.merged == (contraction == 'merged')
    img = plan().double().cpu().numpy()
    err = rel_l2(img, ref)
    print(f'stress_2048 {precision}: rel-L2 {err:.2e}', flush=True)
    assert err < TOL_I
    # the core sits at the commanded off-axis position
    r, c = np.unravel_index(np.argmax(img), img.shape)
    rr, cc = np.unravel_index(np.argmax(ref), ref.shape)
    assert (r, c) == (rr, cc)


def test_stress_2048_full_sweep(L):
    """All 128 wavelengths from a hipGraph: replays equal the eager result,
    the merged bf16x3 image equals the per-tile fp32 image within the gate,
    and its 128 x 128 window around the off-axis core equals the oracle's
    image of all 128 wavelengths on that window (the reference loop with a
    detector mask, propagate.py:245-260, which evaluates only those pixels)."""
    from lentil_amd import workloads
    cfg = workloads.stress_2048()
    wl, wt = cfg['wavelengths'], cfg['weights']
    a = _plan(L, cfg, wl, wt, 'bf16x3', use_graph=True)
    assert a.merged
    out = torch.empty(a.shape_out, dtype=torch.float32, device='cuda')
    first = a(out).clone()          # captures the graph for this image
    for _ in range(2):
        out.fill_(-1.0)
        a(out)                      # replays
        torch.cuda.synchronize()
        assert torch.equal(out, first)
    b = _plan(L, cfg, wl, wt, 'fp32', use_graph=False, contraction='per_tile')
    img = first.double().cpu().numpy()
    err = rel_l2(img, b().double().cpu().numpy())
    print(f'stress_2048 full sweep: merged bf16x3 vs per-tile fp32 rel-L2 {err:.2e}', flush=True)
    assert err < TOL_I
    # the oracle on a window: every wavelength, only the pixels of the mask's box
    n, w = img.shape[0], 128
    r0, c0 = n // 2 + 37 - w // 2, n // 2 - 13 - w // 2   # the core sits at (+37.3, -12.6) px
    mask = np.zeros(img.shape)
    mask[r0:r0 + w, c0:c0 + w] = 1
    print('stress_2048 full sweep: oracle window of 128 wavelengths', flush=True)
    p = optics.Pupil(cfg['amp'], cfg['opd'], cfg['dx'], cfg['focal_length'], mask=cfg['mask'])
    ref = optics.broadband_image(p, wl, wt, cfg['pixelscale'], cfg['shape'], cfg['oversample'], mask=mask,
                                 tilts=[optics.Tilt(*t) for t in cfg['tilts']])
    win = (slice(r0, r0 + w), slice(c0, c0 + w))
    assert np.count_nonzero(ref) == w * w and np.argmax(ref[win]) == np.argmax(img[win])
    err = rel_l2(img[win], ref[win])
    print(f'stress_2048 full sweep: window vs oracle rel-L2 {err:.2e}', flush=True)
    assert err < TOL_I
    a.close()
    b.close()
