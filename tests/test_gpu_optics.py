"""GPU parity of the drop-in optics layer (phasor, Field ops, propagate_dft,
Wavefront.intensity/insert) and of the batched PSFPlan against the oracle
and the golden vectors written from the reference."""
import numpy as np
import pytest
import torch

from oracle import optics, inputs
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu

TOL_I = 1e-5      # PSF intensity gate (rel-L2)


@pytest.fixture(scope='module')
def L():
    import lentil_amd
    assert torch.cuda.is_available()
    return lentil_amd


def _golden_case(L, g, i):
    k = f'prop_{i}_'
    dx, z, du, wl, os_ = g[k + 'params']
    p = L.Pupil(amplitude=g[k + 'amp'], opd=g[k + 'opd'], pixelscale=dx, focal_length=z)
    w = L.Wavefront(wl) * p
    tilt = g[k + 'tilt']
    if not np.any(np.isnan(tilt)):
        w = w * L.Tilt(x=tilt[0], y=tilt[1])
    if (k + 'disp_trace') in g:
        w = w * L.DispersiveTilt(trace=g[k + 'disp_trace'], dispersion=g[k + 'disp_dispersion'])
    ps = tuple(g[k + 'prop_shape'])
    mask = g[k + 'mask'] if (k + 'mask') in g else None
    return w, dict(pixelscale=du, shape=tuple(g[k + 'shape']), prop_shape=None if ps[0] < 0 else ps,
                   oversample=int(os_), mask=mask)


@pytest.mark.parametrize('i', range(9))
def test_propagate_golden(L, golden_small, i):
    g = golden_small
    k = f'prop_{i}_'
    print(f'propagate golden case {i}', flush=True)
    w, kw = _golden_case(L, g, i)
    out = L.propagate_dft(w, **kw)
    assert len(out.data) == int(g[k + 'nfields'])
    for j, f in enumerate(out.data):
        assert np.array_equal(np.asarray(f.offset), g[k + f'field{j}_offset'])
        assert rel_l2(f.data.cpu().numpy(), g[k + f'field{j}_data']) < 2e-6
    ref = g[k + 'intensity']
    got = out.intensity
    assert got.dtype == np.float64 and got.shape == ref.shape
    if np.any(ref):
        assert rel_l2(got, ref) < TOL_I
    else:
        assert np.all(got == 0)


def test_segmented_golden(L, golden_small):
    g = golden_small
    masks = g['seg_masks']
    p = L.Pupil(amplitude=g['seg_amp'], opd=g['seg_opd'], mask=masks, pixelscale=1 / masks.shape[-1],
                focal_length=10)
    p.tilt = [L.Tilt(x=t[0], y=t[1]) for t in g['seg_tilts']]
    w = L.Wavefront(700e-9) * p
    assert len(w.data) == int(g['seg_nphasor'])
    for j, f in enumerate(w.data):
        assert np.array_equal(np.asarray(f.offset), g[f'seg_phasor{j}_offset'])
        assert rel_l2(f.data.cpu().numpy(), g[f'seg_phasor{j}_data']) < 1e-6
    w2 = L.propagate_dft(w, pixelscale=5e-6, shape=20, oversample=2)
    for j, f in enumerate(w2.data):
        assert np.array_equal(np.asarray(f.offset), g[f'seg_field{j}_offset'])
        assert rel_l2(f.data.cpu().numpy(), g[f'seg_field{j}_data']) < 2e-6
    assert rel_l2(w2.intensity, g['seg_intensity']) < TOL_I
    img = np.zeros((40, 40))
    for k, wl in enumerate([650e-9, 720e-9, 810e-9]):
        w = L.propagate_dft(L.Wavefront(wl) * p, pixelscale=5e-6, shape=20, oversample=2)
        img = w.insert(img, weight=0.2 + 0.1 * k)
    assert rel_l2(img, g['seg_broadband']) < TOL_I
    # the batched plan reproduces the same broadband image
    for precision in ('fp32', 'bf16x3'):
        plan = L.PSFPlan(p, [650e-9, 720e-9, 810e-9], [0.2, 0.3, 0.4], pixelscale=5e-6, shape=20, oversample=2,
                         precision=precision, lambda_group=2)
        got = plan().double().cpu().numpy()
        assert rel_l2(got, g['seg_broadband']) < TOL_I, precision


def test_reduce_golden(L, golden_small):
    g = golden_small
    for i in range(4):
        k = f'red_{i}_'
        fields = [L.Field(torch.from_numpy(g[k + f'in{j}_data'].astype(np.complex64)).cuda(), 1,
                          list(g[k + f'in{j}_offset'])) for j in range(int(g[k + 'nin']))]
        red = L.field.reduce(fields)
        assert len(red) == int(g[k + 'nout'])
        for j, f in enumerate(red):
            assert np.array_equal(np.asarray(f.offset), g[k + f'out{j}_offset'])
            assert rel_l2(f.data.cpu().numpy(), g[k + f'out{j}_data']) < 1e-6
        img = np.zeros((48, 48))
        for f in red:
            img = L.field.insert(f, img, intensity=True, weight=0.5)
        assert rel_l2(img, g[k + 'intensity']) < 1e-6


def test_field_multiply_table(L):
    # tests/test_field.py:7-21 with array data on the device
    F = L.Field
    c = F(np.ones((5, 4)), 1, [-2, -2]) * F(np.ones((3, 3)) * 2, 1, [0, -1])
    assert np.array_equal(c.offset, [0, -1])
    assert np.allclose(c.data.cpu().numpy(), 2 * np.ones((2, 2)))
    c = F(1, 1, [0, 0]) * F(np.ones((3, 3)), 1, [-5, -5])
    data = c.data.cpu().numpy() if isinstance(c.data, torch.Tensor) else c.data
    assert np.array_equal(c.offset, [-5, -5]) and np.array_equal(data, np.ones((3, 3)))
    assert (F(np.ones((2, 2)), 1, [0, 0]) * F(np.ones((2, 2)), 1, [4, 4])).size == 0
    rng = np.random.default_rng(9)
    a, b = rng.standard_normal((6, 7)) + 1j, rng.standard_normal((4, 5)) - 2j
    c = F(a, 1, [1, 0]) * F(b, 1, [0, 2])
    ref = optics.Field(a, 1, [1, 0]) * optics.Field(b, 1, [0, 2])
    assert np.array_equal(np.asarray(c.offset), np.asarray(ref.offset))
    assert rel_l2(c.data.cpu().numpy(), ref.data) < 1e-6


def test_merge_known_answer(L):
    c = L.field.merge(L.Field(np.ones((5, 4)), 1, [-2, -2]), L.Field(np.ones((3, 3)), 1, [0, -1]))
    exp = np.zeros((6, 5))
    exp[0:5, 0:4] += 1
    exp[3:6, 2:5] += 1
    assert np.array_equal(c.data.cpu().numpy(), exp)
    assert np.array_equal(c.offset, [-1, -2])


def test_phasor_and_overlapping_segments(L):
    # tests/test_plane.py:119-133: overlapping segment slices give intensity == mask
    seg = L.hexagon((64, 64), 32, antialias=False)[5:60, :]
    segmask = np.zeros((2, 128, 128))
    segmask[0, 0:55, 2:66] = seg
    segmask[1, 29:84, 55:119] = seg
    mask = np.sum(segmask, axis=0)
    pupil = L.Pupil(amplitude=mask, mask=segmask, pixelscale=1 / 256, focal_length=10)
    w = L.Wavefront(500e-9)
    w *= pupil
    assert np.array_equal(mask, w.intensity)
    # phasor values vs amp*exp(2 pi i opd / lambda)
    rng = np.random.default_rng(10)
    amp, opd = rng.uniform(size=(10, 10)), rng.uniform(size=(10, 10)) * 1e-6
    p = L.Plane(pixelscale=1, amplitude=amp, opd=opd)
    w1 = L.Wavefront(650e-9) * p
    slc = L.helper.boundary_slice(p.mask)
    ref = amp[slc] * np.exp(2 * np.pi * 1j * opd[slc] / 650e-9)
    assert rel_l2(w1.data[0].data.cpu().numpy(), ref) < 1e-6


@pytest.mark.parametrize('coeffs,size', [(None, 128), ([0, 1e-6, 2e-6], 64)])
def test_mask_property(L, coeffs, size):
    """tests/test_propagate_mask.py:4-37 with the reference's own tolerance
    (np.allclose defaults, rtol 1e-5, atol 1e-8): the masked propagation equals
    the full one on the window.  The two are separate FP32 computations (their
    outputs are paired about different centres), so dim pixels differ at the
    FP32 rounding of the PSF core, which the reference's atol admits."""
    amp = L.normalize_power(L.circle((256, 256), 120))
    opd = 0 if coeffs is None else L.zernike_compose(amp, coeffs)
    p = L.Pupil(amplitude=amp, opd=opd, pixelscale=1 / 240, focal_length=10)
    psf = L.propagate_dft(L.Wavefront(650e-9) * p, 5e-6, shape=128, oversample=2).intensity
    mask = L.rectangle((256, 256), size, size, shift=(20, -30), antialias=False)
    psf_m = L.propagate_dft(L.Wavefront(650e-9) * p, 5e-6, shape=128, oversample=2, mask=mask).intensity
    assert np.allclose(psf_m, psf * mask)
    assert rel_l2(psf_m, psf * mask) < 1e-5


def test_airy(L):
    amp = L.normalize_power(L.circle((512, 512), 250))
    p = L.Pupil(amplitude=amp, pixelscale=1 / 500, focal_length=10)
    psf = L.propagate_dft(L.Wavefront(650e-9) * p, 5e-6, shape=511, oversample=1).intensity
    psf /= psf.max()
    ref = optics.propagate_dft(optics.Pupil(amp, 0, 1 / 500, 10).__rmul__(optics.Wavefront(650e-9)), 5e-6,
                               shape=511, oversample=1).intensity
    assert rel_l2(psf, ref / ref.max()) < TOL_I


@pytest.mark.parametrize('precision', ['fp32', 'bf16x3'])
def test_mono_256_golden(L, golden_c1, precision):
    from lentil_amd import workloads
    cfg = workloads.mono_256()
    # NumPy's vectorised sin/cos/pow may differ by a few ulp between host CPUs
    assert np.allclose(cfg['opd'], golden_c1['c1_opd'], rtol=0, atol=1e-19)
    p = L.Pupil(amplitude=cfg['amp'], opd=cfg['opd'], pixelscale=cfg['dx'], focal_length=10)
    psf = L.propagate_dft(L.Wavefront(650e-9) * p, 5e-6, shape=128, oversample=2, precision=precision).intensity
    err = rel_l2(psf, golden_c1['c1_psf'])
    print(f'mono_256 {precision} per-call rel-L2 {err:.3e}')
    assert err < TOL_I
    plan = L.PSFPlan(p, [650e-9], [1.0], pixelscale=5e-6, shape=128, oversample=2, precision=precision)
    img = plan().double().cpu().numpy()
    err = rel_l2(img, golden_c1['c1_psf'])
    print(f'mono_256 {precision} plan rel-L2 {err:.3e}')
    assert err < TOL_I


def test_plan_graph_replays(L):
    """Graph replays (with eager work between them) match the eager result."""
    from lentil_amd import workloads
    cfg = workloads.mono_256()
    p = L.Pupil(amplitude=cfg['amp'], opd=cfg['opd'], pixelscale=cfg['dx'], focal_length=10)
    wl = np.linspace(600e-9, 700e-9, 5)
    eager = L.PSFPlan(p, wl, pixelscale=5e-6, shape=128, oversample=2, use_graph=False)().clone()
    plan = L.PSFPlan(p, wl, pixelscale=5e-6, shape=128, oversample=2, use_graph=True)
    out = torch.empty_like(eager)
    for k in range(4):
        out.fill_(float(k))
        plan(out)
        torch.cuda.synchronize()
        _ = L.fourier.dft2(np.ones((8, 8)), 0.1)   # eager work between replays
        assert torch.equal(out, eager), f'replay {k}'
    acc = eager.clone()
    plan(acc, accumulate=True)
    assert torch.allclose(acc, 2 * eager, rtol=1e-6, atol=0)


def test_plan_profile_modes(L):
    """profile() times the launch groups in the mode calls run in: a graph
    plan by differences of replayed prefix graphs (each group right after
    its producer), an eager plan with events between launches; the image it
    leaves is the plan's image."""
    from lentil_amd import workloads
    cfg = workloads.mono_256()
    p = L.Pupil(amplitude=cfg['amp'], opd=cfg['opd'], pixelscale=cfg['dx'], focal_length=10)
    wl = np.linspace(600e-9, 700e-9, 6)
    for use_graph, mode in ((True, 'graph replay'), (False, 'eager')):
        plan = L.PSFPlan(p, wl, pixelscale=5e-6, shape=128, oversample=2, use_graph=use_graph, lambda_group=4)
        ref = plan().clone()
        out = torch.full_like(ref, -1.0)
        ms = plan.profile(out, repeats=8)
        torch.cuda.synchronize()
        assert ms['mode'] == mode
        t = [ms[k] for k in ('phasor', 'stage1', 'stage2', 'reduce')]
        assert all(np.isfinite(t)) and sum(t) > 0
        if not use_graph:
            assert all(x > 0 for x in t)   # differences of replay times may dip below zero for tiny groups
        assert torch.equal(out, ref)
        plan.close()


def _c2_reference():
    from lentil_amd import workloads
    cfg = workloads.broadband_512()
    p = optics.Pupil(cfg['amp'], cfg['opd'], cfg['dx'], cfg['focal_length'])
    return cfg, optics.broadband_image(p, cfg['wavelengths'], cfg['weights'], cfg['pixelscale'], cfg['shape'],
                                       cfg['oversample'])


@pytest.fixture(scope='module')
def c2_ref():
    return _c2_reference()


@pytest.mark.parametrize('precision', ['fp32', 'bf16x3'])
def test_broadband_512_full(L, c2_ref, golden_c2, precision):
    cfg, ref = c2_ref
    # the oracle itself agrees with the reference's summary
    assert abs(ref.sum() - golden_c2['c2_sum']) < 1e-9
    assert rel_l2(ref[448:576, 448:576], golden_c2['c2_center128']) < 1e-10
    p = L.Pupil(amplitude=cfg['amp'], opd=cfg['opd'], pixelscale=cfg['dx'], focal_length=cfg['focal_length'])
    plan = L.PSFPlan(p, cfg['wavelengths'], cfg['weights'], pixelscale=cfg['pixelscale'], shape=cfg['shape'],
                     oversample=cfg['oversample'], precision=precision)
    img = plan().double().cpu().numpy()
    err = rel_l2(img, ref)
    print(f'broadband_512 {precision}: rel-L2 {err:.3e}; info {plan.info}', flush=True)
    assert err < TOL_I
    assert tuple(np.unravel_index(np.argmax(img), img.shape)) == tuple(golden_c2['c2_argmax'])
    assert rel_l2(img.reshape(64, 16, 64, 16).sum(-1).sum(1), golden_c2['c2_rebin16']) < TOL_I


def _small_segmented(L, prop_shape=None, seg_tilts=True):
    rng = np.random.default_rng(21)
    masks = inputs.hex_segments(rings=1, seg_radius=14, seg_gap=2, drop=(0,))
    n = masks.shape[-1]
    amp = np.sum(masks, axis=0)
    opd = inputs.zernike_compose(amp, np.r_[0, rng.normal(0, 50e-9, size=10)])
    tilts = [(rng.normal(0, 2e-6), rng.normal(0, 2e-6)) for _ in range(len(masks))] if seg_tilts else None
    wl = np.linspace(600e-9, 800e-9, 6)
    wt = rng.uniform(0.5, 1.0, size=6)
    ref_p = optics.Pupil(amp, opd, 1 / n, 10, mask=masks,
                         tilt=[optics.Tilt(*t) for t in tilts] if tilts else None)
    ref = optics.broadband_image(ref_p, wl, wt, 5e-6, 40, 2, prop_shape=prop_shape,
                                 tilts=[optics.Tilt(0.3e-6, -0.2e-6)])
    p = L.Pupil(amplitude=amp, opd=opd, mask=masks, pixelscale=1 / n, focal_length=10)
    if tilts:
        p.tilt = [L.Tilt(*t) for t in tilts]
    return p, wl, wt, ref


@pytest.mark.parametrize('precision', ['fp32', 'bf16x3'])
@pytest.mark.parametrize('contraction', ['per_tile', 'merged'])
def test_segmented_plan_vs_oracle(L, precision, contraction):
    # coherent fast path: all tiles share one output grid, contracted per
    # tile (K concatenation) or summed into their union box
    p, wl, wt, ref = _small_segmented(L, seg_tilts=False)
    plan = L.PSFPlan(p, wl, wt, pixelscale=5e-6, shape=40, oversample=2, tilts=[L.Tilt(0.3e-6, -0.2e-6)],
                     precision=precision, lambda_group=4, contraction=contraction)
    assert plan.info['coherent_fast'] == 1
    assert plan.merged == (contraction == 'merged')
    assert len(plan.tile_shapes) == (1 if plan.merged else 6)
    assert rel_l2(plan().double().cpu().numpy(), ref) < TOL_I


def test_merge_refused_when_segments_split(L):
    """Per-segment tilts give each segment its own output grid: a merged
    contraction would not be exact, so 'merged' raises and 'auto' keeps the
    tiles."""
    p, wl, wt, _ = _small_segmented(L, seg_tilts=True)
    with pytest.raises(ValueError):
        L.PSFPlan(p, wl, wt, pixelscale=5e-6, shape=40, oversample=2, contraction='merged')
    plan = L.PSFPlan(p, wl, wt, pixelscale=5e-6, shape=40, oversample=2)
    assert not plan.merged and len(plan.tile_shapes) == 6
    plan.close()


@pytest.mark.parametrize('precision', ['fp32', 'bf16x3'])
def test_segmented_plan_partial_overlap(L, precision):
    # per-segment tilts and a clipped prop_shape: output Fields overlap partially
    p, wl, wt, ref = _small_segmented(L, prop_shape=24, seg_tilts=True)
    plan = L.PSFPlan(p, wl, wt, pixelscale=5e-6, shape=40, prop_shape=24, oversample=2,
                     tilts=[L.Tilt(0.3e-6, -0.2e-6)], precision=precision)
    assert plan.info['coherent_fast'] == 0
    assert rel_l2(plan().double().cpu().numpy(), ref) < TOL_I
    # and the drop-in per-call path agrees too
    img = np.zeros((80, 80))
    for w_, t_ in zip(wl, wt):
        w = L.Wavefront(w_) * p * L.Tilt(0.3e-6, -0.2e-6)
        w = L.propagate_dft(w, 5e-6, shape=40, prop_shape=24, oversample=2, precision=precision)
        img = w.insert(img, weight=t_)
    assert rel_l2(img, ref) < TOL_I


@pytest.mark.parametrize('tile_n', [64, 128])
@pytest.mark.parametrize('lambda_group', [1, 3, 8])
def test_plan_schedules(L, tile_n, lambda_group):
    """Every schedule knob gives the oracle's image (uneven segment tiles, both
    coherent paths)."""
    p, wl, wt, ref = _small_segmented(L, seg_tilts=False)
    plan = L.PSFPlan(p, wl, wt, pixelscale=5e-6, shape=40, oversample=2, tilts=[L.Tilt(0.3e-6, -0.2e-6)],
                     precision='bf16x3', lambda_group=lambda_group, tile_n=tile_n, contraction='per_tile')
    assert plan.info['coherent_fast'] == 1
    assert rel_l2(plan().double().cpu().numpy(), ref) < TOL_I
    p, wl, wt, ref = _small_segmented(L, prop_shape=24, seg_tilts=True)
    plan = L.PSFPlan(p, wl, wt, pixelscale=5e-6, shape=40, prop_shape=24, oversample=2,
                     tilts=[L.Tilt(0.3e-6, -0.2e-6)], precision='fp32', lambda_group=lambda_group, tile_n=tile_n)
    assert plan.info['coherent_fast'] == 0
    assert rel_l2(plan().double().cpu().numpy(), ref) < TOL_I


def test_plan_bad_tile_n(L):
    p, wl, wt, _ = _small_segmented(L, seg_tilts=False)
    with pytest.raises(ValueError):
        L.PSFPlan(p, wl, wt, pixelscale=5e-6, shape=40, oversample=2, tile_n=96)


def test_autotune(L, tmp_path, monkeypatch):
    """The autotuner times every candidate, keeps one of them, and the tuned
    plan still matches the oracle; later plans of the same shapes take the
    choice from the process table, or, in a new process, from the cache
    file."""
    from lentil_amd import batched, workloads
    monkeypatch.setenv('LENTIL_AMD_TUNE_CACHE', str(tmp_path / 'tune.json'))
    monkeypatch.setattr(batched, '_TUNED', {})
    cfg = workloads.mono_256()
    p = L.Pupil(amplitude=cfg['amp'], opd=cfg['opd'], pixelscale=cfg['dx'], focal_length=10)
    wl = np.linspace(600e-9, 700e-9, 8)
    plan = L.PSFPlan(p, wl, pixelscale=5e-6, shape=128, oversample=2, autotune=True)
    assert plan.tuned_from == 'timed'
    assert set(plan.tuning) == set(batched._CANDIDATES)
    assert all(t > 0 for t in plan.tuning.values())
    assert (plan.lambda_group, plan.tile_n) == min(plan.tuning, key=plan.tuning.get)
    ref_p = optics.Pupil(cfg['amp'], cfg['opd'], cfg['dx'], 10)
    ref = optics.broadband_image(ref_p, wl, np.ones_like(wl), 5e-6, 128, 2)
    assert rel_l2(plan().double().cpu().numpy(), ref) < TOL_I
    # a second plan of the same shapes reuses the choice without timing again
    again = L.PSFPlan(p, wl, pixelscale=5e-6, shape=128, oversample=2, autotune=True)
    assert again.tuned_from == 'process' and again.tuning == {}
    assert (again.lambda_group, again.tile_n) == (plan.lambda_group, plan.tile_n)
    # with an empty process table (a new process) the file supplies it
    monkeypatch.setattr(batched, '_TUNED', {})
    third = L.PSFPlan(p, wl, pixelscale=5e-6, shape=128, oversample=2, autotune=True)
    assert third.tuned_from == 'file' and third.tuning == {}
    assert (third.lambda_group, third.tile_n) == (plan.lambda_group, plan.tile_n)
    for q in (plan, again, third):
        q.close()


def test_plan_copies(L):
    """copies=3: three independent pupils (different OPDs) side by side in one
    detector equal three separate plans, image for image."""
    from lentil_amd import workloads
    cfg = workloads.mono_256()
    wl = np.array([600e-9, 650e-9, 700e-9])
    wt = np.array([0.2, 0.5, 0.3])
    pupil = L.Pupil(amplitude=cfg['amp'], opd=cfg['opd'], pixelscale=cfg['dx'], focal_length=cfg['focal_length'])
    kw = dict(pixelscale=cfg['pixelscale'], shape=cfg['shape'], oversample=cfg['oversample'], precision='fp32',
              lambda_group=4)
    multi = L.PSFPlan(pupil, wl, wt, copies=3, **kw)
    rows, cols = multi.image_shape
    assert multi.shape_out == (rows, 3 * cols)
    rng = np.random.default_rng(21)
    opds = [cfg['opd'] + 1e-8 * k * rng.standard_normal(cfg['opd'].shape) for k in range(3)]
    for k, o in enumerate(opds):
        multi.set_opd(0, o, copy=k)
    wide = multi()
    imgs = multi.images(wide)
    assert imgs.shape == (3, rows, cols)
    for k, o in enumerate(opds):
        single = L.PSFPlan(L.Pupil(amplitude=cfg['amp'], opd=o, pixelscale=cfg['dx'],
                                   focal_length=cfg['focal_length']), wl, wt, **kw)
        ref = single().double().cpu().numpy()
        assert rel_l2(imgs[k].double().cpu().numpy(), ref) < 1e-6
        single.close()
    multi.close()


@pytest.mark.parametrize('precision', ['fp32', 'bf16x3'])
def test_plan_fields_vs_oracle(L, precision):
    """PSFPlan.fields: each problem's complex field (times a weight map and a
    scale) from the operands of the plan's last call equals the complex128
    dft2 of its phasor with the problem's geometry; calling it before
    set_fields raises."""
    from oracle import mft
    n = 64
    amp = inputs.normalize_power(inputs.circle((n, n), n // 2 - 3))
    opd = inputs.zernike_compose(amp, [0, 0, 20e-9, -15e-9, 30e-9, 10e-9])
    pupil = L.Pupil(amplitude=amp, opd=opd, pixelscale=1 / n, focal_length=10.0)
    wl = np.array([550e-9, 650e-9, 750e-9])
    plan = L.PSFPlan(pupil, wl, np.full(3, 1 / 3), pixelscale=5e-6, shape=16, oversample=2, precision=precision,
                     lambda_group=2)
    with pytest.raises(ValueError):
        plan.fields()
    rng = np.random.default_rng(33)
    dst = [torch.empty((p['M'], p['N']), dtype=torch.complex64, device='cuda') for p in plan.problems]
    wmaps = [torch.tensor(rng.uniform(0.5, 1.5, (p['M'], p['N'])), dtype=torch.float32, device='cuda')
             for p in plan.problems]
    scale = [0.5 + k for k in range(plan.n_problems)]
    plan.set_fields(dst, scale, weights=wmaps)
    plan()
    plan.fields()
    torch.cuda.synchronize()
    for i, p in enumerate(plan.problems):
        a_t, o_t = (t.cpu().numpy().astype(np.float64) for t in plan._amp_opd[p['tile']])
        f = a_t * np.exp(2j * np.pi * o_t / plan.wavelengths[p['wave']])
        ref = mft.dft2(f, p['alpha'], (p['M'], p['N']), p['shift'], p['offset'], unitary=False)
        ref = ref * scale[i] * wmaps[i].double().cpu().numpy()
        assert rel_l2(dst[i].cpu().numpy(), ref) < 1e-5, i
    plan.close()


def test_plan_fields_coherent_groups(L):
    """For a segmented pupil whose segments are summed coherently, a group's
    field (the sum of its members' dft2 on the shared grid) goes to the
    output of its first problem."""
    from oracle import mft
    p, wl, wt, _ = _small_segmented(L, seg_tilts=False)
    plan = L.PSFPlan(p, wl, wt, pixelscale=5e-6, shape=40, oversample=2, tilts=[L.Tilt(0.3e-6, -0.2e-6)],
                     precision='fp32', lambda_group=4, contraction='per_tile')
    assert plan.info['coherent_fast'] == 1
    dst = [torch.zeros((q['M'], q['N']), dtype=torch.complex64, device='cuda') for q in plan.problems]
    scale = [1.0 + 0.1 * k for k in range(plan.n_problems)]
    plan.set_fields(dst, scale)
    plan()
    plan.fields()
    torch.cuda.synchronize()
    groups = {}
    for i, q in enumerate(plan.problems):
        groups.setdefault(q['group'], []).append(i)
    assert len(groups) == len(wl)
    for members in groups.values():
        first = members[0]
        q0 = plan.problems[first]
        ref = np.zeros((q0['M'], q0['N']), dtype=complex)
        for i in members:
            q = plan.problems[i]
            assert (q['M'], q['N'], q['out_r0'], q['out_c0']) == (q0['M'], q0['N'], q0['out_r0'], q0['out_c0'])
            a_t, o_t = (t.cpu().numpy().astype(np.float64) for t in plan._amp_opd[q['tile']])
            f = a_t * np.exp(2j * np.pi * o_t / plan.wavelengths[q['wave']])
            ref += mft.dft2(f, q['alpha'], (q['M'], q['N']), q['shift'], q['offset'], unitary=False)
        assert rel_l2(dst[first].cpu().numpy(), ref * scale[first]) < 1e-5
    plan.close()
