"""GPU parity of the Plane front end (hooks, freeze/thaw, cached device
tiles), of the batched plan under DispersiveTilt (wavelength-dependent
sub-pixel shifts), of plans writing into strided detector views, and of the
C-ABI error paths that need a device."""
import ctypes

import numpy as np
import pytest
import torch

from oracle import optics, inputs
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu

TOL_I = 1e-5


@pytest.fixture(scope='module')
def L():
    import lentil_amd
    assert torch.cuda.is_available()
    return lentil_amd


def _segments(seed=31):
    rng = np.random.default_rng(seed)
    masks = inputs.hex_segments(rings=1, seg_radius=14, seg_gap=2, drop=(0,))
    n = masks.shape[-1]
    amp = np.sum(masks, axis=0)
    opd = inputs.zernike_compose(amp, np.r_[0, rng.normal(0, 50e-9, size=10)])
    return masks, amp, opd, n


def _oracle_loop(amp, opd, n, masks, wl, wt, tilts=(), seg_tilts=None, shape=40, prop_shape=None, oversample=2,
                 pixelscale=5e-6):
    p = optics.Pupil(amp, opd, 1 / n, 10, mask=masks, tilt=seg_tilts)
    return optics.broadband_image(p, wl, wt, pixelscale, shape, oversample, prop_shape=prop_shape, tilts=tilts)


def _device_loop(L, plane, wl, wt, tilts=(), shape=40, prop_shape=None, oversample=2, pixelscale=5e-6):
    img = None
    for w_, t_ in zip(wl, wt):
        w = L.Wavefront(w_) * plane
        for t in tilts:
            w = w * t
        w = L.propagate_dft(w, pixelscale, shape=shape, prop_shape=prop_shape, oversample=oversample)
        img = w.insert(np.zeros(w.shape) if img is None else img, weight=t_)
    return img


# ------------------------------------------------------------------ Plane front end

def test_plane_hooks_vs_oracle(L):
    """A Pupil subclass supplying amplitude, OPD and a 3-D segment mask only
    through __amp__/__opd__/__mask__ (plane.py:40-57) gives the oracle's
    broadband image, per call and batched; the hooks are read on every
    multiply of an unfrozen plane."""
    masks, amp, opd, n = _segments()

    class HookPupil(L.Pupil):
        def __init__(self):
            super().__init__(pixelscale=1 / n, focal_length=10)
            self.reads = 0

        def __amp__(self):
            self.reads += 1
            return amp

        def __opd__(self):
            return opd

        def __mask__(self):
            return masks

    p = HookPupil()
    assert p.size == len(masks) and p.shape == amp.shape
    wl = np.linspace(600e-9, 800e-9, 4)
    wt = np.array([0.4, 0.7, 1.0, 0.6])
    ref = _oracle_loop(amp, opd, n, masks, wl, wt)
    print('plane hooks: per-call loop', flush=True)
    before = p.reads
    assert rel_l2(_device_loop(L, p, wl, wt), ref) < TOL_I
    assert p.reads >= before + len(wl)
    print('plane hooks: batched plan', flush=True)
    plan = L.PSFPlan(p, wl, wt, pixelscale=5e-6, shape=40, oversample=2, precision='bf16x3', lambda_group=2)
    assert rel_l2(plan().double().cpu().numpy(), ref) < TOL_I
    plan.close()


def test_frozen_tiles_reused_and_invalidated(L):
    """A frozen segmented plane uploads its per-segment device tiles once and
    reuses them for every wavelength; thaw drops them, and a new OPD set
    after thaw is what the next image sees."""
    masks, amp, opd, n = _segments()
    p = L.Pupil(amplitude=amp, opd=opd, mask=masks, pixelscale=1 / n, focal_length=10)
    wl = np.linspace(620e-9, 780e-9, 3)
    wt = np.ones(3)
    ref = _oracle_loop(amp, opd, n, masks, wl, wt)
    unfrozen = _device_loop(L, p, wl, wt)
    assert p._cache is None
    print('frozen plane: cached tiles', flush=True)
    p.freeze()
    tiles = p._device_tiles()
    assert p._device_tiles() is tiles
    assert len(tiles) == len(masks)
    frozen = _device_loop(L, p, wl, wt)
    assert p._device_tiles() is tiles          # every wavelength reused them
    assert np.array_equal(frozen, unfrozen)     # same device arithmetic
    assert rel_l2(frozen, ref) < TOL_I
    # the batched plan built from the frozen plane agrees too
    plan = L.PSFPlan(p, wl, wt, pixelscale=5e-6, shape=40, oversample=2, precision='fp32', lambda_group=4)
    assert rel_l2(plan().double().cpu().numpy(), ref) < TOL_I
    plan.close()
    print('frozen plane: thaw and a new OPD', flush=True)
    p.thaw()
    assert p._cache is None
    opd2 = opd + inputs.zernike_compose(amp, [0, 0, 0, 40e-9, -30e-9])
    p.opd = opd2
    ref2 = _oracle_loop(amp, opd2, n, masks, wl, wt)
    got2 = _device_loop(L, p, wl, wt)
    assert rel_l2(got2, ref2) < TOL_I
    assert rel_l2(got2, ref) > 1e-3


def test_frozen_plane_follows_a_new_mask(L):
    """A frozen plane whose __mask__ hook returns another mask gets new slices
    and tiles for it, as the reference (which reads the mask on every
    multiply) does, while the frozen amplitude stays what it was."""
    masks, amp, opd, n = _segments()
    keep = masks[[0, 2, 4]]

    class MaskPupil(L.Pupil):
        current = masks

        def __mask__(self):
            return self.current

    p = MaskPupil(amplitude=amp, opd=opd, pixelscale=1 / n, focal_length=10)
    p.freeze()
    wl, wt = [700e-9], [1.0]
    assert rel_l2(_device_loop(L, p, wl, wt), _oracle_loop(amp, opd, n, masks, wl, wt)) < TOL_I
    first = p._device_tiles()
    p.current = keep
    assert p._device_tiles() is not first and len(p._device_tiles()) == len(keep)
    assert rel_l2(_device_loop(L, p, wl, wt), _oracle_loop(amp, opd, n, keep, wl, wt)) < TOL_I

    class VaryingAmp(L.Pupil):
        scale = 1.0

        def __amp__(self):
            return amp * self.scale

    q = VaryingAmp(opd=opd, mask=masks, pixelscale=1 / n, focal_length=10)
    a = _device_loop(L, q, wl, wt)
    q.scale = 2.0
    b = _device_loop(L, q, wl, wt)
    assert rel_l2(b, 4.0 * a) < 1e-6  # unfrozen: the amplitude hook is read on every multiply
    q.freeze()
    c = _device_loop(L, q, wl, wt)
    q.scale = 3.0
    assert np.array_equal(_device_loop(L, q, wl, wt), c)  # frozen (plane.py:73-83): the stored amplitude
    assert rel_l2(c, b) < 1e-6


def test_phasor_branches(L):
    """The three kinds of _device_tiles entries: scalar amplitude and OPD, a
    complex amplitude (its phase restored after the device phasor), and a
    real array; each equals amp*exp(2*pi*i*opd/lambda) (plane.py:507)."""
    wl = 650e-9
    w = L.Wavefront(wl) * L.Plane(amplitude=2.0, opd=1e-7)
    d = w.data[0].data
    d = d.cpu().numpy() if isinstance(d, torch.Tensor) else np.asarray(d)
    assert np.allclose(d, 2.0 * np.exp(2j * np.pi * 1e-7 / wl), rtol=1e-12)
    rng = np.random.default_rng(5)
    camp = rng.uniform(0.5, 1.0, (12, 10)) * np.exp(1j * rng.uniform(-3, 3, (12, 10)))
    opd = rng.uniform(-1e-7, 1e-7, (12, 10))
    w = L.Wavefront(wl) * L.Plane(amplitude=camp, opd=opd, mask=np.ones((12, 10)), pixelscale=1)
    assert rel_l2(w.data[0].data.cpu().numpy(), camp * np.exp(2j * np.pi * opd / wl)) < 1e-6
    ramp = rng.uniform(0.0, 1.0, (12, 10))
    p = L.Plane(amplitude=ramp, opd=opd, mask=np.ones((12, 10)), pixelscale=1)
    p.freeze()
    w = L.Wavefront(wl) * p
    assert rel_l2(w.data[0].data.cpu().numpy(), ramp * np.exp(2j * np.pi * opd / wl)) < 1e-6


# ------------------------------------------------------------------ DispersiveTilt in the batched plan

def _dispersion_case(nwave):
    """docs/examples/dispersion.rst:14-59: circular pupil, DispersiveTilt
    trace y = x, 3e-4 m per unit wavelength about 550 nm, prop_shape 32."""
    amp = inputs.circle((256, 256), 120)
    wl = np.linspace(475e-9, 625e-9, nwave)
    return amp, wl, dict(trace=[1, 0], dispersion=[3e-4, 550e-9])


@pytest.mark.parametrize('precision', ['fp32', 'bf16x3'])
def test_dispersion_plan_vs_loop_and_oracle(L, precision):
    """The reference's broadband dispersion loop (150 wavelengths, each with
    its own sub-pixel shift) through PSFPlan equals the per-call
    propagate_dft loop and the oracle; neighbouring wavelengths that share
    the integer placement but not the sub-pixel shift stay in separate
    register accumulations."""
    amp, wl, disp = _dispersion_case(150)
    wt = np.ones_like(wl)
    kw = dict(pixelscale=5.5e-6, shape=(128, 128), prop_shape=(32, 32), oversample=3)
    ref = optics.broadband_image(optics.Pupil(amp, 0, 1 / 240, 10), wl, wt, kw['pixelscale'], kw['shape'],
                                 kw['oversample'], prop_shape=kw['prop_shape'],
                                 tilts=[optics.DispersiveTilt(disp['trace'], disp['dispersion'])])
    pupil = L.Pupil(amplitude=amp, opd=0, pixelscale=1 / 240, focal_length=10)
    tilt = L.DispersiveTilt(**disp)
    print(f'dispersion plan {precision}', flush=True)
    plan = L.PSFPlan(pupil, wl, wt, tilts=[tilt], precision=precision, lambda_group=8, **kw)
    assert plan.info['coherent_fast'] == 1
    got = plan().double().cpu().numpy()
    err = rel_l2(got, ref)
    print(f'dispersion plan {precision}: rel-L2 {err:.3e}', flush=True)
    assert err < TOL_I
    print(f'dispersion per-call loop {precision}', flush=True)
    img = np.zeros(ref.shape)
    for w_ in wl:
        w = L.Wavefront(w_) * pupil * tilt
        w = L.propagate_dft(w, precision=precision, **kw)
        img = w.insert(img)
    assert rel_l2(img, ref) < TOL_I
    plan.close()


@pytest.mark.parametrize('precision', ['fp32', 'bf16x3'])
def test_dispersion_segmented_partial_overlap(L, precision):
    """Segments with their own tilts behind a DispersiveTilt and a clipped
    prop_shape: each wavelength's Fields overlap only partially, at offsets
    that move with wavelength, so the plan takes the complex fallback."""
    masks, amp, opd, n = _segments(seed=41)
    rng = np.random.default_rng(42)
    seg_t = [(rng.normal(0, 2e-6), rng.normal(0, 2e-6)) for _ in range(len(masks))]
    wl = np.linspace(560e-9, 640e-9, 12)
    wt = rng.uniform(0.5, 1.0, len(wl))
    disp = dict(trace=[0.5, 0], dispersion=[2e-4, 600e-9])
    ref = _oracle_loop(amp, opd, n, masks, wl, wt, tilts=[optics.DispersiveTilt(disp['trace'], disp['dispersion'])],
                       seg_tilts=[optics.Tilt(*t) for t in seg_t], prop_shape=24)
    p = L.Pupil(amplitude=amp, opd=opd, mask=masks, pixelscale=1 / n, focal_length=10)
    p.tilt = [L.Tilt(*t) for t in seg_t]
    print(f'dispersion segmented plan {precision}', flush=True)
    plan = L.PSFPlan(p, wl, wt, pixelscale=5e-6, shape=40, prop_shape=24, oversample=2,
                     tilts=[L.DispersiveTilt(**disp)], precision=precision, lambda_group=4)
    assert plan.info['coherent_fast'] == 0
    assert rel_l2(plan().double().cpu().numpy(), ref) < TOL_I
    plan.close()


# ------------------------------------------------------------------ strided detector views

def test_plan_writes_only_its_view(L):
    """A plan writing into a column view of a wider array touches only the
    view (the zeroing before the reduction and the fallback accumulation
    honour the row stride)."""
    masks, amp, opd, n = _segments()
    cases = [
        ('prop_shape', L.Pupil(amplitude=amp, opd=opd, pixelscale=1 / n, focal_length=10),
         dict(prop_shape=24), 1),
        ('fallback', L.Pupil(amplitude=amp, opd=opd, mask=masks, pixelscale=1 / n, focal_length=10),
         dict(prop_shape=24), 0),
    ]
    cases[1][1].tilt = [L.Tilt(2e-6 * (k - 2), -1e-6 * k) for k in range(len(masks))]
    wl = np.array([600e-9, 700e-9])
    for name, pupil, kw, fast in cases:
        print(f'strided view: {name}', flush=True)
        plan = L.PSFPlan(pupil, wl, [0.5, 0.5], pixelscale=5e-6, shape=40, oversample=2, precision='fp32', **kw)
        assert plan.info['coherent_fast'] == fast
        rows, cols = plan.shape_out
        ref = plan().clone()
        big = torch.full((rows + 1, cols + 7), 7.0, dtype=torch.float32, device='cuda')
        view = big[:rows, 3:3 + cols]
        plan(view)
        torch.cuda.synchronize()
        assert torch.equal(view, ref), name
        assert torch.all(big[:, :3] == 7.0) and torch.all(big[:, 3 + cols:] == 7.0), name
        assert torch.all(big[rows:, :] == 7.0), name
        plan.close()


def test_host_tensors_take_the_host_path(L):
    """CPU torch tensors are host data, like NumPy arrays: an input is copied
    to the device and an output is written back, never handed to a kernel
    as a pointer; outputs a kernel writes directly must be on the device."""
    from oracle import mft
    rng = np.random.default_rng(3)
    f = rng.normal(size=(12, 10)) + 1j * rng.normal(size=(12, 10))
    kw = dict(shape=(16, 14), shift=(0.3, -0.1))
    ref = mft.dft2(f, 0.07, **kw)
    got = L.fourier.dft2(torch.tensor(f), 0.07, **kw)
    assert isinstance(got, np.ndarray)
    err = rel_l2(got, ref)
    assert err < 1e-5, err
    o = torch.zeros((16, 14), dtype=torch.complex128)
    assert L.fourier.dft2(f, 0.07, out=o, **kw) is o
    err = rel_l2(o.numpy(), ref)
    assert err < 1e-5, err

    n = 32
    amp = inputs.circle((n, n), n // 2 - 1)
    opd = inputs.zernike_compose(amp, [0, 0, 0, 60e-9])
    pupil = L.Pupil(amplitude=amp, opd=opd, pixelscale=1 / n, focal_length=10)
    w = L.propagate_dft(L.Wavefront(650e-9) * pupil, 5e-6, shape=16, oversample=2)
    host = w.insert(np.zeros((32, 32)), weight=0.5)
    cpu_t = torch.zeros((32, 32), dtype=torch.float64)
    assert w.insert(cpu_t, weight=0.5) is cpu_t
    assert np.array_equal(cpu_t.numpy(), host)

    plan = L.PSFPlan(pupil, [650e-9], [1.0], pixelscale=5e-6, shape=16, oversample=2)
    with pytest.raises(ValueError):
        plan(torch.zeros(plan.shape_out, dtype=torch.float32))
    with pytest.raises(ValueError):
        plan.profile(torch.zeros(plan.shape_out, dtype=torch.float32))
    with pytest.raises(ValueError):
        plan.set_fields([torch.zeros((p['M'], p['N']), dtype=torch.complex64) for p in plan.problems],
                        [1.0] * plan.n_problems)
    img = plan()
    torch.cuda.synchronize()
    err = rel_l2(img.double().cpu().numpy(), host * 2)
    assert err < 1e-5, err
    plan.close()


def test_model_images_are_fresh_tensors(L):
    """With one diversity image the model still returns a tensor of its own:
    the next forward call does not overwrite an image already returned."""
    from lentil_amd.autograd import BroadbandPSFModel
    n = 64
    amp = inputs.circle((n, n), n // 2 - 2)
    model = BroadbandPSFModel(amp, [np.zeros((n, n))], [600e-9, 700e-9], [1.0, 1.0], 1 / n, 5e-6, 10.0, 16,
                              precision='fp32')
    x1 = torch.zeros((n, n), dtype=torch.float64, device='cuda')
    x2 = torch.tensor(inputs.zernike_compose(amp, [0, 0, 0, 80e-9]), device='cuda')
    a = model(x1)
    a_copy = a.clone()
    b = model(x2)
    torch.cuda.synchronize()
    assert a.shape == (1, 32, 32)
    assert torch.equal(a, a_copy)
    assert not torch.equal(a, b)
    assert a.data_ptr() != b.data_ptr()
    model.close()


# ------------------------------------------------------------------ one tile in several problems

def _raw_plan_image(tile_arrays, wl, specs, image_shape, precision, lambda_group=0):
    """A plan built straight through mftx_plan_psf (one problem per spec) and
    executed once: the image as float64 numpy."""
    from lentil_amd import _lib
    lib, ctx = _lib.load(), _lib.context()
    keep = []
    tiles = []
    for amp, opd in tile_arrays:
        ta = torch.tensor(amp, dtype=torch.float32, device='cuda')
        to = torch.tensor(opd, dtype=torch.float64, device='cuda')
        keep += [ta, to]
        tiles.append(_lib.MftxTile(ta.data_ptr(), to.data_ptr(), _lib.DT_F32, _lib.DT_F64, ta.stride(0), to.stride(0),
                                   amp.shape[0], amp.shape[1]))
    probs = [_lib.MftxProblem(s['tile'], s['wave'], s['group'], s['M'], s['N'], s['alpha'], s['alpha'], s['shift'][0],
                              s['shift'][1], s['offset'][0], s['offset'][1], s['out'][0], s['out'][1], s['weight'])
             for s in specs]
    tarr = (_lib.MftxTile * len(tiles))(*tiles)
    parr = (_lib.MftxProblem * len(probs))(*probs)
    warr = (ctypes.c_double * len(wl))(*wl)
    h = ctypes.c_void_p()
    _lib.check(lib.mftx_plan_psf(ctx, tarr, len(tiles), warr, len(wl), parr, len(probs), image_shape[0],
                                 image_shape[1], _lib.precision_code(precision), lambda_group, 0, ctypes.byref(h)),
               ctx, 'mftx_plan_psf')
    img = torch.full(image_shape, -1.0, dtype=torch.float32, device='cuda')
    try:
        _lib.check(lib.mftx_plan_execute(h, ctypes.c_void_p(img.data_ptr()), _lib.DT_F32, image_shape[1], 0), ctx,
                   'mftx_plan_execute')
        torch.cuda.synchronize()
    finally:
        lib.mftx_plan_destroy(h)
    return img.double().cpu().numpy()


@pytest.mark.parametrize('precision', ['fp32', 'bf16x3'])
def test_plan_one_tile_twice_at_one_wavelength(L, precision):
    """The C ABI lets one tile appear at one wavelength in several problems
    with different output grids (here one pupil placed twice on the detector
    with different sub-pixel shifts and Field offsets, one placement clipped
    by the image edge, plus a third problem at another wavelength).  Each
    placement needs an operand carrying its own grid's folds: the image is
    the weighted sum of |dft2|^2 of every problem (complex128 oracle), and
    equals the sum of one-problem plans."""
    from oracle import mft
    rng = np.random.default_rng(5)
    m = 48
    amp = inputs.circle((m, m), 22)
    opd = inputs.zernike_compose(amp, np.r_[0, rng.normal(0, 60e-9, size=8)])
    wl = [650e-9, 540e-9]
    alpha = 1.0 / (2 * m)
    specs = [dict(tile=0, wave=0, group=0, M=64, N=64, alpha=alpha, shift=(0.3, -0.2), offset=(0.0, 0.0),
                  out=(4, 6), weight=1.0),
             dict(tile=0, wave=0, group=1, M=64, N=64, alpha=alpha, shift=(-0.45, 0.15), offset=(1.5, -2.0),
                  out=(40, -10), weight=0.7),
             dict(tile=0, wave=1, group=2, M=64, N=64, alpha=alpha * 650 / 540, shift=(0.1, 0.4), offset=(0.0, 0.0),
                  out=(20, 30), weight=0.5)]
    shape = (100, 90)
    ref = np.zeros(shape)
    for s in specs:
        f = amp * np.exp(2j * np.pi * opd / wl[s['wave']])
        F = mft.dft2(f, alpha=s['alpha'], shape=(s['M'], s['N']), shift=s['shift'], offset=s['offset'])
        r0, c0 = s['out']
        rs, cs = slice(max(r0, 0), min(r0 + s['M'], shape[0])), slice(max(c0, 0), min(c0 + s['N'], shape[1]))
        ref[rs, cs] += s['weight'] * np.abs(F[rs.start - r0:rs.stop - r0, cs.start - c0:cs.stop - c0]) ** 2
    print(f'one tile twice: {precision}', flush=True)
    img = _raw_plan_image([(amp, opd)], wl, specs, shape, precision)
    err = rel_l2(img, ref)
    assert err < TOL_I, err
    singles = sum(_raw_plan_image([(amp, opd)], wl, [s], shape, precision) for s in specs)
    err = rel_l2(img, singles)
    assert err < 1e-6, err


@pytest.mark.parametrize('precision', ['fp32', 'bf16x3'])
def test_plan_mixes_intensity_and_coherent_groups(L, precision):
    """A plan whose coherent groups differ: at one wavelength two tiles summed
    coherently with different shifts and placements (two output grids: their
    fields are summed before |.|^2), at another a lone problem (the
    intensity path).  The image is the oracle's: |F_a + F_b|^2 on the union
    box plus w |F_c|^2."""
    from oracle import mft
    from lentil_amd import _lib
    rng = np.random.default_rng(9)
    m = 40
    amp_a = inputs.circle((m, m), 18)
    amp_b = inputs.circle((m, m), 12, shift=(3, -2))
    opd_a = inputs.zernike_compose(amp_a, np.r_[0, rng.normal(0, 50e-9, size=6)])
    opd_b = inputs.zernike_compose(amp_b, np.r_[0, rng.normal(0, 50e-9, size=6)])
    wl = [600e-9, 700e-9]
    alpha = 1.0 / (2 * m)
    specs = [dict(tile=0, wave=0, group=0, M=48, N=48, alpha=alpha, shift=(0.3, -0.2), offset=(0.0, 0.0),
                  out=(4, 6), weight=1.0),
             dict(tile=1, wave=0, group=0, M=48, N=48, alpha=alpha, shift=(0.1, 0.25), offset=(1.0, -1.5),
                  out=(10, -3), weight=1.0),
             dict(tile=0, wave=1, group=1, M=48, N=48, alpha=alpha * 6 / 7, shift=(-0.2, 0.4), offset=(0.0, 0.0),
                  out=(30, 40), weight=0.6)]
    shape = (90, 80)
    tiles = [(amp_a, opd_a), (amp_b, opd_b)]

    def field(sp):
        amp, opd = tiles[sp['tile']]
        return mft.dft2(amp * np.exp(2j * np.pi * opd / wl[sp['wave']]), alpha=sp['alpha'], shape=(sp['M'], sp['N']),
                        shift=sp['shift'], offset=sp['offset'])

    def place(canvas, F, out):
        r0, c0 = out
        rs = slice(max(r0, 0), min(r0 + F.shape[0], shape[0]))
        cs = slice(max(c0, 0), min(c0 + F.shape[1], shape[1]))
        canvas[rs, cs] += F[rs.start - r0:rs.stop - r0, cs.start - c0:cs.stop - c0]

    coherent = np.zeros(shape, complex)
    for sp in specs[:2]:
        place(coherent, field(sp), sp['out'])
    ref = np.abs(coherent) ** 2
    lone = np.zeros(shape)
    place(lone, specs[2]['weight'] * np.abs(field(specs[2])) ** 2, specs[2]['out'])
    ref += lone
    print(f'mixed plan: {precision}', flush=True)
    img = _raw_plan_image(tiles, wl, specs, shape, precision)
    err = rel_l2(img, ref)
    assert err < TOL_I, err
    # the lone problem's window carries only its intensity (no fallback double count)
    win = (slice(60, 78), slice(70, 80))
    err = rel_l2(img[win], ref[win])
    assert err < TOL_I, err


# ------------------------------------------------------------------ error paths with a device

def test_status_codes_with_a_context(L):
    """Argument errors raise ValueError (MFTX_EINVAL), a workspace larger than
    the device raises MemoryError (MFTX_ENOMEM) without launching anything,
    and the context works normally afterwards."""
    from lentil_amd import _lib
    lib = _lib.load()
    ctx = _lib.context()
    f = torch.ones((8, 8), dtype=torch.complex64, device='cuda')
    F = torch.empty((8, 8), dtype=torch.complex64, device='cuda')

    def dft2(m, n, M, N, ldf=None, ldF=None, prec=0):
        return lib.mftx_dft2(ctx, ctypes.c_void_p(f.data_ptr()), _lib.DT_C64, m, n, ldf or n,
                             ctypes.c_void_p(F.data_ptr()), M, N, ldF or N, 0.1, 0.1, 0.0, 0.0, 0.0, 0.0, 1.0, 0,
                             prec)
    assert dft2(0, 8, 8, 8) == _lib.EINVAL
    assert dft2(8, 8, 8, 8, ldf=4) == _lib.EINVAL
    assert dft2(8, 8, 8, 8, prec=7) == _lib.EINVAL
    assert b'precision' in lib.mftx_last_error(ctx)
    assert dft2((1 << 20) + 1, 8, 8, 8) == _lib.EINVAL
    assert b'too large' in lib.mftx_last_error(ctx)
    print('status codes: oversized workspace', flush=True)
    big = 1 << 18  # 2^18 x 2^18 packed operand: 512 GiB, more than the device holds
    rc = dft2(big, big, big, big)
    assert rc == _lib.ENOMEM, rc
    with pytest.raises(MemoryError):
        _lib.check(rc, ctx, 'mftx_dft2')
    # the failed allocation leaves no pending error: the next transform runs
    out = L.fourier.dft2(np.ones((8, 8)), 0.125)
    ref = optics_dft2(np.ones((8, 8)), 0.125)
    assert rel_l2(out, ref) < 1e-6
    # a rebin factor that does not divide the image is refused (the reference's reshape fails)
    img = torch.ones((10, 12), dtype=torch.float64, device='cuda')
    small = torch.empty((3, 4), dtype=torch.float64, device='cuda')
    assert lib.mftx_rebin(ctx, ctypes.c_void_p(img.data_ptr()), _lib.DT_F64, 1, 10, 12, 12, 3,
                          ctypes.c_void_p(small.data_ptr()), _lib.DT_F64) == _lib.EINVAL
    assert b'divide' in lib.mftx_last_error(ctx)
    # an insert slice wider than its source's or destination's row stride is refused
    fld = torch.zeros((8, 8), dtype=torch.complex64, device='cuda')
    dst = torch.zeros((8, 8), dtype=torch.float64, device='cuda')
    ins = lambda fc0, oc0: lib.mftx_insert(ctx, ctypes.c_void_p(fld.data_ptr()), 8, 0, fc0, 4, 6,
                                           ctypes.c_void_p(dst.data_ptr()), _lib.DT_F64, 8, 0, oc0, 1, 1.0)
    assert ins(2, 2) == _lib.OK
    assert ins(3, 0) == _lib.EINVAL and ins(0, 3) == _lib.EINVAL
    assert b'row stride' in lib.mftx_last_error(ctx)
    # a plan problem with non-finite geometry is refused before anything is built
    amp = torch.ones((8, 8), dtype=torch.float32, device='cuda')
    opd = torch.zeros((8, 8), dtype=torch.float64, device='cuda')
    tiles = (_lib.MftxTile * 1)(_lib.MftxTile(amp.data_ptr(), opd.data_ptr(), _lib.DT_F32, _lib.DT_F64, 8, 8, 8, 8))
    warr = (ctypes.c_double * 1)(600e-9)
    bad = [('alpha_r', float('nan')), ('shift_c', float('nan')), ('offset_r', float('nan')),
           ('weight', float('inf')), ('out_r0', 1 << 31), ('out_c0', -(1 << 31))]
    for field, value in bad:
        p = _lib.MftxProblem(0, 0, 0, 16, 16, 0.05, 0.05, 0.0, 0.0, 0.0, 0.0, 0, 0, 1.0)
        setattr(p, field, value)
        h = ctypes.c_void_p()
        rc = lib.mftx_plan_psf(ctx, tiles, 1, warr, 1, (_lib.MftxProblem * 1)(p), 1, 16, 16, 0, 0, 0, ctypes.byref(h))
        assert rc == _lib.EINVAL and h.value is None, field
        assert b'bad problem' in lib.mftx_last_error(ctx), field
    p = _lib.MftxProblem(0, 0, 0, 16, 16, 0.05, 0.05, 0.0, 0.0, 0.0, 0.0, 0, 0, 1.0)
    h = ctypes.c_void_p()
    rc = lib.mftx_plan_psf(ctx, tiles, 1, warr, 1, (_lib.MftxProblem * 1)(p), 1, 1 << 21, 16, 0, 0, 0, ctypes.byref(h))
    assert rc == _lib.EINVAL and b'image shape' in lib.mftx_last_error(ctx)


def optics_dft2(f, alpha):
    from oracle import mft
    return mft.dft2(f, alpha)
