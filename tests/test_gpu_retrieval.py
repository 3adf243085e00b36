"""GPU gradients of the broadband multi-diversity PSF model (phase retrieval).

Checked against the oracle's complex128 gradient (itself pinned by complex128
finite differences in test_oracle_retrieval.py) and by central finite
differences of the GPU loss.  Tolerances: images rel-L2 <= 1e-5; gradients
rel-L2 <= 2e-4 vs complex128 (FP32 fields, FP64 accumulation); directional
finite differences within 1%.
"""
import numpy as np
import pytest
import torch

from oracle import inputs, retrieval
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu


def _case(n=64, npix=32, os_=2, nwave=3, ndiv=2, seed=50):
    rng = np.random.default_rng(seed)
    amp = inputs.normalize_power(inputs.circle((n, n), n // 2 - 2))
    opd = inputs.zernike_compose(amp, np.r_[0, rng.normal(0, 40e-9, 14)])
    defocus = inputs.zernike_compose(amp, [0, 0, 0, 1.0])
    steps = [-2, -1, 1, 2][:ndiv] if ndiv <= 4 else list(range(ndiv))
    div = [s * 650e-9 / (2 * np.sqrt(3)) * defocus for s in steps]
    wl = np.linspace(550e-9, 750e-9, nwave)
    wt = np.full(nwave, 1 / nwave)
    geom = dict(dx=1 / n, du=5e-6, z=10.0, npix=npix, oversample=os_)
    target = retrieval.psf_stack(amp, opd + rng.normal(0, 10e-9, opd.shape) * (amp > 0), div, wl, wt,
                                 geom['dx'], geom['du'], geom['z'], npix, os_)
    return amp, opd, div, wl, wt, geom, target


def _model(L, amp, div, wl, wt, geom, precision):
    from lentil_amd.autograd import BroadbandPSFModel
    return BroadbandPSFModel(amp, div, wl, wt, geom['dx'], geom['du'], geom['z'], geom['npix'],
                             geom['oversample'], precision=precision)


@pytest.fixture(scope='module')
def L():
    import lentil_amd
    assert torch.cuda.is_available()
    return lentil_amd


@pytest.mark.parametrize('precision', ['fp32', 'bf16x3'])
def test_gradient_vs_complex128(L, precision):
    amp, opd, div, wl, wt, geom, target = _case()
    model = _model(L, amp, div, wl, wt, geom, precision)
    x = torch.tensor(opd, device='cuda', requires_grad=True)
    a = torch.tensor(amp, device='cuda', requires_grad=True)
    I = model(x, a)
    ref_I = retrieval.psf_stack(amp, opd, div, wl, wt, geom['dx'], geom['du'], geom['z'], geom['npix'],
                                geom['oversample'])
    assert rel_l2(I.detach().double().cpu().numpy(), ref_I) < 1e-5
    T = torch.tensor(target, device='cuda', dtype=torch.float64)
    loss = torch.sum((I.double() - T) ** 2)
    loss.backward()
    g_ref, ga_ref = retrieval.psf_stack_grad(amp, opd, div, wl, wt, geom['dx'], geom['du'], geom['z'],
                                             geom['npix'], geom['oversample'], 2 * (ref_I - target))
    e_opd = rel_l2(x.grad.cpu().numpy(), g_ref)
    e_amp = rel_l2(a.grad.cpu().numpy() * (amp > 0), ga_ref * (amp > 0))
    print(f'{precision}: grad rel-L2 opd {e_opd:.2e} amp {e_amp:.2e}')
    assert e_opd < 2e-4 and e_amp < 2e-4


@pytest.mark.parametrize('precision,n,npix', [('fp32', 48, 24), ('bf16x3', 48, 24), ('bf16x3', 128, 64)])
def test_gradient_finite_difference(L, precision, n, npix):
    """Directional derivatives of the GPU loss, central differences of the
    GPU forward, against the GPU gradient (both forms, two pupil sizes)."""
    amp, opd, div, wl, wt, geom, target = _case(n=n, npix=npix, seed=51)
    model = _model(L, amp, div, wl, wt, geom, precision)
    T = torch.tensor(target, device='cuda', dtype=torch.float64)

    def loss_of(o):
        return torch.sum((model(o).double() - T) ** 2)

    x = torch.tensor(opd, device='cuda', requires_grad=True)
    loss_of(x).backward()
    rng = np.random.default_rng(52)
    for _ in range(3):
        # a smooth direction (random low-order Zernike mix, unit rms) and a
        # step of ~2 mrad of phase: small truncation error, loss change far
        # above the FP32 image noise
        v = inputs.zernike_compose(amp, np.r_[0, rng.standard_normal(14)])
        v = torch.tensor(v / np.sqrt(np.mean(v[amp > 0] ** 2)), device='cuda')
        h = 2e-10
        with torch.no_grad():
            fd = (loss_of(x + h * v) - loss_of(x - h * v)).item() / (2 * h)
        an = torch.sum(x.grad * v).item()
        print(f'{precision} {n}^2: directional derivative fd {fd:.6e} analytic {an:.6e}', flush=True)
        assert abs(fd - an) <= 1e-2 * abs(an), (fd, an)


def test_phase_retrieval_512_gradient(L):
    """Full phase-retrieval workload: 512^2 pupil, 4 diversities, 16 wavelengths, npix 256, os 2."""
    from lentil_amd import workloads
    cfg = workloads.phase_retrieval_512()
    amp, opd = cfg['amp'], cfg['opd']
    rng = np.random.default_rng(53)
    wl, wt = cfg['wavelengths'], cfg['weights']
    args = (cfg['diversity'], wl, wt, cfg['dx'], cfg['pixelscale'], cfg['focal_length'], cfg['shape'],
            cfg['oversample'])
    truth = opd + rng.normal(0, 20e-9, opd.shape) * (amp > 0)
    target_t = retrieval.psf_stack(amp, truth, *args)
    photons = 1e5
    target = rng.poisson(target_t / target_t.sum() * photons) / photons * target_t.sum()
    from lentil_amd.autograd import BroadbandPSFModel
    model = BroadbandPSFModel(amp, cfg['diversity'], wl, wt, cfg['dx'], cfg['pixelscale'], cfg['focal_length'],
                              cfg['shape'], cfg['oversample'], precision='bf16x3')
    x = torch.tensor(opd, device='cuda', requires_grad=True)
    I = model(x)
    ref_I = retrieval.psf_stack(amp, opd, *args)
    assert rel_l2(I.detach().double().cpu().numpy(), ref_I) < 1e-5
    loss = torch.sum((I.double() - torch.tensor(target, device='cuda')) ** 2)
    loss.backward()
    g_ref, _ = retrieval.psf_stack_grad(amp, opd, *args, 2 * (ref_I - target))
    err = rel_l2(x.grad.cpu().numpy(), g_ref)
    print(f'phase_retrieval_512 gradient rel-L2 vs complex128: {err:.2e}')
    assert err < 2e-4


def test_gradient_after_another_forward(L):
    """The backward reuses the forward plan's operands only while they belong
    to its own forward call: with a second forward in between (other OPD), the
    first graph's gradient is unchanged, and so is the second's."""
    amp, opd, div, wl, wt, geom, target = _case(seed=51)
    model = _model(L, amp, div, wl, wt, geom, 'bf16x3')
    T = torch.tensor(target, device='cuda', dtype=torch.float64)
    a = torch.tensor(amp, device='cuda')

    def grad(o, interleave=None):
        x = torch.tensor(o, device='cuda', requires_grad=True)
        loss = torch.sum((model(x, a).double() - T) ** 2)
        if interleave is not None:
            y = torch.tensor(interleave, device='cuda', requires_grad=True)
            loss2 = torch.sum((model(y, a).double() - T) ** 2)
            loss.backward()
            loss2.backward()
            return x.grad.cpu().numpy(), y.grad.cpu().numpy()
        loss.backward()
        return x.grad.cpu().numpy()

    opd2 = opd * 0.5
    g1, g2 = grad(opd), grad(opd2)
    h1, h2 = grad(opd, interleave=opd2)
    assert rel_l2(h1, g1) < 1e-6
    assert rel_l2(h2, g2) < 1e-6
    model.close()


def test_amplitude_written_in_place_is_seen(L):
    """The model keeps the last amplitude it uploaded and skips the upload
    while the tensor is unchanged; an in-place write to it is seen by the
    images and by the gradient (compared with a fresh model given the new
    values)."""
    amp, opd, div, wl, wt, geom, target = _case(seed=52)
    model = _model(L, amp, div, wl, wt, geom, 'bf16x3')
    T = torch.tensor(target, device='cuda', dtype=torch.float64)
    a = torch.tensor(amp, device='cuda')

    def run(m, amp_t):
        x = torch.tensor(opd, device='cuda', requires_grad=True)
        img = m(x, amp_t)
        torch.sum((img.double() - T) ** 2).backward()
        return img.detach().double().cpu().numpy(), x.grad.cpu().numpy()

    i1, g1 = run(model, a)
    i1b, g1b = run(model, a)                 # cached amplitude: same results
    assert np.array_equal(i1, i1b) and np.array_equal(g1, g1b)
    with torch.no_grad():
        a.mul_(0.5)                           # in place: must be re-uploaded
    i2, g2 = run(model, a)
    fresh = _model(L, amp * 0.5, div, wl, wt, geom, 'bf16x3')
    i3, g3 = run(fresh, torch.tensor(amp * 0.5, device='cuda'))
    assert rel_l2(i2, i3) < 1e-6 and rel_l2(g2, g3) < 1e-6
    assert rel_l2(i2, i1) > 0.5               # the images did change (a quarter of the power)
    model.close()
    fresh.close()
