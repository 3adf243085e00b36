"""Differentiable broadband PSFs for phase retrieval (no reference counterpart).

``BroadbandPSFModel(amp, diversity, wavelengths, weights, dx, du, z, npix,
oversample)`` maps an OPD map (and optionally the amplitude) to one image per
diversity OPD,

    I_d = sum_l w_l |dft2(amp exp(2 pi i (opd + div_d)/l), alpha_l, npix*os, unitary)|^2

which is the reference's broadband loop (docs/user/diffraction.rst:148-160,
propagate.py:147-242 with a monolithic pupil).  Gradients flow through a
``torch.autograd.Function`` whose backward pass forms the forward fields
again from the forward plan's stage-1 operands (nothing else is stored) and
applies the adjoint MFT:

    D = dL/d conj(f) = dft2_adjoint(w g F),   g = dL/dI
    dL/d opd = sum (2 pi / l) (-2) Im(conj(D) f),  dL/d amp = sum 2 Re(conj(D) e^{i theta})

The forward runs as one batched PSFPlan over every diversity.  The backward
takes the weighted forward fields from that plan's own stage-1 operands
(PSFPlan.fields: one complex-output GEMM), runs the adjoint MFTs as one
persistent grouped plan (mftx_mft_plan, conjugate twiddles) and one
mftx_phase_grad_many per diversity.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .batched import PSFPlan
from .plane import Pupil


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


class BroadbandPSFModel:
    """Multi-diversity broadband PSF model on the device with gradients."""

    def __init__(self, amp, diversity, wavelengths, weights, dx, du, z, npix, oversample=2,
                 precision='bf16x3', lambda_group=8):
        self.dev = _lib.device()
        amp = np.asarray(amp, dtype=float)
        self.shape = amp.shape
        self.wavelengths = np.atleast_1d(np.asarray(wavelengths, dtype=float))
        self.weights = np.broadcast_to(np.asarray(weights, dtype=float), self.wavelengths.shape).copy()
        self.dx, self.du, self.z = float(dx), float(du), float(z)
        self.npix, self.oversample = int(npix), int(oversample)
        self.M = self.npix * self.oversample
        self.precision = precision
        self.amp = torch.from_numpy(amp.copy()).to(self.dev)
        self.div = [torch.as_tensor(np.asarray(d, dtype=float)).to(self.dev) for d in diversity]
        # a full-grid mask keeps the tile the whole array at offset (0, 0)
        pupil = Pupil(amplitude=amp, opd=np.zeros(self.shape), mask=np.ones(self.shape), pixelscale=dx,
                      focal_length=z)
        # one plan for every diversity: the diversity images lie side by side
        # in one detector, so each launch covers all of them
        self.plan = PSFPlan(pupil, self.wavelengths, self.weights, pixelscale=du, shape=npix,
                            oversample=oversample, precision=precision, lambda_group=lambda_group,
                            copies=len(self.div))
        self._wide = torch.empty(self.plan.shape_out, dtype=torch.float32, device=self.dev)
        self._gen = 0  # forward calls so far: the plan's operands belong to the last one
        # the backward reads opd + diversity from the plan's own OPD buffers
        # (tile 0 of every copy is the whole pupil: the mask is the full grid)
        for d in range(len(self.div)):
            ob = self.plan.opd_buffer(0, copy=d)
            assert tuple(ob.shape) == self.shape and ob.stride(1) == 1
        # the amplitude last written into the plan and into the backward's
        # buffer: (source tensor, its version) -- the source is kept, so its
        # memory cannot be reused by another tensor while it is cached
        self._amp_in_plan = None
        self._amp_in_bw = None
        self.alphas = [(self.dx * self.du) / (wl * self.z * self.oversample) for wl in self.wavelengths]

    def __call__(self, opd, amp=None):
        """Images (ndiv, M, M) float32 for an OPD tensor (float64, on the device)."""
        if amp is None:
            amp = self.amp
        return _PSFStack.apply(opd, amp, self)

    @staticmethod
    def _same(cached, t):
        """Whether ``t`` views the cached source tensor's storage unchanged
        (same pointer, layout and dtype, no in-place write since)."""
        if cached is None:
            return False
        src, version = cached
        return (t.data_ptr() == src.data_ptr() and t.dtype == src.dtype and t.shape == src.shape
                and t.stride() == src.stride() and t._version == version)

    # ---------------------------------------------------------------- forward
    def forward_images(self, opd, amp):
        for d, div in enumerate(self.div):
            torch.add(opd, div, out=self.plan.opd_buffer(0, copy=d))
        if not self._same(self._amp_in_plan, amp):
            amp32 = amp.to(torch.float32)  # converted once for every copy
            for d in range(len(self.div)):
                self.plan.set_amplitude(0, amp32, copy=d)
            self._amp_in_plan = (amp, amp._version)
        self.plan(self._wide)
        self._gen += 1
        # a fresh tensor: the next forward call overwrites the plan's detector
        return self.plan.images(self._wide).clone(memory_format=torch.contiguous_format)

    # --------------------------------------------------------------- backward
    def _backward_state(self, m, n, M):
        """Persistent device buffers of the backward pass, the forward plan's
        field outputs and the adjoint MFT plan that point at them (built once:
        a call only refreshes the buffers)."""
        st = getattr(self, '_bw', None)
        if st is not None:
            return st
        nd, nw = len(self.div), len(self.wavelengths)
        dev = self.dev
        st = dict(amp=torch.empty((m, n), dtype=torch.float64, device=dev),
                  gI=torch.empty((nd, M, M), dtype=torch.float32, device=dev),
                  Y=torch.empty((nd * nw, M, M), dtype=torch.complex64, device=dev),
                  D=torch.empty((nd * nw, m, n), dtype=torch.complex64, device=dev))
        adj = (_lib.MftxMft * (nd * nw))()
        for d in range(nd):
            for j, a in enumerate(self.alphas):
                i = d * nw + j
                r = adj[i]
                r.f, r.f_dtype, r.m, r.n, r.ldf = st['Y'][i].data_ptr(), _lib.DT_C64, M, M, M
                r.F, r.M, r.N, r.ldF = st['D'][i].data_ptr(), m, n, n
                r.alpha_r = r.alpha_c = a
                r.scale, r.conj = float(abs(a)), 1     # sqrt(|a * a|), the unitary scale
        st['adj'] = adj
        st['wls'] = (ctypes.c_double * nw)(*[float(w) for w in self.wavelengths])
        # the weighted forward fields Y = unit w_l g_d F come out of the forward
        # plan (problem order: diversity-major copies of one tile per wavelength)
        Ys, Ws, scales = [], [], []
        for p in self.plan.problems:
            d, j = p['tile'] // self.plan._ntile0, p['wave']
            Ys.append(st['Y'][d * nw + j])
            Ws.append(st['gI'][d])
            scales.append(float(abs(self.alphas[j])) * float(self.weights[j]))
        self.plan.set_fields(Ys, scales, weights=Ws)
        # the adjoint MFTs: one persistent plan (descriptors and workspace built once)
        lib = _lib.load()
        ctx = _lib.context()
        prec = _lib.precision_code(self.precision)
        h = ctypes.c_void_p()
        _lib.check(lib.mftx_mft_plan_create(ctx, adj, nd * nw, prec, ctypes.byref(h)), ctx, 'mftx_mft_plan_create')
        st['plans'] = [h.value]
        self._bw = st
        return st

    def close(self):
        """Release the device plans (also done when the model is collected);
        calling the model afterwards raises."""
        st = getattr(self, '_bw', None)
        if st:
            for h in st['plans']:
                _lib.load().mftx_mft_plan_destroy(h)
            self._bw = None
        plan = getattr(self, 'plan', None)
        if plan is not None:
            plan.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def backward_grads(self, opd, amp, grad_images, need_amp, gen=None):
        """Gradients of sum_d <g_d, I_d> with respect to opd (and amp).

        Two GEMM passes cover every (diversity, wavelength) pair: the forward
        plan's fields from the stage-1 operands of the forward call ``gen``
        (restaged from opd and amp when a later forward call replaced them)
        with the weighted cotangent w_l * g_d folded into the store,
        Y = w_l g_d F, then the adjoint MFT D = dft2_adjoint(Y); one
        phase-gradient launch per diversity sums the wavelengths."""
        lib = _lib.load()
        ctx = _lib.context()
        m, n = self.shape
        M = self.M
        nd, nw = len(self.div), len(self.wavelengths)
        st = self._backward_state(m, n, M)
        amp = amp.detach()
        if not self._same(self._amp_in_bw, amp):
            st['amp'].copy_(amp)
            self._amp_in_bw = (amp, amp._version)
        st['gI'].copy_(grad_images.detach())
        g_opd = torch.zeros((m, n), dtype=torch.float64, device=self.dev)
        g_amp = torch.zeros((m, n), dtype=torch.float64, device=self.dev) if need_amp else None
        if gen is None or gen != self._gen:
            self.forward_images(opd.detach(), amp)
        # from here on the plan's operands and OPD buffers (opd + diversity d)
        # belong to this gradient's forward
        self.plan.fields()
        (adj_plan,) = st['plans']
        _lib.check(lib.mftx_mft_plan_execute(adj_plan), ctx, 'mftx_mft_plan_execute (adjoint)')
        for d in range(nd):
            tot = self.plan.opd_buffer(0, copy=d)
            _lib.check(lib.mftx_phase_grad_many(ctx, _ptr(st['amp']), _lib.DT_F64, n, _ptr(tot),
                                                _lib.DT_F64, tot.stride(0), m, n, st['wls'], nw, _ptr(st['D'][d * nw]), n, m * n,
                                                _ptr(g_opd), n, _ptr(g_amp) if need_amp else None, n),
                       ctx, 'mftx_phase_grad_many')
        return g_opd, g_amp


class _PSFStack(torch.autograd.Function):
    @staticmethod
    def forward(ctx, opd, amp, model):
        ctx.model = model
        ctx.save_for_backward(opd, amp)
        images = model.forward_images(opd.detach(), amp.detach())
        ctx.gen = model._gen
        return images

    @staticmethod
    def backward(ctx, grad_images):
        opd, amp = ctx.saved_tensors
        g_opd, g_amp = ctx.model.backward_grads(opd, amp, grad_images, ctx.needs_input_grad[1], gen=ctx.gen)
        return (g_opd.to(opd.dtype) if ctx.needs_input_grad[0] else None,
                g_amp.to(amp.dtype) if g_amp is not None else None, None)
