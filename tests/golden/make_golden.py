"""Write golden input/output vectors from the reference implementation.

Run once, in a container where the reference checkout exists:

    python tests/golden/make_golden.py /root/reference

Only the behaviours that closed forms and the reference's own tests leave
open are recorded here (idft2's unitary scale, propagate_dft edge-case
geometry, Field reduce grouping, seeded pupil inputs, and full-size output
summaries).  The fixtures are data: inputs and the reference outputs.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def _import_reference(path):
    sys.path.insert(0, path)
    import lentil  # noqa: E402
    return lentil


def idft2_cases(lentil):
    rng = np.random.default_rng(11)
    out = {}
    specs = [((6, 7), 0.13, (5, 4), (0.3, -1.2), True),
             ((9, 9), (0.07, 0.11), (12, 10), (2.0, -3.0), True),
             ((8, 5), 0.2, None, (0, 0), True),
             ((5, 6), 0.09, (7, 7), (0.5, 0.25), False)]
    for i, (shp, alpha, oshape, shift, unitary) in enumerate(specs):
        F = rng.standard_normal(shp) + 1j * rng.standard_normal(shp)
        f = lentil.fourier.idft2(F, alpha, shape=oshape, shift=shift, unitary=unitary)
        out[f'idft2_{i}_in'] = F
        out[f'idft2_{i}_alpha'] = np.broadcast_to(alpha, (2,)).astype(float)
        out[f'idft2_{i}_shape'] = np.asarray(f.shape)
        out[f'idft2_{i}_shift'] = np.asarray(shift, dtype=float)
        out[f'idft2_{i}_unitary'] = np.asarray(unitary)
        out[f'idft2_{i}_out'] = f
    return out


def _pupil(lentil, n, radius, coeffs, dx, f, mask=None, shift=(0, 0)):
    amp = lentil.normalize_power(lentil.circle((n, n), radius, shift=shift))
    opd = lentil.zernike_compose(amp, coeffs) if coeffs is not None else 0
    return lentil.Pupil(amplitude=amp, opd=opd, pixelscale=dx, focal_length=f, mask=mask), amp, opd


def propagate_cases(lentil):
    """Edge-case geometries: odd sizes, masks, clipping, tilts, prop_shape,
    first-order dispersive tilt, segmented 3-D masks."""
    rng = np.random.default_rng(12)
    cases = []
    # (n, radius, shape, prop_shape, oversample, tilt(x,y) or None, mask spec, disp)
    cases.append(dict(n=32, radius=14, shape=24, prop_shape=None, os=2, tilt=None, mask=None))
    cases.append(dict(n=33, radius=15, shape=(21, 19), prop_shape=None, os=3, tilt=None, mask=None))
    cases.append(dict(n=32, radius=14, shape=24, prop_shape=12, os=2, tilt=None, mask=None))
    cases.append(dict(n=32, radius=14, shape=24, prop_shape=None, os=2,
                      tilt=(1.3e-6, -0.7e-6), mask=None))
    cases.append(dict(n=31, radius=13, shape=20, prop_shape=14, os=2,
                      tilt=(-2.9e-6, 2.2e-6), mask=None))
    cases.append(dict(n=32, radius=14, shape=24, prop_shape=None, os=2, tilt=None,
                      mask=('rect', 20, 14, (5, -7))))
    cases.append(dict(n=32, radius=14, shape=24, prop_shape=None, os=2, tilt=(0.9e-6, 0.4e-6),
                      mask=('rect', 17, 23, (-9, 4))))
    cases.append(dict(n=32, radius=14, shape=16, prop_shape=None, os=2, tilt=(40e-6, 0.0),
                      mask=None))  # pushed fully off the grid
    cases.append(dict(n=30, radius=13, shape=22, prop_shape=None, os=2, tilt=None, mask=None,
                      disp=([0.3, 0.0], [2.0e-2, 550e-9])))
    out = {}
    dx, z, du = 1 / 28, 10.0, 5e-6
    for i, c in enumerate(cases):
        coeffs = np.zeros(8)
        coeffs[1:] = rng.normal(0, 60e-9, size=7)
        p, amp, opd = _pupil(lentil, c['n'], c['radius'], coeffs, dx, z)
        wl = 600e-9 + 37e-9 * i
        w = lentil.Wavefront(wl)
        w = w * p
        if c['tilt'] is not None:
            w = w * lentil.Tilt(x=c['tilt'][0], y=c['tilt'][1])
        if c.get('disp') is not None:
            tr, dp = c['disp']
            w = w * lentil.DispersiveTilt(trace=tr, dispersion=dp)
        mask = None
        sh = np.broadcast_to(c['shape'], (2,)) * c['os']
        if c['mask'] is not None:
            _, wd, ht, msh = c['mask']
            mask = lentil.rectangle(tuple(sh), wd, ht, shift=msh, antialias=False)
        w2 = lentil.propagate_dft(w, pixelscale=du, shape=c['shape'], prop_shape=c['prop_shape'],
                                  oversample=c['os'], mask=mask)
        k = f'prop_{i}_'
        out[k + 'amp'] = amp
        out[k + 'opd'] = np.asarray(opd, dtype=float)
        out[k + 'params'] = np.array([dx, z, du, wl, c['os']], dtype=float)
        out[k + 'shape'] = np.broadcast_to(c['shape'], (2,)).astype(int)
        out[k + 'prop_shape'] = (np.broadcast_to(c['prop_shape'], (2,)).astype(int)
                                 if c['prop_shape'] is not None else np.array([-1, -1]))
        out[k + 'tilt'] = np.asarray(c['tilt'] if c['tilt'] is not None else [np.nan, np.nan])
        if c.get('disp') is not None:
            out[k + 'disp_trace'] = np.asarray(c['disp'][0], dtype=float)
            out[k + 'disp_dispersion'] = np.asarray(c['disp'][1], dtype=float)
        if mask is not None:
            out[k + 'mask'] = mask
        out[k + 'nfields'] = np.asarray(len(w2.data))
        for j, fld in enumerate(w2.data):
            out[k + f'field{j}_data'] = fld.data
            out[k + f'field{j}_offset'] = np.asarray(fld.offset)
        out[k + 'intensity'] = w2.intensity
    return out


def segmented_case(lentil):
    """Two overlapping segment slices + a 3-segment hex pupil, with per-segment tilt."""
    rng = np.random.default_rng(13)
    masks = lentil.hex_segments(rings=1, seg_radius=9, seg_gap=1.5, drop=(0, 2, 4, 6))
    n = masks.shape[-1]
    amp = np.sum(masks, axis=0)
    opd = lentil.zernike_compose(amp, np.r_[0, rng.normal(0, 40e-9, size=6)])
    tilts = [lentil.Tilt(x=rng.normal(0, 0.4e-6), y=rng.normal(0, 0.4e-6)) for _ in range(len(masks))]
    p = lentil.Pupil(amplitude=amp, opd=opd, mask=masks, pixelscale=1 / n, focal_length=10)
    p.tilt = tilts
    out = {'seg_masks': masks, 'seg_amp': amp, 'seg_opd': opd,
           'seg_tilts': np.array([[t.y, t.x] for t in tilts])}  # stored as constructor (x, y)
    w = lentil.Wavefront(700e-9) * p
    out['seg_nphasor'] = np.asarray(len(w.data))
    for j, fld in enumerate(w.data):
        out[f'seg_phasor{j}_data'] = fld.data
        out[f'seg_phasor{j}_offset'] = np.asarray(fld.offset)
    w2 = lentil.propagate_dft(w, pixelscale=5e-6, shape=20, oversample=2)
    out['seg_nfields'] = np.asarray(len(w2.data))
    for j, fld in enumerate(w2.data):
        out[f'seg_field{j}_data'] = fld.data
        out[f'seg_field{j}_offset'] = np.asarray(fld.offset)
    out['seg_intensity'] = w2.intensity
    img = np.zeros((40, 40))
    for k, wl in enumerate([650e-9, 720e-9, 810e-9]):
        w = lentil.Wavefront(wl) * p
        w = lentil.propagate_dft(w, pixelscale=5e-6, shape=20, oversample=2)
        img = w.insert(img, weight=0.2 + 0.1 * k)
    out['seg_broadband'] = img
    return out


def reduce_cases(lentil):
    """Grouping of partially overlapping Fields and merged offsets."""
    from lentil.field import Field
    rng = np.random.default_rng(14)
    out = {}
    sets = [
        [((5, 4), (-2, -2)), ((3, 3), (0, -1)), ((4, 4), (10, 10)), ((2, 6), (12, 6))],
        [((3, 3), (0, 0)), ((3, 3), (3, 0)), ((3, 3), (6, 1)), ((3, 3), (-8, -8))],
        [((6, 6), (0, 0)), ((2, 2), (20, 20)), ((7, 3), (-4, 2)), ((3, 7), (4, -5)),
         ((5, 5), (21, 18))],
        [((4, 5), (-3, -6)), ((4, 5), (3, 6))],
    ]
    for i, s in enumerate(sets):
        fields = []
        for j, (shp, off) in enumerate(s):
            d = rng.standard_normal(shp) + 1j * rng.standard_normal(shp)
            fields.append(Field(data=d, pixelscale=1, offset=list(off)))
            out[f'red_{i}_in{j}_data'] = d
            out[f'red_{i}_in{j}_offset'] = np.asarray(off)
        red = lentil.field.reduce(fields)
        out[f'red_{i}_nin'] = np.asarray(len(fields))
        out[f'red_{i}_nout'] = np.asarray(len(red))
        for j, f in enumerate(red):
            out[f'red_{i}_out{j}_data'] = f.data
            out[f'red_{i}_out{j}_offset'] = np.asarray(f.offset)
        img = np.zeros((48, 48))
        for f in red:
            img = lentil.field.insert(f, img, intensity=True, weight=0.5)
        out[f'red_{i}_intensity'] = img
    return out


def input_cases(lentil):
    rng = np.random.default_rng(15)
    out = {}
    out['in_circle'] = lentil.circle((64, 61), 23.3, shift=(1.5, -2.25))
    out['in_hexagon'] = lentil.hexagon((64, 64), 25, shift=(-3, 2), rotate=True)
    out['in_hexagon_norot'] = lentil.hexagon((63, 64), 21, shift=(1, 0.5))
    out['in_hexseg'] = lentil.hex_segments(rings=2, seg_radius=8, seg_gap=1, drop=(0, 5))
    amp = lentil.circle((96, 96), 40) - lentil.circle((96, 96), 0.12 * 40)
    coeffs = np.r_[0, rng.normal(0, 30e-9, size=21)]
    out['in_zern_amp'] = amp
    out['in_zern_coeffs'] = coeffs
    out['in_zern_opd'] = lentil.zernike_compose(amp, coeffs)
    out['in_normpow'] = lentil.normalize_power(amp)
    return out


def c1_psf(lentil):
    """Config 1 (256^2 pupil -> 256^2 PSF, 650 nm) in full."""
    rng = np.random.default_rng(1)
    amp = lentil.normalize_power(lentil.circle((256, 256), 128))
    coeffs = np.zeros(11)
    coeffs[1:] = rng.normal(0, 20e-9, size=10)
    opd = lentil.zernike_compose(amp, coeffs)
    p = lentil.Pupil(amplitude=amp, opd=opd, pixelscale=1 / 256, focal_length=10)
    w = lentil.Wavefront(650e-9) * p
    w = lentil.propagate_dft(w, pixelscale=5e-6, shape=128, oversample=2)
    return {'c1_coeffs': coeffs, 'c1_opd': opd, 'c1_psf': w.intensity}


def c2_summary(lentil):
    """Config 2 (512^2 pupil, 64 wavelengths -> 1024^2) output summary."""
    rng = np.random.default_rng(2)
    n = 512
    amp = lentil.normalize_power(lentil.circle((n, n), n // 2) - lentil.circle((n, n), 0.12 * n / 2))
    coeffs = np.zeros(22)
    coeffs[1:] = rng.normal(0, 30e-9, size=21)
    opd = lentil.zernike_compose(amp, coeffs)
    p = lentil.Pupil(amplitude=amp, opd=opd, pixelscale=1 / n, focal_length=10)
    wls = np.linspace(500e-9, 900e-9, 64)
    img = np.zeros((1024, 1024))
    for wl in wls:
        w = lentil.Wavefront(wl) * p
        w = lentil.propagate_dft(w, pixelscale=5e-6, shape=256, oversample=4)
        img = w.insert(img, weight=1 / 64)
    srng = np.random.default_rng(99)
    idx = srng.integers(0, 1024, size=(4096, 2))
    return {'c2_coeffs': coeffs, 'c2_sum': np.sum(img), 'c2_max': np.max(img),
            'c2_argmax': np.asarray(np.unravel_index(np.argmax(img), img.shape)),
            'c2_l2': np.linalg.norm(img),
            'c2_rebin16': img.reshape(64, 16, 64, 16).sum(-1).sum(1),
            'c2_center128': img[448:576, 448:576],
            'c2_sample_idx': idx, 'c2_sample_val': img[idx[:, 0], idx[:, 1]]}


def main(ref_path):
    lentil = _import_reference(ref_path)
    small = {}
    for fn in (idft2_cases, propagate_cases, segmented_case, reduce_cases, input_cases):
        small.update(fn(lentil))
    np.savez_compressed(os.path.join(HERE, 'golden_small.npz'), **small)
    np.savez_compressed(os.path.join(HERE, 'golden_c1.npz'), **c1_psf(lentil))
    np.savez_compressed(os.path.join(HERE, 'golden_c2.npz'), **c2_summary(lentil))
    print('wrote fixtures to', HERE)


if __name__ == '__main__':
    main(sys.argv[1] if len(sys.argv) > 1 else '/root/reference')
