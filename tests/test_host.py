"""CPU tests of the host side: the C-ABI library loads and exports every
declared symbol, and the drop-in bookkeeping (extents, insert placement,
coherent grouping, per-Field propagate geometry, input generators) matches
the oracle.  No compute call reaches the library here."""
import ctypes
import os
import re

import numpy as np
import pytest

from oracle import optics, inputs
from tests.conftest import ROOT

HEADER = os.path.join(ROOT, 'include', 'mftx.h')
LIB = os.path.join(ROOT, 'lentil_amd', 'libmftx.so')


def _declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r'^\s*(?:int|const char\*)\s+(mftx_\w+)\s*\(', text, re.M)))


def test_library_exports_every_declared_symbol():
    if not os.path.exists(LIB):
        pytest.skip('libmftx.so not built (run __graft_entry__.build())')
    lib = ctypes.CDLL(LIB)
    names = _declared()
    assert len(names) >= 14
    for name in names:
        assert hasattr(lib, name), name


def test_binding_covers_header():
    from lentil_amd import _lib
    assert sorted(_lib.SIGNATURES) == _declared()


def test_struct_layouts_match_header():
    from lentil_amd import _lib
    # mftx_tile: 2 pointers, 2 int32, 4 int64
    assert ctypes.sizeof(_lib.MftxTile) == 8 * 2 + 4 * 2 + 8 * 4
    # mftx_problem: 5 int32 (+4 pad), 6 doubles, 2 int64, 1 double
    assert ctypes.sizeof(_lib.MftxProblem) == 5 * 4 + 4 + 8 * 6 + 8 * 2 + 8
    assert ctypes.sizeof(_lib.MftxPlanInfo) == 8 * 8 + 4 * 6


def test_error_mapping():
    from lentil_amd import _lib
    with pytest.raises(ValueError):
        _lib.precision_code('fp16')


# ------------------------------------------------------------- bookkeeping

def test_extent_functions_match_oracle():
    from lentil_amd import extent as E
    rng = np.random.default_rng(0)
    for _ in range(500):
        a = E.array_extent(tuple(rng.integers(1, 30, 2)), tuple(rng.integers(-20, 20, 2)))
        b = E.array_extent(tuple(rng.integers(1, 30, 2)), tuple(rng.integers(-20, 20, 2)))
        assert a == optics.array_extent(*_inv(a))
        assert E.intersect(a, b) == optics.intersect(a, b)
        assert E.intersection_shape(a, b) == optics.intersection_shape(a, b)
        if E.intersect(a, b):
            assert E.intersection_slices(a, b) == optics.intersection_slices(a, b)
            assert E.intersection_shift(a, b) == optics.intersection_shift(a, b)
        assert E.array_center(a) == optics.array_center(a)


def _inv(ext):
    shape = (ext[1] - ext[0] + 1, ext[3] - ext[2] + 1)
    return shape, (ext[0] + shape[0] // 2, ext[2] + shape[1] // 2)


def test_insert_slices_match_oracle():
    from lentil_amd.field import insert_slices
    rng = np.random.default_rng(1)
    for _ in range(500):
        fs = tuple(int(v) for v in rng.integers(1, 40, 2))
        os_ = tuple(int(v) for v in rng.integers(1, 40, 2))
        off = tuple(int(v) for v in rng.integers(-40, 40, 2))
        oslc, fslc = optics.insert_slices(fs, off, os_)
        or0, oc0, fr0, fc0, h, w = insert_slices(fs, off, os_)
        if oslc is Ellipsis:
            assert (or0, oc0, fr0, fc0, h, w) == (0, 0, 0, 0) + os_
            continue
        ref_h = max(oslc[0].stop - oslc[0].start, 0)
        ref_w = max(oslc[1].stop - oslc[1].start, 0)
        if h > 0 and w > 0:
            assert (or0, oc0, fr0, fc0, h, w) == (oslc[0].start, oslc[1].start, fslc[0].start, fslc[1].start,
                                                   ref_h, ref_w)
        else:
            assert ref_h == 0 or ref_w == 0 or np.zeros(os_)[oslc].size == 0


def test_coherent_groups_match_oracle(golden_small):
    from lentil_amd.batched import coherent_groups
    g = golden_small
    for i in range(4):
        k = f'red_{i}_'
        fields = [optics.Field(g[k + f'in{j}_data'], 1, list(g[k + f'in{j}_offset']))
                  for j in range(int(g[k + 'nin']))]
        ref = optics.group_fields(fields)
        got = coherent_groups([f.extent for f in fields])
        assert [[fields.index(f) for f in grp] for grp in ref] == got
    rng = np.random.default_rng(2)
    for _ in range(200):
        fields = [optics.Field(np.ones(tuple(rng.integers(1, 6, 2))), 1, list(rng.integers(-8, 8, 2)))
                  for _ in range(rng.integers(1, 7))]
        ref = optics.group_fields(fields)
        got = coherent_groups([f.extent for f in fields])
        assert [[fields.index(f) for f in grp] for grp in ref] == got


@pytest.mark.parametrize('i', range(9))
def test_plan_geometry_matches_oracle(golden_small, i):
    import lentil_amd as L
    from lentil_amd.batched import plan_geometry
    g = golden_small
    k = f'prop_{i}_'
    dx, z, du, wl, os_ = g[k + 'params']
    tilts = []
    otilts = []
    t = g[k + 'tilt']
    if not np.any(np.isnan(t)):
        tilts.append(L.Tilt(x=t[0], y=t[1]))
        otilts.append(optics.Tilt(x=t[0], y=t[1]))
    if (k + 'disp_trace') in g:
        tilts.append(L.DispersiveTilt(trace=g[k + 'disp_trace'], dispersion=g[k + 'disp_dispersion']))
        otilts.append(optics.DispersiveTilt(g[k + 'disp_trace'], g[k + 'disp_dispersion']))
    ps = tuple(g[k + 'prop_shape'])
    ps = None if ps[0] < 0 else ps
    mask = g[k + 'mask'] if (k + 'mask') in g else None
    p = L.Pupil(amplitude=g[k + 'amp'], opd=g[k + 'opd'], pixelscale=dx, focal_length=z)
    geo = plan_geometry(p, [wl], [1.0], du, tuple(g[k + 'shape']), ps, int(os_), mask, tilts)
    op = optics.Pupil(g[k + 'amp'], g[k + 'opd'], dx, z)
    w = op.__rmul__(optics.Wavefront(wl))
    for tt in otilts:
        w = optics.TiltPlane(tt).__rmul__(w)
    plans, shape_out = optics.plan_fields(w.data, wl, op.pixelscale, du, z, tuple(g[k + 'shape']), ps,
                                          int(os_), mask)
    assert geo['shape_out'] == shape_out
    assert len(geo['problems']) == len(plans) == int(g[k + 'nfields'])
    for prob, ref, j in zip(geo['problems'], plans, range(len(plans))):
        assert (prob['M'], prob['N']) == tuple(ref['shape']) == g[k + f'field{j}_data'].shape
        assert np.allclose(prob['alpha'], ref['alpha'], rtol=1e-15)
        assert np.allclose(prob['shift'], ref['shift'], rtol=0, atol=1e-12)
        assert tuple(prob['out_offset']) == tuple(ref['out_offset']) == tuple(g[k + f'field{j}_offset'])


def test_input_generators_match_oracle():
    import lentil_amd as L
    assert np.array_equal(L.circle((64, 61), 23.3, shift=(1.5, -2.25)), inputs.circle((64, 61), 23.3, (1.5, -2.25)))
    assert np.array_equal(L.hexagon((64, 64), 25, shift=(-3, 2), rotate=True),
                          inputs.hexagon((64, 64), 25, (-3, 2), rotate=True))
    assert np.array_equal(L.hex_segments(2, 8, 1, drop=(0, 5)), inputs.hex_segments(2, 8, 1, drop=(0, 5)))
    assert np.array_equal(L.rectangle((50, 40), 20, 10, (2, -3), 15), inputs.rectangle((50, 40), 20, 10, (2, -3), 15))
    amp = inputs.circle((96, 96), 40) - inputs.circle((96, 96), 4.8)
    c = np.r_[0, np.linspace(-3e-8, 3e-8, 21)]
    assert np.allclose(L.zernike_compose(amp, c), inputs.zernike_compose(amp, c), rtol=0, atol=1e-22)
    from lentil_amd.zernike import noll_to_nm
    for j in range(1, 40):
        assert noll_to_nm(j) == inputs.noll_to_nm(j)
    assert list(L.boundary(L.hexagon((256, 256), 100, rotate=True))) == [28, 228, 41, 215]


def test_workloads_are_seeded():
    from lentil_amd import workloads
    a, b = workloads.mono_256(), workloads.mono_256()
    assert np.array_equal(a['opd'], b['opd'])
    c2 = workloads.broadband_512(nwave=4)
    assert c2['amp'].shape == (512, 512) and len(c2['wavelengths']) == 4
    c3 = workloads.segmented_18(nwave=2)
    assert c3['mask'].shape == (18, 1024, 1024)


def test_bench_refuses_several_gpus():
    """bench.py runs on one GPU: --gpus 2 exits with status 2 and says why,
    before touching a device."""
    import json
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, 'bench.py'), '--gpus', '2'], capture_output=True,
                       text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert 'one GPU' in msg['error']
