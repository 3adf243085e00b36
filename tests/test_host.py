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


def test_bench_reports_unresolved_stages_as_none():
    """A stage whose prefix-graph difference is below 1 us (the few-microsecond
    reduction can difference to zero or below) is reported as None with a
    note, never as a negative time."""
    import importlib.util
    spec = importlib.util.spec_from_file_location('bench_mod', os.path.join(ROOT, 'bench.py'))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    got = bench.resolved({'phasor': 0.035, 'stage1': 0.19, 'stage2': 0.31, 'reduce': -0.0008, 'mode': 'graph replay'})
    assert got['reduce'] is None and 'note' in got and got['stage2'] == 0.31 and got['mode'] == 'graph replay'
    got = bench.resolved({'phasor': 0.035, 'reduce': 0.0055, 'mode': 'eager'})
    assert got == {'phasor': 0.035, 'reduce': 0.0055, 'mode': 'eager'}


def test_null_handles_are_rejected():
    """Every entry point that takes a context, plan or MFT-plan handle first
    returns MFTX_EHANDLE for a null handle, before touching any other
    argument or the device, so the error path runs without a GPU."""
    if not os.path.exists(LIB):
        pytest.skip('libmftx.so not built (run __graft_entry__.build())')
    from lentil_amd import _lib
    lib = _lib.load()
    skip = {'mftx_version', 'mftx_create', 'mftx_last_error'}
    checked = 0
    for name, (_, argtypes) in _lib.SIGNATURES.items():
        if name in skip:
            continue
        args = []
        for t in argtypes:
            if t in (ctypes.c_double,):
                args.append(0.0)
            elif t in (ctypes.c_int, ctypes.c_int32, ctypes.c_int64):
                args.append(0)
            else:
                args.append(None)
        assert getattr(lib, name)(*args) == _lib.EHANDLE, name
        checked += 1
    assert checked == len(_lib.SIGNATURES) - len(skip)
    assert lib.mftx_last_error(None) == b'null context'
    # a null output pointer is an argument error, checked before the device
    assert lib.mftx_create(0, None) == _lib.EINVAL


class _FreezePlane:
    """tests/test_plane.py:183-194: a plane whose OPD is random on every read."""

    @staticmethod
    def make():
        import lentil_amd as L

        class FreezePlane(L.Plane):
            def __opd__(self):
                return np.random.uniform()
        return FreezePlane()


def test_freeze():
    p = _FreezePlane.make()
    p.freeze()
    a = p.opd
    b = p.opd
    assert a == b
    with pytest.raises(RuntimeError):
        p.freeze()
    with pytest.raises(RuntimeError):
        p.opd = 1.0
    p.thaw()
    assert not p.frozen
    assert p.opd != p.opd  # read through the hook again
    # freeze(inplace=False) returns a frozen copy and leaves the plane alone
    q = p.freeze(inplace=False)
    assert q.frozen and not p.frozen
    assert q.opd == q.opd


def test_plane_overload_methods():
    """tests/test_plane.py:40-59: the __amp__/__opd__/__mask__ hooks are what
    the properties return."""
    import lentil_amd as L

    class Plane(L.Plane):
        def __amp__(self):
            return np.array(1)

        def __opd__(self):
            return np.array(2)

        def __mask__(self):
            return np.array(3)

    p = Plane()
    assert p.amplitude == 1 and p.opd == 2 and p.mask == 3


def test_frozen_slices_follow_the_mask():
    """A frozen plane's cached slices are used only for the mask they were
    made from; another mask (a hook returning a new array) gets its own."""
    import lentil_amd as L
    m1 = np.zeros((2, 16, 16))
    m1[0, 2:6, 3:9] = 1
    m1[1, 9:14, 1:4] = 1
    m2 = np.zeros((2, 16, 16))
    m2[0, 1:3, 1:3] = 1
    m2[1, 10:16, 8:16] = 1

    class MaskPlane(L.Plane):
        current = m1

        def __mask__(self):
            return self.current

    p = MaskPlane(amplitude=np.ones((16, 16)), opd=np.zeros((16, 16)), pixelscale=1)
    p.freeze()
    p._cache = {'mask': m1, 'slices': ['cached'], 'tiles': []}
    assert p._slice == ['cached']
    p.current = m1.copy()       # an equal array: still the cache
    assert p._slice == ['cached']
    p.current = m2              # another mask: its own slices
    assert p._slice == [optics.boundary_slice(m) for m in m2]
    p.thaw()
    assert p._cache is None
    assert p._slice == [optics.boundary_slice(m) for m in m2]


def test_merged_contraction_is_exact():
    """The union-box tile of a segmented pupil, transformed once, equals the
    coherent sum of its segment tiles transformed one by one (complex128
    oracle MFTs on the plan geometry, every wavelength), and lands on the
    same output grid; per-segment tilts make the merge inexact, so it is
    refused."""
    import lentil_amd as L
    from lentil_amd.batched import plan_geometry
    from oracle import mft
    rng = np.random.default_rng(3)
    masks = inputs.hex_segments(rings=1, seg_radius=9, seg_gap=1, drop=(0,))
    n = masks.shape[-1]
    amp = np.sum(masks, axis=0) * rng.uniform(0.8, 1.0, masks.shape[-2:])
    opd = inputs.zernike_compose(np.sum(masks, axis=0), np.r_[0, rng.normal(0, 60e-9, size=8)])
    p = L.Pupil(amplitude=amp, opd=opd, mask=masks, pixelscale=1 / n, focal_length=10)
    wl = np.array([600e-9, 710e-9, 830e-9])
    args = (p, wl, None, 5e-6, 24, None, 2, None, [L.Tilt(0.4e-6, -0.25e-6)])
    gt = plan_geometry(*args, contraction='per_tile')
    gm = plan_geometry(*args, contraction='merged')
    ga = plan_geometry(*args)
    assert not gt['merged'] and len(gt['tiles']) == len(masks)
    assert gm['merged'] and len(gm['tiles']) == 1 and len(gm['problems']) == len(wl)
    assert ga['merged'] == (ga['cost_merged'] < ga['cost_per_tile'])
    (ua, uo), = gm['tiles']
    for w, lam in enumerate(wl):
        pm = gm['problems'][w]
        ref = np.zeros((pm['M'], pm['N']), dtype=complex)
        for q in (q for q in gt['problems'] if q['wave'] == w):
            a, o = gt['tiles'][q['tile']]
            assert (q['M'], q['N'], q['alpha'], q['shift'], q['out_r0'], q['out_c0']) == \
                (pm['M'], pm['N'], pm['alpha'], pm['shift'], pm['out_r0'], pm['out_c0'])
            ref += mft.dft2(a * np.exp(2j * np.pi * o / lam), q['alpha'], (q['M'], q['N']), q['shift'], q['offset'])
        got = mft.dft2(ua * np.exp(2j * np.pi * uo / lam), pm['alpha'], (pm['M'], pm['N']), pm['shift'],
                       pm['offset'])
        assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-12
    # a planted mistake (the union box placed one row off) is caught
    bad = mft.dft2(ua * np.exp(2j * np.pi * uo / wl[0]), pm['alpha'], (pm['M'], pm['N']), pm['shift'],
                   (gm['problems'][0]['offset'][0] + 1, gm['problems'][0]['offset'][1]))
    ref0 = sum(mft.dft2(gt['tiles'][q['tile']][0] * np.exp(2j * np.pi * gt['tiles'][q['tile']][1] / wl[0]),
                        q['alpha'], (q['M'], q['N']), q['shift'], q['offset'])
               for q in gt['problems'] if q['wave'] == 0)
    assert np.linalg.norm(bad - ref0) / np.linalg.norm(ref0) > 1e-3
    # per-segment tilts split the segments' output grids: no merge
    p.tilt = [L.Tilt(1e-6 * k, -0.5e-6 * k) for k in range(len(masks))]
    assert not plan_geometry(*args)['merged']
    with pytest.raises(ValueError):
        plan_geometry(*args, contraction='merged')


def test_tune_cache_file(tmp_path, monkeypatch):
    """The persistent autotune cache: entries round-trip through the JSON
    file, keys carry the device and library identity, an unreadable file or
    an entry that is not a candidate schedule reads as absent, and the
    environment variable moves or turns off the file."""
    from lentil_amd import batched
    path = str(tmp_path / 'sub' / 'tune.json')
    key = ('bf16x3', (1024, 1024), ((512, 512),), np.int64(64), ((0, np.int64(1024), 1024, 0),))
    k1 = batched.tune_cache_key(key, 'gpu A', 'abc')
    assert k1 != batched.tune_cache_key(key, 'gpu B', 'abc') != batched.tune_cache_key(key, 'gpu A', 'abd')
    assert batched.tune_cache_load(path) == {}
    batched.tune_cache_store(path, k1, (16, 128))
    batched.tune_cache_store(path, 'other', (4, 64))
    assert batched.tune_cache_load(path) == {k1: (16, 128), 'other': (4, 64)}
    with open(path, 'w') as fh:
        fh.write('{"x": [3, 7], "y": [8, 128], "z": "no"}')
    assert batched.tune_cache_load(path) == {'y': (8, 128)}
    with open(path, 'w') as fh:
        fh.write('not json')
    assert batched.tune_cache_load(path) == {}
    monkeypatch.setenv('LENTIL_AMD_TUNE_CACHE', path)
    assert batched.tune_cache_path() == path
    monkeypatch.setenv('LENTIL_AMD_TUNE_CACHE', '')
    assert batched.tune_cache_path() is None


def test_retrieval_amplitude_cache_key():
    """The retrieval model skips re-uploading an amplitude only when the
    tensor views the cached storage unchanged: an in-place write, another
    tensor with equal values, or another view of it all count as new."""
    import torch
    from lentil_amd.autograd import BroadbandPSFModel
    same = BroadbandPSFModel._same
    a = torch.ones((4, 5), dtype=torch.float64)
    cached = (a, a._version)
    assert not same(None, a)
    assert same(cached, a) and same(cached, a.detach())
    assert not same(cached, a.clone())
    assert not same(cached, a.t())
    assert not same(cached, a.float())
    a.mul_(2.0)
    assert not same(cached, a) and not same(cached, a.detach())
