"""Batched broadband / segmented PSFs on one MI355X.

``PSFPlan`` runs the reference user loop (docs/user/diffraction.rst:148-160)

    img = zeros(shape * oversample)
    for wl, wt in zip(wavelengths, weights):
        w = Wavefront(wl) * pupil            # one phasor Field per segment
        for t in tilts: w = w * t
        w = propagate_dft(w, pixelscale, shape, prop_shape, oversample, mask)
        img = w.insert(img, weight=wt)       # coherent merge, then |E|^2

as a handful of grouped launches (phasors for every (segment, wavelength),
one stage-1 GEMM launch, one stage-2 GEMM launch with |E|^2 accumulated in
registers across wavelengths, one reduction onto the detector), optionally
replayed from a hipGraph.  The per-Field geometry is the same host code as
``propagate_dft`` (``propagate.field_geometry``), so both paths agree on
every shift, clip and offset.
"""
import ctypes
import hashlib
import json
import os
from itertools import combinations

import numpy as np
import torch

from . import _lib, extent as _ext, helper
from .field import extents_boundary, tilt_shift
from .plane import Pupil, _mul_pixelscale
from .propagate import field_geometry, output_extent


class _TileField:
    """Geometry of a phasor Field: offset and tilt list (what propagate_dft reads)."""

    def __init__(self, shape, offset, tilt):
        self.shape = shape
        self.offset = offset
        self.tilt = tilt

    def shift(self, z, wavelength, pixelscale, oversample, indexing='ij'):
        return tilt_shift(self.tilt, z, wavelength, pixelscale, oversample, indexing)


def coherent_groups(extents):
    """Indices of transitively overlapping extents, grouped in the reference's
    join-first-pair-and-rescan order (field.py:451-461, field.boundary quirks
    included)."""
    groups = [([i], e) for i, e in enumerate(extents)]
    joined = True
    while joined:
        joined = False
        for a, b in combinations(range(len(groups)), 2):
            if _ext.intersect(groups[a][1], groups[b][1]):
                idx = groups[a][0] + groups[b][0]
                groups[a] = (idx, extents_boundary([extents[i] for i in idx]))
                groups.pop(b)
                joined = True
                break
    return [g[0] for g in groups]


def _round_up(v, q):
    return -(-int(v) // q) * q


def contraction_cost(shapes, M, N):
    """Complex MACs the two GEMM stages execute for tiles of ``shapes``
    (m, n) summed coherently onto one M x N grid, their stage-2 K
    concatenated, with the kernels' padding (tiled rows to 128, K to 32),
    for the cheaper of the two contraction orders."""
    row_first = (sum(_round_up(m, 128) * _round_up(n, 32) * N for m, n in shapes)
                 + _round_up(N, 128) * sum(_round_up(m, 32) for m, _ in shapes) * M)
    col_first = (sum(_round_up(n, 128) * _round_up(m, 32) * M for m, n in shapes)
                 + _round_up(M, 128) * sum(_round_up(n, 32) for _, n in shapes) * N)
    return min(row_first, col_first)


def _problems(tile_fields, wavelengths, weights, dx, du, z, prop_out, out_ext, oversample, shape_out):
    problems = []
    gid = 0
    for w, (wl, wt) in enumerate(zip(wavelengths, weights)):
        geo = []
        for t, tf in enumerate(tile_fields):
            g = field_geometry(tf, wl, dx, du, z, prop_out, out_ext, oversample)
            if g is not None:
                geo.append((t, g))
        for members in coherent_groups([_ext.array_extent(g['shape'], g['out_offset']) for _, g in geo]):
            for i in members:
                t, g = geo[i]
                M, N = (int(v) for v in g['shape'])
                problems.append(dict(
                    tile=t, wave=w, group=gid, M=M, N=N, alpha=tuple(float(a) for a in g['alpha']),
                    shift=tuple(float(v) for v in g['shift']), offset=tuple(float(v) for v in g['offset']),
                    out_offset=tuple(int(v) for v in g['out_offset']),
                    out_r0=shape_out[0] // 2 - M // 2 + int(g['out_offset'][0]),
                    out_c0=shape_out[1] // 2 - N // 2 + int(g['out_offset'][1]), weight=float(wt)))
            gid += 1
    return problems


def _grid_key(p):
    return p['M'], p['N'], p['alpha'], p['shift'], p['out_r0'], p['out_c0']


def _mergeable(problems, ntiles, nwave):
    """True when, at every wavelength, every tile lands in one coherent group
    with one output grid (the same M, N, alpha, shift and placement): then
    the group's coherent sum is one MFT of the tiles summed into their union
    box (linearity; docs/user/segmented.rst:19-22)."""
    by_wave = {}
    for p in problems:
        by_wave.setdefault(p['wave'], []).append(p)
    if ntiles < 2 or len(by_wave) != nwave:
        return False
    for ps in by_wave.values():
        if len(ps) != ntiles or len({p['group'] for p in ps}) != 1 or len({_grid_key(p) for p in ps}) != 1:
            return False
    return True


def _union_tile(pupil, slices, tiles, tilt):
    """The tiles summed into the bounding box of their slices of the pupil
    grid: amplitude sum_t amp_t (each already times its segment mask, so
    overlapping segments add as their coherent sum does), the pupil's OPD
    over the box, and the box's offset from the pupil centre."""
    r0 = min(s[0].start for s in slices)
    r1 = max(s[0].stop for s in slices)
    c0 = min(s[1].start for s in slices)
    c1 = max(s[1].stop for s in slices)
    amp = np.zeros((r1 - r0, c1 - c0))
    for (a, _), s in zip(tiles, slices):
        amp[s[0].start - r0:s[0].stop - r0, s[1].start - c0:s[1].stop - c0] += a
    opd_all = np.asarray(pupil.opd)
    if opd_all.size == 1:
        opd = np.full(amp.shape, float(opd_all))
    else:
        opd = np.broadcast_to(opd_all, pupil.shape)[r0:r1, c0:c1]
    box = np.s_[r0:r1, c0:c1]
    return (amp, opd), _TileField(amp.shape, helper.slice_offset(box, pupil.shape), tilt)


def plan_geometry(pupil, wavelengths, weights=None, pixelscale=None, shape=None, prop_shape=None, oversample=2,
                  mask=None, tilts=(), contraction='auto'):
    """Host bookkeeping of the reference loop for every (segment, wavelength).

    Returns dict(shape_out, wavelengths, weights, tiles=[(amp, opd)],
    problems=[dict(tile, wave, group, M, N, alpha, shift, offset, out_offset,
    out_r0, out_c0, weight)], merged, cost_per_tile, cost_merged) where
    ``group`` numbers the coherent groups that Wavefront.insert would merge
    before squaring (field.py:413-461) and (out_r0, out_c0) is where F[0][0]
    lands on the detector (field.py:267, before clipping).

    ``contraction`` chooses how the segments of a coherent group are
    transformed: 'per_tile' runs one MFT per segment tile (their stage-2
    GEMMs concatenated along K); 'merged' sums the tiles into their union
    box first and runs one MFT per wavelength, which is exact when every
    wavelength's tiles share one output grid (ValueError otherwise); 'auto'
    merges when that is exact and executes fewer MACs.
    """
    if contraction not in ('auto', 'per_tile', 'merged'):
        raise ValueError("contraction must be 'auto', 'per_tile' or 'merged'")
    if not isinstance(pupil, Pupil):
        raise TypeError('PSFPlan needs a Pupil')
    if pupil.focal_length is None:
        raise ValueError('pupil focal_length must be set')
    if pixelscale is None:
        raise ValueError('pixelscale must be given')
    wavelengths = np.atleast_1d(np.asarray(wavelengths, dtype=float))
    if weights is None:
        weights = np.ones_like(wavelengths)
    weights = np.broadcast_to(np.asarray(weights, dtype=float), wavelengths.shape).copy()
    dx = _mul_pixelscale(pupil.pixelscale, None)
    if dx is None:
        raise ValueError('pupil pixelscale must be set')
    du = np.broadcast_to(pixelscale, (2,))
    z = pupil.focal_length
    shape = np.asarray(pupil.shape) if shape is None else np.broadcast_to(shape, (2,))
    prop_shape = np.asarray(shape) if prop_shape is None else np.broadcast_to(prop_shape, (2,))
    shape_out, out_ext = output_extent(shape, oversample, mask)
    shape_out = tuple(int(v) for v in shape_out)
    prop_out = prop_shape * oversample

    tiles, tile_fields, slices = [], [], []
    for n, s, amp, opd in pupil.phasor_tiles():
        amp = np.asarray(amp)
        opd = np.asarray(opd)
        if amp.size == 1 and opd.size == 1:
            raise ValueError('PSFPlan needs an array amplitude or OPD')
        if np.iscomplexobj(amp):
            raise TypeError('PSFPlan supports real amplitudes')
        shp = opd.shape if amp.size == 1 else amp.shape
        tiles.append((np.broadcast_to(amp, shp), np.broadcast_to(opd, shp)))
        tilt = ([pupil.tilt[n]] if pupil.tilt else []) + list(tilts)
        tile_fields.append(_TileField(shp, helper.slice_offset(s, pupil.shape), tilt))
        slices.append(s)

    args = (wavelengths, weights, dx, du, z, prop_out, out_ext, oversample, shape_out)
    problems = _problems(tile_fields, *args)
    merged, cost_tiles, cost_merged = False, None, None
    sliced = all(isinstance(s, tuple) and len(s) == 2 and all(isinstance(x, slice) for x in s) for s in slices)
    if contraction != 'per_tile':
        exact = sliced and _mergeable(problems, len(tiles), len(wavelengths))
        if exact:
            utile, ufield = _union_tile(pupil, slices, tiles, tile_fields[0].tilt)
            uproblems = _problems([ufield], *args)
            # the union box must land on the very grid the tiles did
            exact = len(uproblems) == len(wavelengths) and all(
                _grid_key(u) == _grid_key(next(p for p in problems if p['wave'] == u['wave'])) for u in uproblems)
        if exact:
            cost_tiles = sum(contraction_cost([tiles[p['tile']][0].shape for p in problems if p['wave'] == w],
                                              *_grid_key(u)[:2]) for w, u in enumerate(uproblems))
            cost_merged = sum(contraction_cost([utile[0].shape], *_grid_key(u)[:2]) for u in uproblems)
            if contraction == 'merged' or cost_merged < cost_tiles:
                tiles, problems, merged = [utile], uproblems, True
        elif contraction == 'merged':
            raise ValueError("contraction='merged' needs every wavelength's segment tiles to share one output "
                             'grid (no per-segment tilts that split them)')
    return dict(shape_out=shape_out, wavelengths=wavelengths, weights=weights, tiles=tiles, problems=problems,
                merged=merged, cost_per_tile=cost_tiles, cost_merged=cost_merged)


# (lambda_group, tile_n) chosen by PSFPlan(autotune=True), keyed by the
# shapes that decide the schedule: kept for the process and in a JSON file
# (LENTIL_AMD_TUNE_CACHE, default ~/.cache/lentil_amd/autotune.json; an empty
# value turns the file off) whose entries also name the device and the built
# library's hash, so another GPU model or a rebuilt library tunes afresh.
_TUNED = {}
_CANDIDATES = [(lg, tn) for tn in (64, 128) for lg in (4, 8, 16)]
_LIB_HASH = None


def tune_cache_path():
    """Path of the persistent autotune cache, or None when it is turned off."""
    path = os.environ.get('LENTIL_AMD_TUNE_CACHE')
    if path is None:
        path = os.path.join(os.path.expanduser('~'), '.cache', 'lentil_amd', 'autotune.json')
    return path or None


def _library_hash():
    global _LIB_HASH
    if _LIB_HASH is None:
        with open(_lib.LIB_PATH, 'rb') as fh:
            _LIB_HASH = hashlib.sha1(fh.read()).hexdigest()[:16]
    return _LIB_HASH


def tune_cache_key(key, device_name, lib_hash):
    """The cache file's key for one schedule choice: a stable text form of the
    shape key with the device name and the library hash."""
    return json.dumps([device_name, lib_hash, key], separators=(',', ':'), default=int)


def tune_cache_load(path):
    """Entries of the cache file ({key: [lambda_group, tile_n]}); an absent or
    unreadable file is an empty cache."""
    try:
        with open(path) as fh:
            data = json.load(fh)
    except (OSError, ValueError):
        return {}
    if not isinstance(data, dict):
        return {}
    return {k: (int(v[0]), int(v[1])) for k, v in data.items()
            if isinstance(v, list) and len(v) == 2 and tuple(v) in _CANDIDATES}


def tune_cache_store(path, ckey, choice):
    """Add one entry, rewriting the file atomically (another process's
    entries written in between may be lost; they are re-tuned on demand)."""
    data = tune_cache_load(path)
    data[ckey] = [int(choice[0]), int(choice[1])]
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    tmp = f'{path}.{os.getpid()}.tmp'
    with open(tmp, 'w') as fh:
        json.dump(data, fh, indent=0, sort_keys=True)
    os.replace(tmp, path)


class PSFPlan:
    """Weighted broadband (and segmented) PSF accumulated on the device.

    Parameters mirror the reference loop: ``pupil`` is a lentil_amd
    ``Pupil`` (2-D mask: one tile; 3-D mask: one tile per segment), ``tilts``
    are Tilt/DispersiveTilt planes applied after the pupil.  ``precision`` is
    'fp32' or 'bf16x3'; ``lambda_group`` wavelengths share one output tile
    accumulation in registers; ``tile_n`` is the GEMM workgroup width (64 or
    128 columns, 0 = library default); ``autotune`` times every
    (lambda_group, tile_n) candidate on this device and keeps the fastest
    (remembered per problem shape for the process); ``use_graph`` replays the
    launches from a hipGraph after the first call.  ``copies`` > 1 makes that
    many independent copies of the pupil's tiles (each with its own device
    amplitude and OPD, see ``set_opd``) whose images lie side by side in one
    (rows, copies * cols) detector, all computed by the same launches.
    ``contraction`` ('auto', 'per_tile' or 'merged', see ``plan_geometry``)
    decides whether a segmented pupil's tiles are transformed one by one or
    summed into their union box first; ``merged`` tells which was chosen,
    and ``tile_shapes`` lists the tiles the plan transforms (the ones
    ``set_opd`` and ``set_amplitude`` address).
    """

    def __init__(self, pupil, wavelengths, weights=None, pixelscale=None, shape=None, prop_shape=None,
                 oversample=2, mask=None, tilts=(), precision='bf16x3', lambda_group=8, tile_n=0,
                 use_graph=True, autotune=False, copies=1, contraction='auto'):
        geo = plan_geometry(pupil, wavelengths, weights, pixelscale, shape, prop_shape, oversample, mask, tilts,
                            contraction)
        self.merged = geo['merged']
        self.contraction_cost = dict(per_tile=geo['cost_per_tile'], merged=geo['cost_merged'])
        copies = int(copies)
        if copies < 1:
            raise ValueError('copies must be positive')
        self.copies = copies
        self.image_shape = geo['shape_out']
        ntile0, ngroup0 = len(geo['tiles']), 1 + max((p['group'] for p in geo['problems']), default=-1)
        if copies > 1:
            width = geo['shape_out'][1]
            geo['tiles'] = geo['tiles'] * copies
            geo['problems'] = [dict(p, tile=p['tile'] + k * ntile0, group=p['group'] + k * ngroup0,
                                    out_c0=p['out_c0'] + k * width)
                               for k in range(copies) for p in geo['problems']]
            geo['shape_out'] = (geo['shape_out'][0], copies * width)
        self._ntile0 = ntile0
        self.wavelengths = geo['wavelengths']
        self.weights = geo['weights']
        self.precision = precision
        self.oversample = oversample
        self.shape_out = geo['shape_out']
        self.tile_shapes = [a.shape for a, _ in geo['tiles']]
        self._plan = None

        # device tiles (kept alive: the plan holds raw pointers)
        dev = _lib.device()
        self._amp_opd = []
        tiles = []
        for amp, opd in geo['tiles']:
            ta = torch.from_numpy(np.array(amp, dtype=np.float32, order='C', copy=True)).to(dev)
            to = torch.from_numpy(np.array(opd, dtype=np.float64, order='C', copy=True)).to(dev)
            self._amp_opd.append((ta, to))
            tiles.append(_lib.MftxTile(ta.data_ptr(), to.data_ptr(), _lib.DT_F32, _lib.DT_F64, ta.stride(0),
                                       to.stride(0), amp.shape[0], amp.shape[1]))
        problems = [_lib.MftxProblem(p['tile'], p['wave'], p['group'], p['M'], p['N'], p['alpha'][0],
                                     p['alpha'][1], p['shift'][0], p['shift'][1], p['offset'][0], p['offset'][1],
                                     p['out_r0'], p['out_c0'], p['weight']) for p in geo['problems']]
        self.n_problems = len(problems)
        self.problems = geo['problems']
        self._tarr = (_lib.MftxTile * len(tiles))(*tiles)
        self._warr = (ctypes.c_double * len(self.wavelengths))(*self.wavelengths)
        self._parr = (_lib.MftxProblem * max(len(problems), 1))(*problems)
        self._ntiles = len(tiles)

        self.tuning = {}
        self.tuned_from = None
        if autotune:
            key = (precision, self.shape_out, tuple(self.tile_shapes), len(self.wavelengths),
                   tuple((p['tile'], p['M'], p['N'], p['group']) for p in self.problems))
            self.tuned_from = 'process'
            if key not in _TUNED:
                path = tune_cache_path()
                ckey = tune_cache_key(key, torch.cuda.get_device_name(dev), _library_hash()) if path else None
                cached = tune_cache_load(path).get(ckey) if path else None
                if cached is not None:
                    _TUNED[key], self.tuned_from = cached, 'file'
                else:
                    _TUNED[key], self.tuned_from = self._autotune(), 'timed'
                    if path:
                        try:
                            tune_cache_store(path, ckey, _TUNED[key])
                        except OSError:
                            pass   # a read-only home: the choice still holds for this process
            lambda_group, tile_n = _TUNED[key]
        self.lambda_group, self.tile_n = int(lambda_group), int(tile_n)
        self._plan = self._build(self.lambda_group, self.tile_n)
        _lib.load().mftx_plan_set_graph(self._plan, 1 if use_graph else 0)
        info = _lib.MftxPlanInfo()
        _lib.load().mftx_plan_get_info(self._plan, ctypes.byref(info))
        self.info = {k: (list(getattr(info, k)) if k in ('gemm_flops', 'mfma_flops') else getattr(info, k))
                     for k, _ in _lib.MftxPlanInfo._fields_}

    def _build(self, lambda_group, tile_n):
        ctx = _lib.context()
        h = ctypes.c_void_p()
        rc = _lib.load().mftx_plan_psf(ctx, self._tarr, self._ntiles, self._warr, len(self.wavelengths), self._parr,
                                       self.n_problems, self.shape_out[0], self.shape_out[1],
                                       _lib.precision_code(self.precision), int(lambda_group), int(tile_n),
                                       ctypes.byref(h))
        _lib.check(rc, ctx, 'mftx_plan_psf')
        return h.value

    def _autotune(self, repeats=8, passes=2):
        """Time each (lambda_group, tile_n) candidate eagerly on the current
        stream and return the fastest; self.tuning keeps the timings (ms).
        The candidates are timed in passes of alternating order (the second
        runs them in reverse), and a candidate's time is the mean over its
        passes: the power-managed clock drifts while the GPU warms, and a
        single pass in a fixed order favours the candidates timed first."""
        lib = _lib.load()
        ctx = _lib.context()
        img = torch.empty(self.shape_out, dtype=torch.float32, device=_lib.device())
        times = {c: 0.0 for c in _CANDIDATES}
        for p in range(passes):
            for lg, tn in (_CANDIDATES if p % 2 == 0 else _CANDIDATES[::-1]):
                h = self._build(lg, tn)
                try:
                    lib.mftx_plan_set_graph(h, 0)
                    for _ in range(2):
                        _lib.check(lib.mftx_plan_execute(h, ctypes.c_void_p(img.data_ptr()), _lib.DT_F32,
                                                         img.stride(0), 0), ctx, 'mftx_plan_execute')
                    start = torch.cuda.Event(enable_timing=True)
                    stop = torch.cuda.Event(enable_timing=True)
                    start.record()
                    for _ in range(repeats):
                        lib.mftx_plan_execute(h, ctypes.c_void_p(img.data_ptr()), _lib.DT_F32, img.stride(0), 0)
                    stop.record()
                    stop.synchronize()
                    times[(lg, tn)] += start.elapsed_time(stop) / repeats / passes
                finally:
                    lib.mftx_plan_destroy(h)
        self.tuning = times
        return min(times, key=times.get)

    def __call__(self, out=None, accumulate=False):
        """Compute the image; returns ``out`` (a float32/float64 device tensor)."""
        if out is None:
            out = torch.empty(self.shape_out, dtype=torch.float32, device=_lib.device())
            accumulate = False
        if (tuple(out.shape) != self.shape_out or out.dim() != 2 or out.stride(1) != 1
                or out.device != _lib.device()):
            raise ValueError(f'out must be a {self.shape_out} row-major array on {_lib.device()}')
        code = _lib.dtype_code(out)
        if code not in (_lib.DT_F32, _lib.DT_F64):
            raise TypeError('out must be float32 or float64')
        ctx = _lib.context()
        rc = _lib.load().mftx_plan_execute(self._plan, ctypes.c_void_p(out.data_ptr()), code, out.stride(0),
                                           1 if accumulate else 0)
        _lib.check(rc, ctx, 'mftx_plan_execute')
        return out

    def profile(self, out=None, repeats=5):
        """Mean milliseconds of (phasors, stage-1 GEMM, stage-2 GEMM, reduction)
        in the mode calls run in: with ``use_graph``, differences of the mean
        replay times of prefix graphs (the first k launch groups, ``repeats``
        back-to-back replays each between two stream events), so each group
        runs right after its producer as in a replayed call; otherwise HIP
        events between eagerly queued groups.  ``mode`` says which was used."""
        if out is None:
            out = torch.empty(self.shape_out, dtype=torch.float32, device=_lib.device())
        if (tuple(out.shape) != self.shape_out or out.dim() != 2 or out.stride(1) != 1
                or out.device != _lib.device()):
            raise ValueError(f'out must be a {self.shape_out} row-major array on {_lib.device()}')
        ms = (ctypes.c_double * 5)()
        ctx = _lib.context()
        rc = _lib.load().mftx_plan_profile(self._plan, ctypes.c_void_p(out.data_ptr()), _lib.dtype_code(out),
                                           out.stride(0), int(repeats), ms)
        _lib.check(rc, ctx, 'mftx_plan_profile')
        return dict(phasor=ms[0], stage1=ms[1], stage2=ms[2], reduce=ms[3],
                    mode='graph replay' if ms[4] else 'eager')

    def set_opd(self, tile, opd, copy=0):
        """Overwrite a tile's OPD (of one copy) in place (the plan re-reads it on every call)."""
        self._amp_opd[copy * self._ntile0 + tile][1].copy_(torch.as_tensor(opd, dtype=torch.float64))

    def opd_buffer(self, tile, copy=0):
        """The float64 device OPD a tile (of one copy) is read from, for in-place updates."""
        return self._amp_opd[copy * self._ntile0 + tile][1]

    def set_amplitude(self, tile, amp, copy=0):
        """Overwrite a tile's (masked) amplitude (of one copy) in place."""
        self._amp_opd[copy * self._ntile0 + tile][0].copy_(torch.as_tensor(amp, dtype=torch.float32))

    def images(self, out):
        """The per-copy images of a (rows, copies * cols) result, as a
        (copies, rows, cols) view."""
        rows, cols = self.image_shape
        return out.view(rows, self.copies, cols).permute(1, 0, 2)

    def set_fields(self, dst, scale, weights=None):
        """Register, once, where ``fields()`` writes the complex field of each
        problem (problem order): ``dst`` M x N complex64 device tensors with
        one row stride, ``scale`` a number per problem, ``weights`` optional
        real maps of the same layout multiplying the fields.  The tensors are
        kept alive by the plan."""
        n = self.n_problems
        if len(dst) != n or len(scale) != n or (weights is not None and len(weights) != n):
            raise ValueError(f'set_fields needs one entry per problem ({n})')
        ld = dst[0].stride(0)
        for t, p in zip(dst, self.problems):
            if (t.dtype != torch.complex64 or t.device != _lib.device() or tuple(t.shape) != (p['M'], p['N'])
                    or t.stride(1) != 1 or t.stride(0) != ld):
                raise ValueError('fields must be M x N complex64 device tensors with one row stride')
        wdt, ldw = _lib.DT_F32, 0
        if weights is not None:
            wdt = _lib.dtype_code(weights[0])
            ldw = weights[0].stride(0)
            for t, p in zip(weights, self.problems):
                if (_lib.dtype_code(t) != wdt or wdt not in (_lib.DT_F32, _lib.DT_F64) or t.device != _lib.device()
                        or tuple(t.shape) != (p['M'], p['N']) or t.stride(1) != 1 or t.stride(0) != ldw):
                    raise ValueError('weight maps must be M x N real device tensors with one row stride')
        self._fields_keep = (list(dst), None if weights is None else list(weights))
        d = (ctypes.c_void_p * n)(*[t.data_ptr() for t in dst])
        w = None if weights is None else (ctypes.c_void_p * n)(*[t.data_ptr() for t in weights])
        sc = (ctypes.c_double * n)(*[float(x) for x in scale])
        ctx = _lib.context()
        _lib.check(_lib.load().mftx_plan_set_fields(self._plan, d, ld, w, wdt, ldw, sc), ctx, 'mftx_plan_set_fields')

    def fields(self):
        """Write each problem's field F[u][v] (times its weight map and scale)
        into the tensors given to ``set_fields``, from the operands of this
        plan's last call (one complex-output GEMM launch)."""
        ctx = _lib.context()
        _lib.check(_lib.load().mftx_plan_execute_fields(self._plan), ctx, 'mftx_plan_execute_fields')

    def close(self):
        if getattr(self, '_plan', None):
            _lib.load().mftx_plan_destroy(self._plan)
            self._plan = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
