"""Input generators for seeded configurations (oracle only).

Restates the reference's shape, segment and Zernike generators so the tests
can build the same pupils the reference would:

* mesh            helper.py:7-17
* circle          shape.py:6-30
* hexagon         shape.py:33-75
* rectangle       shape.py:78-114
* hex_segments    segmented.py:8-80, 86-133 (hex-grid maths)
* centroid        util.py:6-28
* normalize_power util.py:378-411
* noll_to_nm      zernike.py:348-391
* zernike         zernike.py:11-93 (Noll order), R zernike.py:96-109
* zernike_coordinates zernike.py:444-480
* zernike_compose zernike.py:112-180
"""
from math import factorial

import numpy as np


def mesh(shape, shift=(0, 0), angle=0):
    nr, nc = shape[0], shape[1]
    rr, cc = np.meshgrid(np.arange(nr) - np.floor(nr / 2.0) - shift[0],
                         np.arange(nc) - np.floor(nc / 2.0) - shift[1], indexing='ij')
    a = np.deg2rad(angle)
    return rr * np.cos(a) + cc * np.sin(a), rr * -np.sin(a) + cc * np.cos(a)


def circle(shape, radius, shift=(0, 0), antialias=True):
    rr, cc = mesh(shape)
    r = np.sqrt(np.square(rr - shift[0]) + np.square(cc - shift[1]))
    mask = np.clip(radius + 0.5 - r, 0.0, 1.0)
    if not antialias:
        mask[mask > 0] = 1
    return mask


def hexagon(shape, radius, shift=(0, 0), rotate=False, antialias=True):
    shape = np.broadcast_to(shape, (2,))
    r, c = mesh(shape, shift)
    mask = np.ones(shape)
    inner = radius * np.sqrt(3) / 2
    for n in range(6):
        theta = n * np.pi / 3 if rotate else n * np.pi / 3 + np.pi / 6
        rho = r * np.sin(theta) + c * np.cos(theta)
        if antialias:
            slc = np.clip(inner + 0.5 - rho, 0.0, 1.0)
        else:
            slc = np.ones(shape)
            slc[rho > inner] = 0
        mask = np.minimum(mask, slc)
    return mask


def rectangle(shape, width, height, shift=(0, 0), angle=0, antialias=True):
    shape = np.broadcast_to(shape, (2,))
    rr, cc = mesh(shape, shift, angle)
    rect = np.minimum(np.minimum(np.ones(shape), np.clip(0.5 + width / 2 - np.abs(cc), 0, 1)),
                      np.clip(0.5 + height / 2 - np.abs(rr), 0, 1))
    if not antialias:
        rect[rect > 0] = 1
    return rect


def _hex_ring(radius):
    """Cube-coordinate ring walk (segmented.py:105-113)."""
    dirs = [(1, 0, -1), (1, -1, 0), (0, -1, 1), (-1, 0, 1), (-1, 1, 0), (0, 1, -1)]
    h = (-radius, radius, 0)
    out = []
    for i in range(6):
        for _ in range(radius):
            out.append(h)
            h = (h[0] + dirs[i][0], h[1] + dirs[i][1], h[2] + dirs[i][2])
    return out


def _hex_to_rc(h, radius, rotate=False):
    """segmented.py:116-130 ((x, y) -> (r, c) = (-y, x))."""
    q, r = h[0], h[1]
    if rotate:
        x = radius * (np.sqrt(3) * q + np.sqrt(3) / 2 * r)
        y = radius * (3 / 2 * r)
    else:
        x = radius * (3 / 2 * q)
        y = radius * (np.sqrt(3) / 2 * q + np.sqrt(3) * r)
    return -y, x


def hex_segments(rings, seg_radius, seg_gap, rotate=False, antialias=True, flatten=False,
                 pad=2, drop=(0,)):
    inner = seg_radius * np.sqrt(3) / 2
    size = np.ceil((rings * 2 + 1) * inner * 2 + (rings * 2) * seg_gap + pad * 2).astype(int)
    shape = np.broadcast_to(size, (2,))
    masks = []
    if 0 not in drop:
        masks.append(hexagon(shape, seg_radius, shift=(0, 0), antialias=antialias, rotate=rotate))
    seg = 1
    for ring in range(1, rings + 1):
        for h in _hex_ring(ring):
            r, c = _hex_to_rc(h, seg_radius + seg_gap / 2, rotate)
            if seg not in drop:
                masks.append(hexagon(shape, seg_radius, shift=(r, c), antialias=antialias,
                                     rotate=rotate))
            seg += 1
    masks = np.asarray(masks)
    return np.sum(masks, axis=0) if flatten else masks


def centroid(img):
    img = np.asarray(img)
    img = img / np.sum(img)
    nr, nc = img.shape
    rr, cc = np.mgrid[0:nr, 0:nc]
    return np.dot(rr.ravel(), img.ravel()), np.dot(cc.ravel(), img.ravel())


def normalize_power(array, power=1):
    array = np.asarray(array)
    return array * np.sqrt(power / np.sum(np.abs(array) ** 2))


def noll_to_nm(j):
    if j < 1:
        raise ValueError('Noll index must be greater than zero')
    n = int(np.ceil((-1 + np.sqrt(1 + 8 * j)) / 2) - 1)
    if n == 0:
        return 0, 0
    k = (n + 1) * (n + 2) / 2
    r = int(j - k - 1)
    sign = -1 if (j & 1) else 1
    row_m = [1, 1] if (n & 1) else [0]
    for _ in range(n // 2):
        row_m.append(row_m[-1] + 2)
        row_m.append(row_m[-1])
    return int(n), int(row_m[r] * sign)


def _radial(n, m, rho):
    m, n = abs(int(m)), abs(int(n))
    if (n - m) & 1:
        return 0
    out = np.zeros(rho.shape)
    for k in range((n - m) // 2 + 1):
        c = ((-1) ** k * factorial(n - k)
             / (factorial(k) * factorial((n + m) // 2 - k) * factorial((n - m) // 2 - k)))
        out += c * rho ** (n - 2 * k)
    return out


def zernike_coordinates(mask, shift=None, rotate=0):
    mask = np.asarray(mask, dtype=bool)
    if shift is None:
        center = np.asarray(mask.shape) / 2
        cen = centroid(mask)
        shift = (cen[0] - center[0], cen[1] - center[1])
    rr, cc = mesh(mask.shape, shift)
    r = np.abs(rr + 1j * cc)
    rho = r / np.max(r * mask)
    angle = (90 - rotate) * np.pi / 180
    theta = np.angle(-rr * np.exp(1j * angle) + 1j * cc * np.exp(1j * angle))
    return rho, theta


def zernike(mask, index, normalize=True, rho=None, theta=None):
    mask = np.asarray(mask, dtype=bool)
    n, m = noll_to_nm(index)
    if rho is None:
        rho, theta = zernike_coordinates(mask)
    if m == 0:
        if n == 0:
            return mask
        return (np.sqrt(n + 1) if normalize else 1) * _radial(n, m, rho) * mask
    norm = np.sqrt(2) * np.sqrt(n + 1) if normalize else 1
    trig = np.cos(m * theta) if m > 0 else np.sin(m * theta)
    return norm * _radial(n, m, rho) * trig * mask


def zernike_compose(mask, coeffs, normalize=True, rho=None, theta=None):
    mask = np.asarray(mask)
    coeffs = np.asarray(coeffs)
    opd = np.zeros(mask.shape)
    for idx, c in np.ndenumerate(coeffs):
        opd += c * zernike(mask, idx[0] + 1, normalize, rho, theta)
    return opd
