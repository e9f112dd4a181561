"""Seeded synthetic CCTag frames with exact ground truth.

Each marker is a three-ring CCTag drawn through a plane-to-image homography
H (marker plane, unit outer radius -> pixels).  The radial profile follows the
marker bank exactly as the identification step reads it
(Identification.cpp:134-153): at normalized radius rho < 1 the colour is
white when #{j : 1/rr_j <= rho} is even and black otherwise, with a white
quiet zone for 1 <= rho <= 1.4 (markersToPrint/generators/generate.py:18-21).
The ground-truth imaged centre is H.(0,0,1).

Pixel convention: the renderer integrates pixel (x, y) over the square
[x, x+1) x [y, y+1).  Ground-truth centres are reported in the detector's
coordinates, where the value of pixel (x, y) sits at (x, y) (the bilinear
sampler of Identification.hpp:197-222), i.e. H.(0,0,1) - (0.5, 0.5).

Frame k of configuration c uses seed 1000*c + k.  Everything is numpy with a
PCG64 generator, so a frame is byte-identical on any machine with the same
numpy version.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np
from scipy import ndimage

from .bank import bank_ratios

BLACK = 20.0
WHITE = 235.0


@dataclasses.dataclass
class MarkerGT:
    id: int
    center: tuple          # ground-truth imaged centre, detector coordinates
    radius: float          # outer radius in px at zero tilt
    H: np.ndarray          # 3x3 plane -> image
    occluded: bool = False


@dataclasses.dataclass
class FrameConfig:
    width: int
    height: int
    n_markers: tuple       # (min, max) inclusive
    radius: tuple          # outer radius range, px
    max_tilt_deg: float
    blur: tuple            # gaussian sigma range
    noise: float
    occlusion_frac: float = 0.0
    distractors: tuple = (0, 0)
    distractor_radius: tuple = (5, 300)
    motion_blur: float = 0.0
    perlin_amp: float = 20.0
    n_crowns: int = 3


CONFIGS = {
    # 1280x720, 4 markers, r 40-120, tilt <= 45, blur 1, noise 4
    1: FrameConfig(1280, 720, (4, 4), (40, 120), 45.0, (1.0, 1.0), 4.0),
    # 1920x1080, 16 markers, r 25-200, tilt <= 60, 20% occluded, blur 0.5-2, noise 6
    2: FrameConfig(1920, 1080, (16, 16), (25, 200), 60.0, (0.5, 2.0), 6.0,
                   occlusion_frac=0.2),
    # 3840x2160, 32 markers, r 15-400, tilt <= 70, 200 distractors, motion blur <= 7, noise 8
    3: FrameConfig(3840, 2160, (32, 32), (15, 400), 70.0, (0.5, 1.0), 8.0,
                   distractors=(200, 200), motion_blur=7.0),
}

BATCH_SIZES = [(640, 480), (1280, 720), (1920, 1080), (3840, 2160)]


def batch_config(rng: np.random.Generator) -> FrameConfig:
    """One frame of the uneven-size batch: size drawn from BATCH_SIZES,
    1-32 markers and 0-100 distractors."""
    w, h = BATCH_SIZES[int(rng.integers(0, len(BATCH_SIZES)))]
    rmax = min(200.0, 0.18 * min(w, h))
    return FrameConfig(w, h, (1, 32), (15, rmax), 60.0, (0.5, 2.0), 6.0,
                       occlusion_frac=0.2, distractors=(0, 100))


def _perlin(h: int, w: int, cell: int, rng: np.random.Generator) -> np.ndarray:
    """Classic 2-D gradient noise (Perlin 1985) on a square lattice of
    `cell` px, values roughly in [-1, 1]."""
    gh, gw = h // cell + 2, w // cell + 2
    ang = rng.uniform(0, 2 * np.pi, size=(gh, gw))
    gx, gy = np.cos(ang), np.sin(ang)
    ys = (np.arange(h) + 0.5) / cell
    xs = (np.arange(w) + 0.5) / cell
    y0 = np.floor(ys).astype(int)
    x0 = np.floor(xs).astype(int)
    fy = (ys - y0)[:, None]
    fx = (xs - x0)[None, :]

    def fade(t):
        return t * t * t * (t * (t * 6 - 15) + 10)

    def dot(iy, ix, dy, dx):
        return gx[iy[:, None], ix[None, :]] * dx + gy[iy[:, None], ix[None, :]] * dy

    n00 = dot(y0, x0, fy, fx)
    n01 = dot(y0, x0 + 1, fy, fx - 1)
    n10 = dot(y0 + 1, x0, fy - 1, fx)
    n11 = dot(y0 + 1, x0 + 1, fy - 1, fx - 1)
    u, v = fade(fx), fade(fy)
    nx0 = n00 + u * (n01 - n00)
    nx1 = n10 + u * (n11 - n10)
    return (nx0 + v * (nx1 - nx0)) * math.sqrt(2.0)


def marker_homography(cx: float, cy: float, radius: float, tilt: float,
                      tilt_axis: float, rot: float, focal: float) -> np.ndarray:
    """Plane-to-image homography of a unit-radius marker seen by a pinhole
    camera (focal `focal` px, principal point at the marker centre) after the
    plane is rotated in-plane by `rot`, tilted by `tilt` about an axis at
    azimuth `tilt_axis`.  At zero tilt the outer circle images to radius px."""
    def rz(a):
        c, s = math.cos(a), math.sin(a)
        return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1.0]])

    def rx(a):
        c, s = math.cos(a), math.sin(a)
        return np.array([[1, 0, 0], [0, c, -s], [0, s, c]])

    R = rz(tilt_axis) @ rx(tilt) @ rz(-tilt_axis) @ rz(rot)
    D = focal / radius
    K = np.array([[focal, 0, cx], [0, focal, cy], [0, 0, 1.0]])
    M = np.column_stack([R[:, 0], R[:, 1], np.array([0.0, 0.0, D])])
    H = K @ M
    return H / H[2, 2]


def _render_marker(img: np.ndarray, H: np.ndarray, ratios, ss: int = 4) -> None:
    """Draws one marker (and its quiet zone) into img, box-filtered over
    ss x ss sub-samples per pixel."""
    h, w = img.shape
    Hinv = np.linalg.inv(H)
    # bounding box of the quiet-zone boundary rho = 1.42
    t = np.linspace(0, 2 * np.pi, 256, endpoint=False)
    pts = np.stack([1.42 * np.cos(t), 1.42 * np.sin(t), np.ones_like(t)])
    q = H @ pts
    qx, qy = q[0] / q[2], q[1] / q[2]
    x0 = max(int(np.floor(qx.min())) - 1, 0)
    x1 = min(int(np.ceil(qx.max())) + 2, w)
    y0 = max(int(np.floor(qy.min())) - 1, 0)
    y1 = min(int(np.ceil(qy.max())) + 2, h)
    if x0 >= x1 or y0 >= y1:
        return
    inv_r = np.sort(1.0 / np.asarray(ratios, dtype=np.float64))
    sub = (np.arange(ss) + 0.5) / ss
    xs = (np.arange(x0, x1)[:, None] + sub[None, :]).reshape(-1)
    ys = (np.arange(y0, y1)[:, None] + sub[None, :]).reshape(-1)
    X, Y = np.meshgrid(xs, ys)
    u = Hinv[0, 0] * X + Hinv[0, 1] * Y + Hinv[0, 2]
    v = Hinv[1, 0] * X + Hinv[1, 1] * Y + Hinv[1, 2]
    wq = Hinv[2, 0] * X + Hinv[2, 1] * Y + Hinv[2, 2]
    rho = np.hypot(u / wq, v / wq)
    ldum = np.searchsorted(inv_r, rho, side="right")
    val = np.where(ldum % 2 == 0, WHITE, BLACK)
    val = np.where(rho >= 1.0, WHITE, val)
    inside = rho <= 1.4
    nh, nw = y1 - y0, x1 - x0
    val = val.reshape(nh, ss, nw, ss)
    inside = inside.reshape(nh, ss, nw, ss)
    cov = inside.mean(axis=(1, 3))
    col = np.where(inside, val, 0.0).sum(axis=(1, 3)) / np.maximum(inside.sum(axis=(1, 3)), 1)
    patch = img[y0:y1, x0:x1]
    img[y0:y1, x0:x1] = patch * (1 - cov) + col * cov


def _render_distractor(img, rng, rmin, rmax):
    """A filled ellipse or an elliptical ring of random contrast."""
    h, w = img.shape
    a = rng.uniform(rmin, rmax)
    b = a * rng.uniform(0.3, 1.0)
    cx, cy = rng.uniform(0, w), rng.uniform(0, h)
    ang = rng.uniform(0, np.pi)
    level = rng.uniform(10, 245)
    ring = rng.random() < 0.5
    thick = rng.uniform(0.08, 0.3)
    x0, x1 = max(int(cx - a) - 2, 0), min(int(cx + a) + 3, w)
    y0, y1 = max(int(cy - a) - 2, 0), min(int(cy + a) + 3, h)
    if x0 >= x1 or y0 >= y1:
        return
    ss = 2
    sub = (np.arange(ss) + 0.5) / ss
    xs = (np.arange(x0, x1)[:, None] + sub[None, :]).reshape(-1) - cx
    ys = (np.arange(y0, y1)[:, None] + sub[None, :]).reshape(-1) - cy
    X, Y = np.meshgrid(xs, ys)
    c, s = math.cos(ang), math.sin(ang)
    u = (c * X + s * Y) / a
    v = (-s * X + c * Y) / b
    r = np.hypot(u, v)
    m = r <= 1.0
    if ring:
        m &= r >= 1.0 - thick
    cov = m.reshape(y1 - y0, ss, x1 - x0, ss).mean(axis=(1, 3))
    patch = img[y0:y1, x0:x1]
    img[y0:y1, x0:x1] = patch * (1 - cov) + level * cov


def _motion_kernel(length: float, angle: float) -> np.ndarray:
    n = int(math.ceil(length)) | 1
    k = np.zeros((n, n))
    c = n // 2
    for t in np.linspace(-length / 2, length / 2, 8 * n):
        x = c + t * math.cos(angle)
        y = c + t * math.sin(angle)
        k[int(round(y)), int(round(x))] += 1
    return k / k.sum()


def make_frame(cfg: FrameConfig, seed: int, n_markers: int | None = None):
    """Returns (uint8 image HxW, list[MarkerGT])."""
    rng = np.random.Generator(np.random.PCG64(seed))
    w, h = cfg.width, cfg.height
    img = 128.0 + cfg.perlin_amp * _perlin(h, w, 256, rng)
    ratios = bank_ratios(cfg.n_crowns)
    if n_markers is None:
        n_markers = int(rng.integers(cfg.n_markers[0], cfg.n_markers[1] + 1))
    nd = int(rng.integers(cfg.distractors[0], cfg.distractors[1] + 1))
    for _ in range(nd):
        _render_distractor(img, rng, *cfg.distractor_radius)
    focal = 2.0 * max(w, h)
    placed = []  # (cx, cy, bounding radius)
    gts = []
    for i in range(n_markers):
        mid = int(rng.integers(0, len(ratios)))
        r = rng.uniform(*cfg.radius)
        ok = False
        for attempt in range(400):
            if attempt and attempt % 50 == 0:
                r = max(cfg.radius[0], r * 0.8)
            rb = 1.5 * r
            if 2 * rb >= min(w, h):
                r *= 0.8
                continue
            cx = rng.uniform(rb, w - rb)
            cy = rng.uniform(rb, h - rb)
            if all(math.hypot(cx - px, cy - py) > rb + pr for px, py, pr in placed):
                ok = True
                break
        if not ok:
            continue
        tilt = math.radians(rng.uniform(0, cfg.max_tilt_deg))
        H = marker_homography(cx, cy, r, tilt, rng.uniform(0, 2 * np.pi),
                              rng.uniform(0, 2 * np.pi), focal)
        # clear the distractor clutter under the marker's quiet zone
        _render_marker(img, H, ratios[mid])
        placed.append((cx, cy, 1.5 * r))
        occluded = bool(rng.random() < cfg.occlusion_frac)
        if occluded:
            phi = rng.uniform(0, 2 * np.pi)
            p = H @ np.array([math.cos(phi), math.sin(phi), 1.0])
            ox, oy = p[0] / p[2], p[1] / p[2]
            half = rng.uniform(0.15, 0.3) * r
            xa, xb = int(max(ox - half, 0)), int(min(ox + half, w))
            ya, yb = int(max(oy - half, 0)), int(min(oy + half, h))
            img[ya:yb, xa:xb] = 25.0
        gts.append(MarkerGT(mid, (cx - 0.5, cy - 0.5), r, H, occluded))
    sigma = rng.uniform(*cfg.blur)
    if sigma > 0:
        img = ndimage.gaussian_filter(img, sigma, mode="nearest")
    if cfg.motion_blur > 0:
        k = _motion_kernel(rng.uniform(1.0, cfg.motion_blur), rng.uniform(0, np.pi))
        img = ndimage.convolve(img, k, mode="nearest")
    img = img + rng.normal(0.0, cfg.noise, size=img.shape)
    return np.clip(np.rint(img), 0, 255).astype(np.uint8), gts


def border_frame(seed: int = 7000, w: int = 640, h: int = 480):
    """Markers cut by the frame border: four markers (outer radius 40-60 px,
    tilt <= 30 deg) whose centres lie 2-4 px inside the left, right, top and
    bottom borders, and two whole markers inside; blur 1 px, noise 4 DN.
    Returns (uint8 image, list[MarkerGT])."""
    rng = np.random.Generator(np.random.PCG64(seed))
    img = 128.0 + 20.0 * _perlin(h, w, 256, rng)
    ratios = bank_ratios(3)
    focal = 2.0 * max(w, h)
    d = lambda: rng.uniform(2.0, 4.0)
    centres = [(d(), rng.uniform(0.3, 0.7) * h), (w - d(), rng.uniform(0.3, 0.7) * h),
               (rng.uniform(0.3, 0.7) * w, d()), (rng.uniform(0.3, 0.7) * w, h - d()),
               (0.3 * w, 0.5 * h), (0.7 * w, 0.5 * h)]
    gts = []
    for cx, cy in centres:
        mid = int(rng.integers(0, len(ratios)))
        r = rng.uniform(40.0, 60.0)
        H = marker_homography(cx, cy, r, math.radians(rng.uniform(0, 30)),
                              rng.uniform(0, 2 * np.pi), rng.uniform(0, 2 * np.pi), focal)
        _render_marker(img, H, ratios[mid])
        gts.append(MarkerGT(mid, (cx - 0.5, cy - 0.5), r, H))
    img = ndimage.gaussian_filter(img, 1.0, mode="nearest")
    img = img + rng.normal(0.0, 4.0, size=img.shape)
    return np.clip(np.rint(img), 0, 255).astype(np.uint8), gts


FOUR_CROWNS = FrameConfig(1280, 720, (6, 6), (40, 120), 45.0, (1.0, 1.0), 4.0, n_crowns=4)


def four_crown_frame(k: int = 0):
    """A 1280x720 frame with 6 four-ring markers (IDs from the 128-row
    bank), otherwise as the 1280x720 configuration; seed 6000 + k."""
    return make_frame(FOUR_CROWNS, 6000 + k)


def config_frame(config: int, k: int = 0):
    """Frame k of configuration 1, 2 or 3."""
    return make_frame(CONFIGS[config], 1000 * config + k)


def frame_1080p(k: int):
    """Frame k of the 1920x1080 16-marker configuration (frame 0 is the one
    the plain-path benchmark detects)."""
    return config_frame(2, k)


def batch_frame(k: int, seed: int = 4000):
    """Frame k of the uneven-size batch (size drawn from BATCH_SIZES)."""
    rng = np.random.Generator(np.random.PCG64(seed + k))
    return make_frame(batch_config(rng), seed + 10000 + k)


def batch_frames(n: int = 64, seed: int = 4000):
    """The uneven-size batch: n frames, sizes drawn from BATCH_SIZES."""
    return [batch_frame(k, seed) for k in range(n)]


# ---------------------------------------------------------------------------
# Inputs of the centre refinement built from ground truth: the outer ellipse
# (the image of the unit circle), cut stops on it, the ellipse centre as the
# start and the true imaged centre, in detector coordinates (pixel centres at
# integer coordinates).

DETECTOR_SHIFT = np.array([[1.0, 0.0, -0.5], [0.0, 1.0, -0.5], [0.0, 0.0, 1.0]])


def ellipse_of(H: np.ndarray):
    """(cx, cy, a, b, angle) of the image of the unit circle under H, a >= b."""
    Hi = np.linalg.inv(H)
    C = Hi.T @ np.diag([1.0, 1.0, -1.0]) @ Hi
    A = C[:2, :2]
    ctr = np.linalg.solve(A, -C[:2, 2])
    f = C[2, 2] + C[:2, 2] @ ctr
    lam, vec = np.linalg.eigh(A)
    if f > 0:
        lam, f = -lam, -f
    ax = np.sqrt(-f / lam)
    i = int(np.argmax(ax))
    angle = math.atan2(vec[1, i], vec[0, i])
    return float(ctr[0]), float(ctr[1]), float(ax[i]), float(ax[1 - i]), angle


def refine_marker(H_synth: np.ndarray, n_cuts: int = 22, phase: float = 0.1) -> dict:
    H = DETECTOR_SHIFT @ np.asarray(H_synth, np.float64)
    th = phase + 2 * np.pi * np.arange(n_cuts) / n_cuts
    pts = H @ np.stack([np.cos(th), np.sin(th), np.ones_like(th)])
    stops = (pts[:2] / pts[2]).T
    e = ellipse_of(H)
    c = H @ np.array([0.0, 0.0, 1.0])
    return {"ellipse": e, "center0": (e[0], e[1]), "stops": stops,
            "center_gt": (c[0] / c[2], c[1] / c[2]), "H": H / H[2, 2]}


def refine_mosaic(n: int = 4096, tile: int = 160, unique: int = 256, seed: int = 5000):
    """n candidate markers for the refinement workload: `unique` markers
    (outer radius 20-36 px, tilt <= 60 deg seen by a camera whose focal
    length is 4-8 marker radii, so the ellipse centre is off the imaged
    centre; random in-plane rotation and ID) each rendered into a tile x
    tile patch with blur 0.5-1.5 px, tiled over a square mosaic, then noise
    6 DN over the whole image.  Returns (uint8 image, list of refine_marker
    dicts)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    ratios = bank_ratios(3)
    side = int(math.ceil(math.sqrt(n)))
    tiles, hs = [], []
    for _ in range(unique):
        t = np.full((tile, tile), 128.0)
        r = rng.uniform(20.0, 36.0)
        c = tile / 2 + rng.uniform(-3, 3, size=2)
        H = marker_homography(c[0], c[1], r, math.radians(rng.uniform(0, 60)),
                              rng.uniform(0, 2 * np.pi), rng.uniform(0, 2 * np.pi),
                              r * rng.uniform(4.0, 8.0))
        _render_marker(t, H, ratios[int(rng.integers(0, len(ratios)))])
        tiles.append(ndimage.gaussian_filter(t, rng.uniform(0.5, 1.5), mode="nearest"))
        hs.append(H)
    img = np.empty((side * tile, side * tile))
    markers = []
    for k in range(side * side):
        u = k % unique
        ty, tx = divmod(k, side)
        img[ty * tile:(ty + 1) * tile, tx * tile:(tx + 1) * tile] = tiles[u]
        if k < n:
            T = np.array([[1.0, 0.0, tx * tile], [0.0, 1.0, ty * tile], [0.0, 0.0, 1.0]])
            markers.append(refine_marker(T @ hs[u]))
    img += rng.normal(0.0, 6.0, size=img.shape)
    return np.clip(np.rint(img), 0, 255).astype(np.uint8), markers
