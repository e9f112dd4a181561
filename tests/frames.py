"""Test inputs: seeded synthetic CCTag frames plus small edge cases."""
from __future__ import annotations

import functools
import math

import numpy as np
from scipy import ndimage

from cctag_amd import synth
from cctag_amd.bank import bank_ratios


@functools.lru_cache(maxsize=None)
def config_frame(config: int, k: int = 0):
    img, gts = synth.config_frame(config, k)
    img.setflags(write=False)
    return img, gts


def small_marker_frame(seed: int = 7, w: int = 320, h: int = 240):
    """A small frame with two markers, cheap enough for the oracle."""
    cfg = synth.FrameConfig(w, h, (2, 2), (30, 50), 40.0, (1.0, 1.0), 4.0)
    return synth.make_frame(cfg, seed)


def edge_case(name: str) -> np.ndarray:
    rng = np.random.Generator(np.random.PCG64(99))
    if name == "blank":
        return np.full((64, 96), 128, np.uint8)
    if name == "noise":
        return rng.integers(0, 256, size=(128, 160), dtype=np.uint8)
    if name == "checker":
        y, x = np.mgrid[0:96, 0:128]
        return np.where(((x // 8) + (y // 8)) % 2 == 0, 30, 220).astype(np.uint8)
    if name == "tiny":
        return rng.integers(0, 256, size=(16, 16), dtype=np.uint8)
    if name == "rings":
        y, x = np.mgrid[0:256, 0:256]
        r = np.hypot(x - 127.3, y - 121.8)
        return np.where((r.astype(int) // 9) % 2 == 0, 40, 210).astype(np.uint8)
    if name == "spiral":
        # a band along a spiral through many 64x16 tiles, 4 DN above the
        # background (weak edges, m ~ 3) except one 40-degree arc at +60 DN
        # (strong): the whole band's edges come from one strong arc far away
        y, x = np.mgrid[0:300, 0:420]
        dx, dy = x - 211.5, y - 148.5
        r, a = np.hypot(dx, dy), np.arctan2(dy, dx)
        turn = (r - 12.0 * (a + np.pi) / (2 * np.pi)) / 12.0
        band = np.abs(turn - np.round(turn)) < 0.22
        img = np.full(x.shape, 128, np.int32)
        img[band & (r > 10)] = 132
        img[band & (r > 100) & (r < 125) & (np.abs(a - 0.6) < 0.35)] = 188
        return img.astype(np.uint8)
    if name == "stripes":
        # diagonal and near-horizontal stripes: rows of candidates that fill
        # whole 64-pixel tile rows and join across tile edges diagonally
        y, x = np.mgrid[0:200, 0:330]
        a = ((x + y) // 3) % 2
        b = ((7 * x + y) // 4) % 2
        return np.where((x < 165) ^ (y < 100), 60 + 130 * a, 70 + 120 * b).astype(np.uint8)
    raise KeyError(name)


@functools.lru_cache(maxsize=None)
def small_batch():
    """Five frames of four sizes for the batched path: a 640x480 frame of the
    uneven-size batch (markers + distractors), the 1280x720 four-marker
    frame, the small two-marker frame, concentric rings and a blank frame."""
    cfg = synth.FrameConfig(640, 480, (6, 6), (15, 80), 60.0, (0.5, 2.0), 6.0,
                            occlusion_frac=0.2, distractors=(20, 20))
    out = [synth.make_frame(cfg, 4242)[0], config_frame(1)[0], small_marker_frame()[0],
           edge_case("rings"), edge_case("blank")]
    for img in out:
        img.setflags(write=False)
    return out


# Frame sizes whose pyramid does not halve exactly at every level: a 1366x768
# panel (683x384 -> 341x192 -> 170x96), odd sides from level 0 on, and a 1080p
# frame cropped by one column and row.
ODD_SIZES = [(1366, 768), (1001, 751), (1919, 1079)]


@functools.lru_cache(maxsize=None)
def odd_size_frame(w: int, h: int, seed: int = 8100):
    """Markers as in the 1280x720 configuration, at a frame size whose
    floor-halved pyramid has levels that are not exact halves."""
    cfg = synth.FrameConfig(w, h, (5, 5), (30, 120), 45.0, (0.8, 1.5), 5.0,
                            occlusion_frac=0.2, distractors=(10, 10))
    img, gts = synth.make_frame(cfg, seed + w + h)
    img.setflags(write=False)
    return img, gts


@functools.lru_cache(maxsize=None)
def border_frame(seed: int = 7000):
    img, gts = synth.border_frame(seed)
    img.setflags(write=False)
    return img, gts


@functools.lru_cache(maxsize=None)
def batch_frame(k: int):
    img, gts = synth.batch_frame(k)
    img.setflags(write=False)
    return img, gts


# ---------------------------------------------------------------------------
# identification inputs built from ground truth

def outer_points(e, n=150, mag=100.0):
    """n points on the ellipse (cx, cy, a, b, angle) with outward normals
    (x, y, gx, gy)."""
    cx, cy, a, b, ang = e
    t = 2 * np.pi * np.arange(n) / n
    ca, sa = np.cos(ang), np.sin(ang)
    x = cx + a * np.cos(t) * ca - b * np.sin(t) * sa
    y = cy + a * np.cos(t) * sa + b * np.sin(t) * ca
    u, v = (x - cx) * ca + (y - cy) * sa, -(x - cx) * sa + (y - cy) * ca
    gu, gv = u / a ** 2, v / b ** 2
    gx, gy = gu * ca - gv * sa, gu * sa + gv * ca
    nrm = np.hypot(gx, gy)
    return np.column_stack([x, y, mag * gx / nrm, mag * gy / nrm])


def ident_input(H_synth):
    """The identification input of a synthetic marker: its outer ellipse,
    150 points on it and the ellipse centre as the starting imaged centre."""
    m = synth.refine_marker(H_synth)
    e = m["ellipse"]
    return {"ellipse": e, "points": outer_points(e), "center0": m["center0"],
            "center_gt": m["center_gt"]}


def clean_marker(seed, w=320, h=240, radius=(40, 80), tilt=40.0, blur=0.8):
    """One marker rendered without noise: (image, its ID, ident_input)."""
    rng = np.random.default_rng(seed)
    mid = int(rng.integers(0, 32))
    img = np.full((h, w), 128.0)
    H = synth.marker_homography(w / 2 + rng.uniform(-3, 3), h / 2 + rng.uniform(-3, 3),
                                rng.uniform(*radius), math.radians(rng.uniform(0, tilt)),
                                rng.uniform(0, 2 * np.pi), rng.uniform(0, 2 * np.pi),
                                2.0 * max(w, h))
    synth._render_marker(img, H, bank_ratios(3)[mid])
    img = ndimage.gaussian_filter(img, blur, mode="nearest")
    return np.clip(np.rint(img), 0, 255).astype(np.uint8), mid, ident_input(H)


def faded_sector(img, e, lo=0.2, width=np.pi / 2, contrast=0.63):
    """img with the sector [lo, lo + width) of angles about the ellipse
    centre shown at `contrast` of its contrast around 128."""
    f = img.astype(np.float64)
    yy, xx = np.mgrid[0:img.shape[0], 0:img.shape[1]]
    ang = np.arctan2(yy - e[1], xx - e[0])
    sector = (ang > lo) & (ang < lo + width)
    f[sector] = 128 + (f[sector] - 128) * contrast
    return np.clip(np.rint(f), 0, 255).astype(np.uint8)


def failing_marker():
    """A marker input whose cuts are all collected but whose refinement cuts
    all leave the image, so no cut is selected (status -2)."""
    ys = np.linspace(280.0, 320.0, 12)
    return {"ellipse": (6.0, 300.0, 40.0, 40.0, 0.0),
            "points": np.column_stack([np.full(12, 3.0), ys, np.full(12, -100.0), np.zeros(12)]),
            "center0": (6.0, 300.0)}
