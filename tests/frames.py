"""Test inputs: seeded synthetic CCTag frames plus small edge cases."""
from __future__ import annotations

import functools

import numpy as np

from cctag_amd import synth


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
