"""The library's host candidate stage against the oracle's independent
restatement of it (Detection.cpp:69-693, Vote.cpp:249-742,
EllipseGrowing.cpp, Multiresolution.cpp:43-413), on the oracle front end's
levels of the same frame, without a GPU: the same marker candidates in the
same order with bit-identical outer ellipses (level and level 0), centres,
qualities and ring point lists.  The library runs its task-parallel form
(speculative loop-one linking, one task flow for loops two and three and the
re-fit) on several threads, the oracle its serial loops."""
import ctypes

import numpy as np
import pytest

from cctag_amd import _lib, synth
from tests import frames, oracle_lib


def mismatch(img, params=None, threads=8):
    o = oracle_lib.load()
    f = o.oracle_candidate_mismatch
    f.argtypes = [ctypes.POINTER(_lib.Params), ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                  ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.c_char_p, ctypes.c_int]
    img = np.ascontiguousarray(img, np.uint8)
    h, w = img.shape
    p = params if params is not None else _lib.default_params(3)
    n = ctypes.c_int()
    msg = ctypes.create_string_buffer(512)
    rc = f(ctypes.byref(p), img.ctypes.data, w, h, threads, ctypes.byref(n), msg, 512)
    return rc, n.value, msg.value.decode()


def check(img, params=None, threads=8, min_markers=1):
    rc, n, msg = mismatch(img, params, threads)
    assert rc == 0, msg
    assert n >= min_markers


@pytest.mark.parametrize("config,k", [(1, 0), (1, 1), (1, 2), (1, 3), (2, 0), (2, 1)])
def test_candidate_stage_configurations(config, k):
    img, gts = frames.config_frame(config, k)
    check(img, min_markers=len(gts))


@pytest.mark.parametrize("threads", [1, 3, 16])
def test_candidate_stage_thread_counts(threads):
    check(frames.config_frame(2, 2)[0], threads=threads, min_markers=16)


def test_candidate_stage_4k_distractors():
    """3840x2160 with 200 distractor ellipses and motion blur."""
    check(frames.config_frame(3, 0)[0], min_markers=32)


@pytest.mark.parametrize("k", [2, 3, 7])
def test_candidate_stage_batch_frames(k):
    check(frames.batch_frame(k)[0])


def test_candidate_stage_border_markers():
    check(frames.border_frame()[0])


@pytest.mark.parametrize("w,h", frames.ODD_SIZES)
def test_candidate_stage_floor_halved_pyramid(w, h):
    check(frames.odd_size_frame(w, h)[0], min_markers=5)


def test_candidate_stage_four_crowns():
    check(synth.four_crown_frame(1)[0], params=_lib.default_params(4), min_markers=6)


@pytest.mark.parametrize("name", ["blank", "noise", "checker", "rings"])
def test_candidate_stage_edge_cases(name):
    check(frames.edge_case(name), min_markers=0)


def test_candidate_stage_without_segment_search():
    p = _lib.default_params(3)
    p.search_for_another_segment = 0
    check(frames.config_frame(2, 0)[0], params=p, min_markers=16)
