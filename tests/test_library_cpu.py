"""The C-ABI library without a GPU: it loads, exports every entry point
include/cctag_amd.h declares, and its host-only entry points behave as
documented (defaults of Params.hpp:27-66, argument errors, no device)."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from cctag_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cctag_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cctagx_\w+)\s*\(", text)))


def test_header_symbols_are_exported():
    lib = _lib.load()
    names = declared_functions()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_version_and_struct_sizes():
    lib = _lib.load()
    assert lib.cctagx_version().startswith(b"cctag_amd")
    # field-for-field mirrors of the C structs
    assert ctypes.sizeof(_lib.Params) == 27 * 4
    assert ctypes.sizeof(_lib.Marker) == 4 * (7 + 5 + 9)


def test_default_parameters_are_the_references():
    p = _lib.default_params(3)
    exp = dict(canny_thr_low=0.01, canny_thr_high=0.04, dist_search=30,
               thr_gradient_mag_in_vote=2500, angle_voting=0.0, ratio_voting=4.0,
               average_vote_min=0.0, thr_median_distance_ellipse=3.0, maximum_nb_seeds=500,
               maximum_nb_candidates_loop_two=40, n_crowns=3, min_points_segment_candidate=10,
               min_votes_to_select_candidate=3, thresh_robust_estimation_of_outer_ellipse=30.0,
               ellipse_growing_elliptic_hull_width=2.3, window_size_on_inner_elliptic_segment=20,
               number_of_multires_layers=4, number_of_processed_multires_layers=4,
               n_samples_outer_ellipse=150, num_cuts_in_ident_step=22,
               num_samples_outer_edge_points_refinement=20, sample_cut_length=100,
               imaged_center_n_grid_sample=5, imaged_center_neighbour_size=0.2,
               min_ident_proba=1e-6, search_for_another_segment=1, do_identification=1)
    for k, v in exp.items():
        assert getattr(p, k) == pytest.approx(np.float32(v), rel=0, abs=0), k
    assert _lib.default_params(4).n_crowns == 4


def test_argument_errors_without_device():
    lib = _lib.load()
    h = ctypes.c_void_p()
    p = _lib.default_params(3)
    assert lib.cctagx_create(None, 64, 48, ctypes.byref(h)) == -1
    assert lib.cctagx_create(ctypes.byref(p), 0, 48, ctypes.byref(h)) == -1
    bad = _lib.default_params(3)
    bad.n_crowns = 5
    assert lib.cctagx_create(ctypes.byref(bad), 64, 48, ctypes.byref(h)) == -6
    bad = _lib.default_params(3)
    bad.angle_voting = 0.5  # Vote.cpp:95-98 rejects angle voting
    assert lib.cctagx_create(ctypes.byref(bad), 64, 48, ctypes.byref(h)) == -1
    for n in ("cctagx_destroy", "cctagx_stage_count"):
        assert getattr(lib, n)(None) == -1
    n = ctypes.c_int()
    assert lib.cctagx_detect(None, None, 0, 0, 0, 0, None, 0, ctypes.byref(n)) == -1
    assert lib.cctagx_load(None, None, 0, 0, 0) == -1
    assert lib.cctagx_load_bgr(None, None, 0, 0, 0, 3) == -1
    assert lib.cctagx_detect_bgr(None, None, 0, 0, 0, 3, 0, None, 0, ctypes.byref(n)) == -1
    assert lib.cctagx_front_end(None, None, 0, 0, 0) == -1
    assert lib.cctagx_last_markers(None, None, 0, ctypes.byref(n)) == -1
    bank = np.ones((32, 5), np.float32)
    assert lib.cctagx_set_bank(None, bank.ctypes.data, 32, 5) == -1
    assert lib.cctagx_identify(None, None, 0, None) == -1


def test_create_without_device_fails_cleanly():
    lib = _lib.load()
    if lib.cctagx_device_count() > 0:
        pytest.skip("a device is visible")
    h = ctypes.c_void_p()
    p = _lib.default_params(3)
    assert lib.cctagx_create(ctypes.byref(p), 64, 48, ctypes.byref(h)) == -5
    assert not h.value
    assert b"device" in lib.cctagx_last_error()


MIRRORS = [("Params", "cctagx_params"), ("LevelView", "cctagx_level_view"),
           ("Marker", "cctagx_marker"), ("IdentMarker", "cctagx_ident_marker"),
           ("IdentResult", "cctagx_ident_result"), ("DetectStats", "cctagx_detect_stats"),
           ("BatchStats", "cctagx_batch_stats"), ("RefineMarker", "cctagx_refine_marker")]


def test_ctypes_mirrors_match_the_c_layout(tmp_path):
    """Every ctypes mirror has the size and field offsets the C compiler
    gives the header's struct (a probe built with gcc against the header)."""
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "cctag_amd.h"',
             'int main(void) {']
    for py, c in MIRRORS:
        lines.append(f'  printf("{py} size %zu\\n", sizeof({c}));')
        for name, _ in getattr(_lib, py)._fields_:
            lines.append(f'  printf("{py} {name} %zu\\n", offsetof({c}, {name}));')
    lines += ['  return 0;', '}']
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.dirname(HEADER),
                    str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    c_layout = {}
    for line in out.splitlines():
        py, name, v = line.split()
        c_layout[(py, name)] = int(v)
    for py, _ in MIRRORS:
        cls = getattr(_lib, py)
        assert c_layout[(py, "size")] == ctypes.sizeof(cls), py
        for name, _ in cls._fields_:
            assert c_layout[(py, name)] == getattr(cls, name).offset, (py, name)
