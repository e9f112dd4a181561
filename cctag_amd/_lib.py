"""ctypes bindings of libcctag_amd.so (include/cctag_amd.h).

The library is loaded from this package directory only; a missing library is
an error, never a silent fallback.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcctag_amd.so")

OK = 0
ERRORS = {
    -1: "bad argument",
    -2: "HIP runtime error",
    -3: "capacity exceeded",
    -4: "call order violated",
    -5: "no gfx950 device",
    -6: "unsupported size or parameter",
}

c_i32 = ctypes.c_int32
c_f32 = ctypes.c_float
P = ctypes.POINTER


class Params(ctypes.Structure):
    """cctagx_params, field for field (Params.hpp:108-199)."""
    _fields_ = [
        ("canny_thr_low", c_f32),
        ("canny_thr_high", c_f32),
        ("dist_search", c_i32),
        ("thr_gradient_mag_in_vote", c_i32),
        ("angle_voting", c_f32),
        ("ratio_voting", c_f32),
        ("average_vote_min", c_f32),
        ("thr_median_distance_ellipse", c_f32),
        ("maximum_nb_seeds", c_i32),
        ("maximum_nb_candidates_loop_two", c_i32),
        ("n_crowns", c_i32),
        ("min_points_segment_candidate", c_i32),
        ("min_votes_to_select_candidate", c_i32),
        ("thresh_robust_estimation_of_outer_ellipse", c_f32),
        ("ellipse_growing_elliptic_hull_width", c_f32),
        ("window_size_on_inner_elliptic_segment", c_i32),
        ("number_of_multires_layers", c_i32),
        ("number_of_processed_multires_layers", c_i32),
        ("n_samples_outer_ellipse", c_i32),
        ("num_cuts_in_ident_step", c_i32),
        ("num_samples_outer_edge_points_refinement", c_i32),
        ("sample_cut_length", c_i32),
        ("imaged_center_n_grid_sample", c_i32),
        ("imaged_center_neighbour_size", c_f32),
        ("min_ident_proba", c_f32),
        ("search_for_another_segment", c_i32),
        ("do_identification", c_i32),
    ]


class LevelView(ctypes.Structure):
    """cctagx_level_view."""
    _fields_ = [
        ("width", c_i32), ("height", c_i32), ("shape_w", c_i32), ("shape_h", c_i32),
        ("img", P(ctypes.c_uint8)), ("dx", P(ctypes.c_int16)), ("dy", P(ctypes.c_int16)),
        ("edges", P(ctypes.c_uint8)), ("edge_map", P(c_i32)),
        ("n_edges", c_i32),
        ("edge_x", P(ctypes.c_int16)), ("edge_y", P(ctypes.c_int16)),
        ("edge_dx", P(ctypes.c_int16)), ("edge_dy", P(ctypes.c_int16)),
        ("before", P(c_i32)), ("after", P(c_i32)),
        ("chosen", P(c_i32)), ("total_dist", P(c_f32)),
        ("voters_index", P(c_i32)), ("voters_list", P(c_i32)),
        ("flow_length", P(c_f32)), ("is_max", P(c_i32)),
        ("n_seeds", c_i32), ("seeds", P(c_i32)),
    ]


class Marker(ctypes.Structure):
    """cctagx_marker (ICCTag.hpp:28-82 plus quality / level / homography)."""
    _fields_ = [
        ("x", c_f32), ("y", c_f32), ("id", c_i32), ("status", c_i32), ("quality", c_f32),
        ("pyramid_level", c_i32), ("scale", c_f32), ("ellipse", c_f32 * 5),
        ("homography", c_f32 * 9),
    ]

    def as_dict(self) -> dict:
        return {"x": self.x, "y": self.y, "id": self.id, "status": self.status,
                "quality": self.quality, "pyramid_level": self.pyramid_level,
                "scale": self.scale, "ellipse": list(self.ellipse),
                "homography": list(self.homography)}


class IdentMarker(ctypes.Structure):
    """cctagx_ident_marker."""
    _fields_ = [
        ("ellipse", c_f32 * 5), ("center", c_f32 * 2), ("scale", c_f32), ("n_points", c_i32),
        ("points", P(c_f32)),
    ]


class IdentResult(ctypes.Structure):
    """cctagx_ident_result."""
    _fields_ = [
        ("status", c_i32), ("id", c_i32), ("x", c_f32), ("y", c_f32), ("quality", c_f32),
        ("homography", c_f32 * 9), ("n_cuts", c_i32), ("stops", (c_f32 * 2) * 32),
    ]

    def as_dict(self) -> dict:
        n = self.n_cuts
        return {"status": self.status, "id": self.id, "x": self.x, "y": self.y,
                "quality": self.quality, "homography": list(self.homography),
                "stops": np.array([[self.stops[k][0], self.stops[k][1]] for k in range(n)],
                                  np.float32).reshape(n, 2)}


def ident_markers(markers):
    """A cctagx_ident_marker array for marker dicts with "ellipse" (cx, cy,
    a, b, angle), "center0" (x, y), "points" (n x 4: x, y, gx, gy) and an
    optional "scale"; returns (array, the point buffers it refers to)."""
    arr = (IdentMarker * max(len(markers), 1))()
    keep = []
    for i, m in enumerate(markers):
        pts = np.ascontiguousarray(np.asarray(m["points"], np.float32).reshape(-1, 4))
        keep.append(pts)
        arr[i].ellipse[:] = [float(v) for v in m["ellipse"]]
        arr[i].center[:] = [float(v) for v in m["center0"]]
        arr[i].scale = float(m.get("scale", 1.0))
        arr[i].n_points = len(pts)
        arr[i].points = pts.ctypes.data_as(P(c_f32))
    return arr, keep


class DetectStats(ctypes.Structure):
    """cctagx_detect_stats."""
    _fields_ = [
        ("front_end_us", c_f32), ("stage_us", c_f32), ("host_us", c_f32),
        ("select_us", c_f32), ("refine_us", c_f32),
        ("n_select", c_i32), ("n_refine", c_i32),
        ("select_samples", ctypes.c_int64), ("refine_samples", ctypes.c_int64),
        ("refine_flops", ctypes.c_int64),
        ("load_us", c_f32), ("loop_one_us", c_f32), ("loop_two_us", c_f32), ("labels_us", c_f32),
        ("loop_three_us", c_f32), ("rescale_us", c_f32), ("ident_us", c_f32),
        ("finish_us", c_f32), ("ident_prep_us", c_f32), ("ident_select_us", c_f32),
        ("ident_refine_us", c_f32), ("ident_post_us", c_f32),
        ("staged_bytes", ctypes.c_int64), ("staged_two_step", c_i32),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class BatchStats(ctypes.Structure):
    """cctagx_batch_stats."""
    _fields_ = [
        ("front_end_us", c_f32), ("stage_us", c_f32), ("candidates_us", c_f32),
        ("ident_us", c_f32), ("select_us", c_f32), ("refine_us", c_f32), ("tune_us", c_f32),
        ("filter_variant", c_i32), ("edge_blocks", c_i32), ("graph", c_i32),
        ("n_markers", c_i32),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class RefineMarker(ctypes.Structure):
    """cctagx_refine_marker."""
    _fields_ = [
        ("ellipse", c_f32 * 5), ("center0", c_f32 * 2), ("n_cuts", c_i32),
        ("stops", (c_f32 * 2) * 32),
    ]


def _arr(p, n, dtype):
    if n == 0 or not p:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(p, shape=(n,)).astype(dtype, copy=True)


def level_view_to_dict(v: LevelView) -> dict:
    """Copies a level view into numpy arrays."""
    w, h, n = v.width, v.height, v.n_edges
    vi = _arr(v.voters_index, n + 1, np.int32)
    nv = int(vi[-1]) if n + 1 > 0 and len(vi) else 0
    return {
        "width": w, "height": h, "shape": (v.shape_w, v.shape_h),
        "img": _arr(v.img, w * h, np.uint8).reshape(h, w),
        "dx": _arr(v.dx, w * h, np.int16).reshape(h, w),
        "dy": _arr(v.dy, w * h, np.int16).reshape(h, w),
        "edges": _arr(v.edges, w * h, np.uint8).reshape(h, w),
        "edge_map": _arr(v.edge_map, w * h, np.int32).reshape(h, w),
        "n_edges": n,
        "edge_x": _arr(v.edge_x, n, np.int16), "edge_y": _arr(v.edge_y, n, np.int16),
        "edge_dx": _arr(v.edge_dx, n, np.int16), "edge_dy": _arr(v.edge_dy, n, np.int16),
        "before": _arr(v.before, n, np.int32), "after": _arr(v.after, n, np.int32),
        "chosen": _arr(v.chosen, n, np.int32), "total_dist": _arr(v.total_dist, n, np.float32),
        "voters_index": vi, "voters_list": _arr(v.voters_list, nv, np.int32),
        "flow_length": _arr(v.flow_length, n, np.float32), "is_max": _arr(v.is_max, n, np.int32),
        "seeds": _arr(v.seeds, v.n_seeds, np.int32),
    }


class CCTagError(RuntimeError):
    def __init__(self, rc: int, what: str, detail: str = ""):
        msg = "%s failed: %s (%d)" % (what, ERRORS.get(rc, "error"), rc)
        if detail:
            msg += ": " + detail
        super().__init__(msg)
        self.rc = rc


_lib = None


def load() -> ctypes.CDLL:
    """Loads the in-tree library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(
            "libcctag_amd.so is not built (%s); run python -m cctag_amd.build" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    lib.cctagx_params_default.argtypes = [P(Params), ctypes.c_int]
    lib.cctagx_params_default.restype = None
    lib.cctagx_create.argtypes = [P(Params), ctypes.c_int, ctypes.c_int, P(vp)]
    lib.cctagx_destroy.argtypes = [vp]
    lib.cctagx_front_end.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_size_t]
    lib.cctagx_set_bank.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int]
    lib.cctagx_identify.argtypes = [vp, P(IdentMarker), ctypes.c_int, P(IdentResult)]
    lib.cctagx_last_markers.argtypes = [vp, P(Marker), ctypes.c_int, P(ctypes.c_int)]
    lib.cctagx_get_level.argtypes = [vp, ctypes.c_int, P(LevelView)]
    lib.cctagx_set_stage_timing.argtypes = [vp, ctypes.c_int]
    lib.cctagx_stage_count.argtypes = [vp]
    lib.cctagx_stage_time.argtypes = [vp, ctypes.c_int, P(ctypes.c_char_p), P(c_f32)]
    lib.cctagx_time_front_end.argtypes = [vp, ctypes.c_int, P(c_f32)]
    lib.cctagx_detect.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_size_t,
                                  ctypes.c_uint64, P(Marker), ctypes.c_int, P(ctypes.c_int)]
    lib.cctagx_load.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_size_t]
    lib.cctagx_load_bgr.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_size_t,
                                    ctypes.c_int]
    lib.cctagx_detect_bgr.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_size_t,
                                      ctypes.c_int, ctypes.c_uint64, P(Marker), ctypes.c_int,
                                      P(ctypes.c_int)]
    lib.cctagx_detect_loaded.argtypes = [vp, ctypes.c_uint64, P(Marker), ctypes.c_int,
                                         P(ctypes.c_int)]
    lib.cctagx_detect_times.argtypes = [vp, P(c_f32), P(c_f32), P(c_f32)]
    lib.cctagx_get_detect_stats.argtypes = [vp, P(DetectStats)]
    lib.cctagx_set_host_threads.argtypes = [vp, ctypes.c_int]
    lib.cctagx_set_refinement.argtypes = [vp, ctypes.c_int]
    lib.cctagx_batch_set_refinement.argtypes = [vp, ctypes.c_int]
    lib.cctagx_get_host_threads.argtypes = [vp]
    lib.cctagx_batch_create.argtypes = [P(Params), ctypes.c_int, P(c_i32), P(c_i32), P(vp)]
    lib.cctagx_batch_destroy.argtypes = [vp]
    lib.cctagx_batch_load.argtypes = [vp, ctypes.c_int, vp, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_size_t]
    lib.cctagx_batch_set_partition.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    lib.cctagx_batch_detect.argtypes = [vp, P(Marker), ctypes.c_int, P(c_i32), P(ctypes.c_int)]
    lib.cctagx_batch_get_stats.argtypes = [vp, P(BatchStats)]
    lib.cctagx_refiner_create.argtypes = [P(Params), ctypes.c_int, P(vp)]
    lib.cctagx_refiner_destroy.argtypes = [vp]
    lib.cctagx_refiner_set_image.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_size_t]
    lib.cctagx_refiner_set_markers.argtypes = [vp, P(RefineMarker), ctypes.c_int]
    lib.cctagx_refiner_cost_grad.argtypes = [vp, vp, ctypes.c_int, vp, vp]
    lib.cctagx_refiner_homography.argtypes = [vp, vp, vp]
    lib.cctagx_refiner_descend.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp, vp]
    lib.cctagx_refiner_refine.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp, vp]
    lib.cctagx_refiner_grid.argtypes = [vp, vp, vp]
    lib.cctagx_refiner_times.argtypes = [vp, P(c_f32), P(c_f32)]
    lib.cctagx_refine_grid_static.argtypes = [P(Params), vp, ctypes.c_int, ctypes.c_int,
                                              ctypes.c_size_t, P(RefineMarker), ctypes.c_int, vp,
                                              vp]
    lib.cctagx_params_parse_xml.argtypes = [ctypes.c_char_p, P(Params)]
    lib.cctagx_params_read_xml.argtypes = [ctypes.c_char_p, P(Params)]
    lib.cctagx_params_write_xml.argtypes = [P(Params), ctypes.c_char_p, ctypes.c_size_t,
                                            P(ctypes.c_size_t)]
    lib.cctagx_bank_parse.argtypes = [ctypes.c_char_p, vp, ctypes.c_int, ctypes.c_int,
                                      P(ctypes.c_int), P(ctypes.c_int)]
    lib.cctagx_bank_read.argtypes = [ctypes.c_char_p, vp, ctypes.c_int, ctypes.c_int,
                                     P(ctypes.c_int), P(ctypes.c_int)]
    lib.cctagx_version.restype = ctypes.c_char_p
    lib.cctagx_last_error.restype = ctypes.c_char_p
    lib.cctagx_device_count.restype = ctypes.c_int
    _lib = lib
    return lib


def check(rc: int, what: str) -> None:
    if rc != OK:
        detail = load().cctagx_last_error().decode(errors="replace")
        raise CCTagError(rc, what, detail)


def default_params(n_crowns: int = 3) -> Params:
    p = Params()
    load().cctagx_params_default(ctypes.byref(p), n_crowns)
    return p


def params_from_xml(text: str | None = None, path: str | None = None) -> Params:
    """cctag::Parameters from its XML archive (the `parameterFile` of
    cctagDetection): the archive text or a file path."""
    p = Params()
    lib = load()
    if path is not None:
        check(lib.cctagx_params_read_xml(path.encode(), ctypes.byref(p)), "cctagx_params_read_xml")
    else:
        check(lib.cctagx_params_parse_xml(text.encode(), ctypes.byref(p)), "cctagx_params_parse_xml")
    return p


def params_to_xml(p: Params) -> str:
    """The XML archive text of p (Parameters::serialize's members)."""
    lib = load()
    n = ctypes.c_size_t()
    check(lib.cctagx_params_write_xml(ctypes.byref(p), None, 0, ctypes.byref(n)),
          "cctagx_params_write_xml")
    buf = ctypes.create_string_buffer(n.value + 1)
    check(lib.cctagx_params_write_xml(ctypes.byref(p), buf, n.value + 1, ctypes.byref(n)),
          "cctagx_params_write_xml")
    return buf.value.decode()


def bank_from_text(text: str | None = None, path: str | None = None) -> np.ndarray:
    """A marker bank (rows of radius ratios) from its text format
    (CCTagMarkersBank::read): the text or a file path."""
    lib = load()
    rows, cols = ctypes.c_int(), ctypes.c_int()
    if path is not None:
        fn, arg, what = lib.cctagx_bank_read, path.encode(), "cctagx_bank_read"
    else:
        fn, arg, what = lib.cctagx_bank_parse, text.encode(), "cctagx_bank_parse"
    check(fn(arg, None, 0, 0, ctypes.byref(rows), ctypes.byref(cols)), what)
    out = np.zeros((rows.value, cols.value), np.float32)
    check(fn(arg, out.ctypes.data, rows.value, cols.value, ctypes.byref(rows), ctypes.byref(cols)),
          what)
    return out
