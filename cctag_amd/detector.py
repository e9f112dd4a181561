"""Python host mirror of the reference's detection entry points.

`Detector` owns one library handle (one frame size), the counterpart of a
reference `TagPipe` selected by pipeId (Detection.cpp:732-760).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib


class Detector:
    def __init__(self, width: int, height: int, params: _lib.Params | None = None,
                 n_crowns: int = 3):
        self._lib = _lib.load()
        self.params = params if params is not None else _lib.default_params(n_crowns)
        self.width, self.height = int(width), int(height)
        h = ctypes.c_void_p()
        _lib.check(self._lib.cctagx_create(ctypes.byref(self.params), self.width, self.height,
                                           ctypes.byref(h)), "cctagx_create")
        self._h = h

    def close(self) -> None:
        if self._h:
            self._lib.cctagx_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _gray(img: np.ndarray) -> np.ndarray:
        img = np.ascontiguousarray(img)
        if img.dtype != np.uint8 or img.ndim != 2:
            raise ValueError("expected a 2-D uint8 gray frame")  # Detection.cpp:815-817
        return img

    def front_end(self, img: np.ndarray) -> None:
        img = self._gray(img)
        h, w = img.shape
        _lib.check(self._lib.cctagx_front_end(self._h, img.ctypes.data, w, h, img.strides[0]),
                   "cctagx_front_end")

    def _markers(self, buf, n: int, max_markers: int) -> list:
        if n > max_markers:  # every marker, as the reference returns them all
            buf = (_lib.Marker * n)()
            got = ctypes.c_int()
            _lib.check(self._lib.cctagx_last_markers(self._h, buf, n, ctypes.byref(got)),
                       "cctagx_last_markers")
        return [buf[i].as_dict() for i in range(n)]

    def detect(self, img: np.ndarray, frame: int = 0, max_markers: int = 1024) -> list:
        """cctagDetection (ICCTag.hpp:124-131): every marker found, sorted by
        id; only status == 1 markers are identified.  A frame of another size
        than the handle's is rejected (the reference stops the process,
        Detection.cpp:753-757)."""
        img = self._gray(img)
        h, w = img.shape
        buf = (_lib.Marker * max_markers)()
        n = ctypes.c_int()
        _lib.check(self._lib.cctagx_detect(self._h, img.ctypes.data, w, h, img.strides[0], frame,
                                           buf, max_markers, ctypes.byref(n)), "cctagx_detect")
        return self._markers(buf, n.value, max_markers)

    def load(self, img: np.ndarray) -> None:
        """TagPipe::load (cuda/tag.cu:154-159): the frame stays resident."""
        img = self._gray(img)
        h, w = img.shape
        _lib.check(self._lib.cctagx_load(self._h, img.ctypes.data, w, h, img.strides[0]),
                   "cctagx_load")

    @staticmethod
    def _color(img: np.ndarray) -> np.ndarray:
        img = np.asarray(img)
        if img.dtype != np.uint8 or img.ndim != 3 or img.shape[2] not in (3, 4):
            raise ValueError("expected an h x w x 3 (BGR) or h x w x 4 (BGRA) uint8 frame")
        if img.strides[1:] != (img.shape[2], 1):
            img = np.ascontiguousarray(img)
        return img

    def detect_bgr(self, img: np.ndarray, frame: int = 0, max_markers: int = 1024) -> list:
        """detect() of a BGR / BGRA frame, converted to gray on the device as
        the reference's applications convert it (cv::cvtColor BGR2GRAY)."""
        img = self._color(img)
        h, w, c = img.shape
        buf = (_lib.Marker * max_markers)()
        n = ctypes.c_int()
        _lib.check(self._lib.cctagx_detect_bgr(self._h, img.ctypes.data, w, h, img.strides[0], c,
                                               frame, buf, max_markers, ctypes.byref(n)),
                   "cctagx_detect_bgr")
        return self._markers(buf, n.value, max_markers)

    def load_bgr(self, img: np.ndarray) -> None:
        """load() of a BGR / BGRA frame (converted on the device)."""
        img = self._color(img)
        h, w, c = img.shape
        _lib.check(self._lib.cctagx_load_bgr(self._h, img.ctypes.data, w, h, img.strides[0], c),
                   "cctagx_load_bgr")

    def detect_loaded(self, frame: int = 0, max_markers: int = 1024) -> list:
        """Detection on the resident frame (no host-to-device copy)."""
        buf = (_lib.Marker * max_markers)()
        n = ctypes.c_int()
        _lib.check(self._lib.cctagx_detect_loaded(self._h, frame, buf, max_markers,
                                                  ctypes.byref(n)), "cctagx_detect_loaded")
        return self._markers(buf, n.value, max_markers)

    def set_bank(self, ratios) -> None:
        """The marker bank identification reads (rows of radius ratios,
        innermost ring first), as cctagDetection's pBank supplies it."""
        r = np.ascontiguousarray(np.asarray(ratios, np.float32))
        if r.ndim != 2:
            raise ValueError("a bank is a 2-D table of radius ratios")
        _lib.check(self._lib.cctagx_set_bank(self._h, r.ctypes.data, r.shape[0], r.shape[1]),
                   "cctagx_set_bank")

    def identify(self, markers) -> list:
        """The identification of given markers on the frame of the last front
        end (cctagx_identify): dicts with "ellipse", "center0", "points" and
        an optional "scale"; returns per marker status, id, centre, quality,
        homography and the selected cut stops."""
        arr, keep = _lib.ident_markers(markers)
        out = (_lib.IdentResult * max(len(markers), 1))()
        _lib.check(self._lib.cctagx_identify(self._h, arr, len(markers), out), "cctagx_identify")
        del keep
        return [out[i].as_dict() for i in range(len(markers))]

    def detect_times(self) -> dict:
        host, sel, ref = ctypes.c_float(), ctypes.c_float(), ctypes.c_float()
        _lib.check(self._lib.cctagx_detect_times(self._h, ctypes.byref(host), ctypes.byref(sel),
                                                 ctypes.byref(ref)), "cctagx_detect_times")
        return {"host_us": host.value, "select_us": sel.value, "refine_us": ref.value}

    def set_refinement(self, gradient_steps: int = 0) -> None:
        """0: the reference grid search; n > 0: its first iteration, then n
        descent steps on the differentiable score (cctagx_set_refinement)."""
        _lib.check(self._lib.cctagx_set_refinement(self._h, gradient_steps), "cctagx_set_refinement")

    def set_host_threads(self, threads: int) -> None:
        """Candidate-stage host threads (0 = default); results do not change."""
        _lib.check(self._lib.cctagx_set_host_threads(self._h, threads), "cctagx_set_host_threads")

    def host_threads(self) -> int:
        return self._lib.cctagx_get_host_threads(self._h)

    def detect_stats(self) -> dict:
        """Time and work of the last detection (cctagx_get_detect_stats)."""
        st = _lib.DetectStats()
        _lib.check(self._lib.cctagx_get_detect_stats(self._h, ctypes.byref(st)),
                   "cctagx_get_detect_stats")
        return st.as_dict()

    def level(self, level: int) -> dict:
        v = _lib.LevelView()
        _lib.check(self._lib.cctagx_get_level(self._h, level, ctypes.byref(v)),
                   "cctagx_get_level")
        return _lib.level_view_to_dict(v)

    def set_stage_timing(self, per_kernel: bool) -> None:
        """Event marks after every front-end kernel (profiling passes) or only
        around the whole front end (the default)."""
        _lib.check(self._lib.cctagx_set_stage_timing(self._h, int(per_kernel)),
                   "cctagx_set_stage_timing")

    def stage_times(self) -> list:
        out = []
        n = self._lib.cctagx_stage_count(self._h)
        for i in range(max(n, 0)):
            name = ctypes.c_char_p()
            us = ctypes.c_float()
            if self._lib.cctagx_stage_time(self._h, i, ctypes.byref(name), ctypes.byref(us)) == 0:
                out.append((name.value.decode(), us.value))
        return out

    def time_front_end(self, iters: int) -> float:
        us = ctypes.c_float()
        _lib.check(self._lib.cctagx_time_front_end(self._h, iters, ctypes.byref(us)),
                   "cctagx_time_front_end")
        return us.value


class BatchDetector:
    """Batched detection of many frames of uneven sizes (cctagx_batch_*):
    one captured graph launch for the front end of every frame, the
    candidate stage parallel over frames, identification batched."""

    def __init__(self, sizes, params: _lib.Params | None = None, n_crowns: int = 3):
        self._lib = _lib.load()
        self.params = params if params is not None else _lib.default_params(n_crowns)
        self.sizes = [(int(w), int(h)) for w, h in sizes]
        n = len(self.sizes)
        ws = (_lib.c_i32 * n)(*[w for w, _ in self.sizes])
        hs = (_lib.c_i32 * n)(*[h for _, h in self.sizes])
        h = ctypes.c_void_p()
        _lib.check(self._lib.cctagx_batch_create(ctypes.byref(self.params), n, ws, hs,
                                                 ctypes.byref(h)), "cctagx_batch_create")
        self._h = h

    def close(self) -> None:
        if self._h:
            self._lib.cctagx_batch_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def load(self, frame: int, img: np.ndarray) -> None:
        img = np.ascontiguousarray(img)
        w, h = self.sizes[frame]
        if img.dtype != np.uint8 or img.shape != (h, w):
            raise ValueError("frame %d must be a %dx%d uint8 image" % (frame, w, h))
        _lib.check(self._lib.cctagx_batch_load(self._h, frame, img.ctypes.data, w, h,
                                               img.strides[0]), "cctagx_batch_load")

    def set_refinement(self, gradient_steps: int = 0) -> None:
        """As Detector.set_refinement, for every frame of the batch."""
        _lib.check(self._lib.cctagx_batch_set_refinement(self._h, gradient_steps),
                   "cctagx_batch_set_refinement")

    def set_partition(self, filter_variant: int = -1, edge_blocks: int = 0) -> None:
        _lib.check(self._lib.cctagx_batch_set_partition(self._h, filter_variant, edge_blocks),
                   "cctagx_batch_set_partition")

    def detect(self, max_markers: int = 1 << 14) -> list:
        """Markers of every frame: a list (one per frame) of marker dicts."""
        n = len(self.sizes)
        buf = (_lib.Marker * max_markers)()
        first = (_lib.c_i32 * (n + 1))()
        tot = ctypes.c_int()
        _lib.check(self._lib.cctagx_batch_detect(self._h, buf, max_markers, first,
                                                 ctypes.byref(tot)), "cctagx_batch_detect")
        if tot.value > max_markers:
            raise RuntimeError("more than %d markers in the batch" % max_markers)
        return [[buf[i].as_dict() for i in range(first[f], first[f + 1])] for f in range(n)]

    def stats(self) -> dict:
        st = _lib.BatchStats()
        _lib.check(self._lib.cctagx_batch_get_stats(self._h, ctypes.byref(st)),
                   "cctagx_batch_get_stats")
        return st.as_dict()


class Refiner:
    """Differentiable centre refinement (cctagx_refiner_*): the grid search's
    score as a function of the homography, its gradient from a forward /
    adjoint kernel pair, a batched descent on the imaged centres and the
    reference grid search on the same markers."""

    def __init__(self, max_markers: int, params: _lib.Params | None = None, n_crowns: int = 3):
        self._lib = _lib.load()
        self.params = params if params is not None else _lib.default_params(n_crowns)
        h = ctypes.c_void_p()
        _lib.check(self._lib.cctagx_refiner_create(ctypes.byref(self.params), int(max_markers),
                                                   ctypes.byref(h)), "cctagx_refiner_create")
        self._h = h
        self.n = 0

    def close(self) -> None:
        if self._h:
            self._lib.cctagx_refiner_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_image(self, img: np.ndarray) -> None:
        img = np.ascontiguousarray(img, dtype=np.uint8)
        h, w = img.shape
        _lib.check(self._lib.cctagx_refiner_set_image(self._h, img.ctypes.data, w, h, img.strides[0]),
                   "cctagx_refiner_set_image")

    def set_markers(self, markers) -> None:
        """markers: dicts with ellipse (cx, cy, a, b, angle), center0 (x, y), stops (k x 2)."""
        arr = (_lib.RefineMarker * max(len(markers), 1))()
        for i, m in enumerate(markers):
            arr[i].ellipse[:] = [float(v) for v in m["ellipse"]]
            arr[i].center0[:] = [float(v) for v in m["center0"]]
            st = np.asarray(m["stops"], np.float32)
            arr[i].n_cuts = len(st)
            for k, (x, y) in enumerate(st):
                arr[i].stops[k][0] = float(x)
                arr[i].stops[k][1] = float(y)
        _lib.check(self._lib.cctagx_refiner_set_markers(self._h, arr, len(markers)),
                   "cctagx_refiner_set_markers")
        self.n = len(markers)

    def cost_grad(self, H: np.ndarray, fp64: bool = False):
        H = np.ascontiguousarray(np.asarray(H, np.float32).reshape(self.n, 9))
        cost = np.zeros(self.n, np.float64)
        grad = np.zeros((self.n, 10), np.float64)
        _lib.check(self._lib.cctagx_refiner_cost_grad(self._h, H.ctypes.data, int(fp64),
                                                      cost.ctypes.data, grad.ctypes.data),
                   "cctagx_refiner_cost_grad")
        return cost, grad

    def homography(self, centers: np.ndarray) -> np.ndarray:
        c = np.ascontiguousarray(np.asarray(centers, np.float32).reshape(self.n, 2))
        H = np.zeros((self.n, 9), np.float32)
        _lib.check(self._lib.cctagx_refiner_homography(self._h, c.ctypes.data, H.ctypes.data),
                   "cctagx_refiner_homography")
        return H.reshape(self.n, 3, 3)

    def refine(self, steps: int = 20, fp64: bool = False):
        """The grid search's first iteration, then `steps` descent steps."""
        c = np.zeros((self.n, 2), np.float32)
        cost = np.zeros(self.n, np.float64)
        _lib.check(self._lib.cctagx_refiner_refine(self._h, steps, int(fp64), c.ctypes.data,
                                                   cost.ctypes.data), "cctagx_refiner_refine")
        return c, cost

    def descend(self, steps: int = 20, fp64: bool = False):
        c = np.zeros((self.n, 2), np.float32)
        cost = np.zeros(self.n, np.float64)
        _lib.check(self._lib.cctagx_refiner_descend(self._h, steps, int(fp64), c.ctypes.data,
                                                    cost.ctypes.data), "cctagx_refiner_descend")
        return c, cost

    def grid(self):
        c = np.zeros((self.n, 2), np.float32)
        ok = np.zeros(self.n, np.int32)
        _lib.check(self._lib.cctagx_refiner_grid(self._h, c.ctypes.data, ok.ctypes.data),
                   "cctagx_refiner_grid")
        return c, ok.astype(bool)

    def times(self) -> dict:
        d, g = ctypes.c_float(), ctypes.c_float()
        _lib.check(self._lib.cctagx_refiner_times(self._h, ctypes.byref(d), ctypes.byref(g)),
                   "cctagx_refiner_times")
        return {"descend_us": d.value, "grid_us": g.value}
