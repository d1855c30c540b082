"""ctypes bindings of the MI355X DIC engine.

`Engine` mirrors the reference's Python binding `semper.Engine`
(bindings/python/semper/__init__.py) over the C ABI of include/semper_c.h:
set_reference(), run() -> (points[n, 8], metrics[24]), cancel(). `Dev` wraps the
kernel-level interface of include/semper_dev.h used by the tests and the
benchmark. The libraries are the in-tree builds next to this file; loading
fails loudly when they are missing.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_C = os.path.join(HERE, "libsemper_c.so")
LIB_DEV = os.path.join(HERE, "libsemper_dev.so")
# the learned seeder's trained weights (train_seeder.py)
SEEDER_WEIGHTS = os.path.join(HERE, "seeder_weights.bin")

FLOATS_PER_POINT = 8
METRICS_LEN = 24
ERR_ROI, ERR_INIT, ERR_STRAIN_WINDOW, ERR_SEED, ERR_CANCELLED = -2, -3, -4, -5, -99
EXPORTS = [
    "semper_cancel", "semper_create", "semper_destroy", "semper_incremental_cancel",
    "semper_incremental_create", "semper_incremental_destroy", "semper_incremental_init",
    "semper_incremental_process", "semper_run", "semper_set_reference", "semper_version",
]


class Params(ctypes.Structure):
    """semper_params: 8 x int32 (reference include/semper/semper_c.h:46-52)."""
    _fields_ = [(n, ctypes.c_int32) for n in (
        "rect_x", "rect_y", "rect_w", "rect_h", "step", "subset_size", "strain_window",
        "use_6x6_interpolator")]


class IncInfo(ctypes.Structure):
    """semper_incremental_info: 6 x int32 (reference semper_c.h:155-162)."""
    _fields_ = [(n, ctypes.c_int32) for n in (
        "reference_frame", "reference_updated", "tracked_points", "frame0_points",
        "taylor_fallbacks", "lost_this_frame")]


PROGRESS_CB = ctypes.CFUNCTYPE(None, ctypes.c_int, ctypes.c_void_p)
_VP = ctypes.c_void_p


def _ptr(a):
    return a.ctypes.data_as(_VP) if a is not None else None


def _bind_c(lib):
    lib.semper_create.restype = _VP
    lib.semper_destroy.argtypes = [_VP]
    lib.semper_set_reference.argtypes = [_VP, _VP, ctypes.c_size_t, _VP, ctypes.c_size_t, ctypes.c_int,
                                         ctypes.c_int]
    lib.semper_run.argtypes = [_VP, _VP, ctypes.c_size_t, _VP, ctypes.c_size_t, ctypes.POINTER(Params), _VP,
                               ctypes.c_int, _VP, ctypes.c_int, PROGRESS_CB, _VP]
    lib.semper_cancel.argtypes = [_VP]
    lib.semper_version.restype = ctypes.c_char_p
    lib.semper_incremental_create.restype = _VP
    lib.semper_incremental_destroy.argtypes = [_VP]
    lib.semper_incremental_init.argtypes = [_VP, _VP, ctypes.c_size_t, _VP, ctypes.c_size_t, ctypes.c_int,
                                            ctypes.c_int, ctypes.POINTER(Params), ctypes.c_int32]
    lib.semper_incremental_process.argtypes = [_VP, _VP, ctypes.c_size_t, _VP, ctypes.c_int, _VP, ctypes.c_int,
                                               ctypes.POINTER(IncInfo), PROGRESS_CB, _VP]
    lib.semper_incremental_cancel.argtypes = [_VP]
    return lib


def load_c(path=LIB_C):
    if not os.path.exists(path):
        raise RuntimeError(f"libsemper_c.so is not built ({path}); run __graft_entry__.build()")
    return _bind_c(ctypes.CDLL(path))


def load_dev(path=LIB_DEV):
    if not os.path.exists(path):
        raise RuntimeError(f"libsemper_dev.so is not built ({path}); run __graft_entry__.build()")
    lib = _bind_c(ctypes.CDLL(path))
    lib.sdk_device_count.restype = ctypes.c_int
    lib.sdk_malloc.restype = _VP
    lib.sdk_malloc.argtypes = [ctypes.c_size_t]
    lib.sdk_free.argtypes = [_VP]
    lib.sdk_host_alloc.restype = _VP
    lib.sdk_host_alloc.argtypes = [ctypes.c_size_t]
    lib.sdk_host_free.argtypes = [_VP]
    lib.sdk_memcpy.argtypes = [_VP, _VP, ctypes.c_size_t, ctypes.c_int]
    lib.sdk_solve_nodes.argtypes = [_VP, _VP, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_float, ctypes.c_int, _VP, _VP, _VP, _VP, _VP]
    lib.sdk_solve_nodes_noalloc.argtypes = [_VP, _VP, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_int, _VP, _VP, _VP,
                                            _VP, _VP, _VP]
    lib.sdk_engine_global_model.argtypes = [_VP, _VP]
    lib.sdk_pool.argtypes = [_VP, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _VP, _VP,
                             _VP, _VP]
    lib.sdk_ref_planes.argtypes = [_VP, _VP, ctypes.c_int, ctypes.c_int, _VP, _VP, _VP]
    lib.sdk_engine_create.restype = _VP
    lib.sdk_engine_destroy.argtypes = [_VP]
    lib.sdk_engine_set_reference.argtypes = [_VP, _VP, ctypes.c_int, ctypes.c_int, ctypes.c_int, _VP]
    lib.sdk_engine_run.argtypes = [_VP, _VP, ctypes.c_int, _VP, _VP, ctypes.c_int, _VP, ctypes.c_int,
                                   ctypes.c_float]
    lib.sdk_engine_timings.argtypes = [_VP, _VP]
    lib.sdk_engine_launch_log.argtypes = [_VP, _VP, ctypes.c_int]
    lib.sdk_engine_set_flag.argtypes = [_VP, ctypes.c_int, ctypes.c_int]
    lib.sdk_wave_sum_probe.argtypes = [_VP, ctypes.c_int, _VP, _VP]
    lib.sdk_interp_probe.argtypes = [_VP, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _VP, _VP, _VP]
    lib.sdk_compose64_probe.argtypes = ([_VP, _VP] + [ctypes.c_int] * 5 + [_VP] * 5 + [ctypes.c_int] * 5 +
                                        [_VP] * 6)
    lib.sdk_synth_frame.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_int, _VP,
                                    ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_float, _VP, ctypes.c_int]
    lib.sdk_synth_pair.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_int, _VP, ctypes.c_int,
                                   ctypes.c_double, ctypes.c_double, ctypes.c_float, _VP, _VP, ctypes.c_int]
    lib.sdk_decode.argtypes = [_VP, ctypes.c_size_t, ctypes.c_int, ctypes.c_int, ctypes.c_int, _VP, ctypes.c_size_t,
                               _VP, _VP]
    lib.sdk_image_dimensions.argtypes = [_VP, ctypes.c_size_t, _VP, _VP]
    lib.sdk_image_dimensions.restype = None
    lib.sdk_to_gray8.argtypes = [_VP, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _VP]
    lib.sdk_to_mask8.argtypes = [_VP, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _VP]
    lib.sdk_seeder_grad.argtypes = [_VP, _VP, _VP, ctypes.c_int, _VP, _VP]
    lib.sdk_seeder_forward.argtypes = [_VP, _VP, ctypes.c_int, _VP]
    lib.sdk_seeder_forward_mfma.argtypes = [_VP, _VP, ctypes.c_int, _VP]
    lib.sdk_seeder_train.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_char_p, _VP,
                                     ctypes.c_int]
    lib.sdk_engine_load_seeder.argtypes = [_VP, ctypes.c_char_p]
    lib.sdk_engine_seed_stats.argtypes = [_VP, _VP]
    lib.sdk_engine_set_subset_map.argtypes = [_VP, ctypes.c_int, ctypes.c_int, _VP]
    lib.sdk_engine_tuning.argtypes = [_VP, _VP, _VP]
    lib.sdk_engine_set_region_map.argtypes = [_VP, ctypes.c_int, ctypes.c_int, _VP]
    lib.sdk_incremental_set_region_map.argtypes = [_VP, ctypes.c_int, ctypes.c_int, _VP]
    lib.sdk_engine_graph_stats.argtypes = [_VP, _VP]
    lib.sdk_engine_node_report.argtypes = [_VP, _VP, _VP, ctypes.c_int]
    lib.sdk_engine_node_planes.argtypes = [_VP, ctypes.c_int, _VP, ctypes.c_int]
    lib.sdk_incremental_set_flag.argtypes = [_VP, ctypes.c_int, ctypes.c_int]
    lib.sdk_incremental_graph_stats.argtypes = [_VP, _VP]
    lib.sdk_incremental_timings.argtypes = [_VP, _VP]
    lib.sdk_incremental_engine_timings.argtypes = [_VP, _VP]
    lib.sdk_incremental_anchor_stats.argtypes = [_VP, _VP]
    return lib


_DEPTHS = {np.dtype(np.uint8): 0, np.dtype(np.uint16): 1, np.dtype(np.int16): 2, np.dtype(np.int32): 3,
           np.dtype(np.float32): 4, np.dtype(np.float64): 5}


def decode(data, expected_w=0, expected_h=0, mask=False, lib=None):
    """The boundary's decode of image bytes (io::decode_gray, or decode_mask
    with mask=True): a 2-D uint8 array, or None when nothing decodes."""
    lib = lib or load_dev()
    buf = np.frombuffer(bytes(data), np.uint8) if not isinstance(data, np.ndarray) else data
    w, h = ctypes.c_int32(0), ctypes.c_int32(0)
    if lib.sdk_decode(_ptr(buf), buf.size, expected_w, expected_h, int(mask), None, 0, ctypes.byref(w),
                      ctypes.byref(h)) != 0:
        return None
    out = np.empty((h.value, w.value), np.uint8)
    rc = lib.sdk_decode(_ptr(buf), buf.size, expected_w, expected_h, int(mask), _ptr(out), out.size,
                        ctypes.byref(w), ctypes.byref(h))
    return out if rc == 0 else None


def image_dimensions(data, lib=None):
    lib = lib or load_dev()
    buf = np.frombuffer(bytes(data), np.uint8) if data is not None else None
    w, h = ctypes.c_int32(-1), ctypes.c_int32(-1)
    lib.sdk_image_dimensions(_ptr(buf), 0 if buf is None else buf.size, ctypes.byref(w), ctypes.byref(h))
    return w.value, h.value


def _convert(fn, img):
    a = np.ascontiguousarray(img)
    if a.dtype not in _DEPTHS or a.ndim not in (2, 3) or (a.ndim == 3 and a.shape[2] != 3):
        raise ValueError("expected an (h, w) or (h, w, 3) array of u8/u16/s16/s32/f32/f64")
    out = np.empty(a.shape[:2], np.uint8)
    if fn(_ptr(a), a.shape[1], a.shape[0], 1 if a.ndim == 2 else 3, _DEPTHS[a.dtype], _ptr(out)) != 0:
        return None
    return out


def to_gray8(img, lib=None):
    """io::to_gray8 of an (h, w) or (h, w, 3 = R, G, B) array."""
    return _convert((lib or load_dev()).sdk_to_gray8, img)


def to_mask8(img, lib=None):
    """io::to_mask8 of an (h, w) or (h, w, 3 = R, G, B) array."""
    return _convert((lib or load_dev()).sdk_to_mask8, img)


def _image_buffer(img):
    """(contiguous uint8 buffer, (h, w) or None): an array is a gray plane,
    bytes are passed through for the engine to decode."""
    if isinstance(img, (bytes, bytearray, memoryview)):
        return np.frombuffer(img, np.uint8), None
    a = np.ascontiguousarray(img, dtype=np.uint8)
    return a, a.shape[:2]


def _mask_buffer(mask):
    """Mask bytes for the C ABI (reference bindings/python/semper/__init__.py:
    Engine docstring): bytes pass through; a bool or non-uint8 array, or a
    uint8 array of only 0 and 1, is boolean (nonzero included, as 255); any
    other uint8 array excludes the pixels below 128."""
    if mask is None:
        return None
    if isinstance(mask, (bytes, bytearray, memoryview)):
        return np.frombuffer(mask, np.uint8)
    m = np.asarray(mask)
    if m.dtype != np.uint8 or (m.size and int(m.max()) <= 1):
        m = np.where(m != 0, 255, 0).astype(np.uint8)
    return np.ascontiguousarray(m, dtype=np.uint8)


def params(rect, step, subset, strain_window, use6=False):
    p = Params()
    p.rect_x, p.rect_y, p.rect_w, p.rect_h = [int(v) for v in rect]
    p.step, p.subset_size, p.strain_window = int(step), int(subset), int(strain_window)
    p.use_6x6_interpolator = 1 if use6 else 0
    return p


class Engine:
    """One semper_engine handle (reference bindings/python/semper/__init__.py:83-230)."""

    def __init__(self, lib=None):
        self.lib = lib or load_c()
        self.h = self.lib.semper_create()
        if not self.h:
            raise MemoryError("semper_create failed")

    def close(self):
        if self.h:
            self.lib.semper_destroy(self.h)
            self.h = None

    __del__ = close

    def set_reference(self, img, mask=None, expected_w=0, expected_h=0):
        """Cache the reference: a 2-D uint8 array, or image bytes (raw gray or
        RGBA of expected_w x expected_h, or an encoded file)."""
        buf, shape = _image_buffer(img)
        if shape is not None:
            expected_h, expected_w = shape
        m = _mask_buffer(mask)
        rc = self.lib.semper_set_reference(self.h, _ptr(buf), buf.size, _ptr(m), 0 if m is None else m.size,
                                           int(expected_w), int(expected_h))
        self.shape = shape if shape is not None else None
        return rc

    def run(self, img, p, mask=None, capacity=None, metrics_len=METRICS_LEN, progress=None):
        buf, shape = _image_buffer(img)
        # default capacity: every grid node of the rectangle clipped to the image
        ih, iw = shape if shape is not None else (max(p.rect_h, 0), max(p.rect_w, 0))
        st = max(p.step, 1)
        rw, rh = min(max(p.rect_w, 0), iw), min(max(p.rect_h, 0), ih)
        cap = capacity if capacity is not None else max(1, (rw // st + 1) * (rh // st + 1)) * FLOATS_PER_POINT
        out = np.zeros(max(cap, 1), np.float32)
        met = np.zeros(max(metrics_len, 1), np.float32)
        m = _mask_buffer(mask)
        cb = PROGRESS_CB(progress) if progress else PROGRESS_CB()
        n = self.lib.semper_run(self.h, _ptr(buf), buf.size, _ptr(m), 0 if m is None else m.size,
                                ctypes.byref(p), _ptr(out), cap, _ptr(met), metrics_len, cb, None)
        pts = out[: max(n, 0) * FLOATS_PER_POINT].reshape(-1, FLOATS_PER_POINT).copy()
        return n, pts, met

    def cancel(self):
        self.lib.semper_cancel(self.h)


class Incremental:
    """One semper_incremental handle (reference include/semper/incremental.hpp,
    bindings/python/semper/__init__.py): init(), process() -> (points[n, 8],
    metrics[24], info)."""

    def __init__(self, lib=None):
        self.lib = lib or load_c()
        self.h = self.lib.semper_incremental_create()
        if not self.h:
            raise MemoryError("semper_incremental_create failed")

    def close(self):
        if self.h:
            self.lib.semper_incremental_destroy(self.h)
            self.h = None

    __del__ = close

    def init(self, img, p, update_interval, mask=None, expected_w=0, expected_h=0):
        buf, shape = _image_buffer(img)
        if shape is not None:
            expected_h, expected_w = shape
        h, w = (expected_h, expected_w) if shape is None else shape
        w, h = (w or p.rect_w), (h or p.rect_h)
        m = _mask_buffer(mask)
        self.cap = max(1, (min(p.rect_w, w) // max(p.step, 1) + 1) * (min(p.rect_h, h) // max(p.step, 1) + 1)) * 8
        return self.lib.semper_incremental_init(self.h, _ptr(buf), buf.size, _ptr(m), 0 if m is None else m.size,
                                                int(expected_w), int(expected_h), ctypes.byref(p), update_interval)

    def process(self, img, capacity=None, metrics_len=METRICS_LEN):
        img, _ = _image_buffer(img)
        cap = capacity if capacity is not None else self.cap
        out = np.zeros(max(cap, 1), np.float32)
        met = np.zeros(max(metrics_len, 1), np.float32)
        info = IncInfo()
        n = self.lib.semper_incremental_process(self.h, _ptr(img), img.size, _ptr(out), cap, _ptr(met), metrics_len,
                                                ctypes.byref(info), PROGRESS_CB(), None)
        pts = out[: max(n, 0) * FLOATS_PER_POINT].reshape(-1, FLOATS_PER_POINT).copy()
        return n, pts, met, info

    def cancel(self):
        self.lib.semper_incremental_cancel(self.h)
