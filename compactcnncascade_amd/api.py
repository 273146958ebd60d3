"""Python binding of libccnn.so (the C ABI in include/ccnn.h).

`FaceDetector` mirrors the reference's CompactCNNLib::FaceDetector
(prebuild/include/CompactCNNLib.h:36-194): same members, same argument meaning and
return codes, with numpy arrays in place of ImageData/Face buffers. Everything runs on
the GPU through the HIP library; if the library or a HIP device is missing, loading or
Init fails loudly -- there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libccnn.so")
# the profiling build with the kernels' phase timestamps (build.py --phase-stamps); load()
# takes it when LIB_PATH is set to it before the first load
PHASES_LIB_PATH = os.path.join(_HERE, "libccnn_phases.so")

_P = ctypes.c_void_p
_I = ctypes.c_int
_F = ctypes.c_float

PIPELINE_CPU, PIPELINE_GPU, PIPELINE_GPU_CPU = 0, 1, 2
DETECT_DISABLE, DETECT_SYNC, DETECT_ASYNC = 0, 1, 2
PRECISION_DEF, PRECISION_LOW, PRECISION_NORMAL, PRECISION_HIGH, PRECISION_ULTRA = 0, 1, 2, 3, 4
# compute precisions of this library (ccnn_set_compute_precision)
FP32, FP16_BF16, FP16_FP8, FP32_BF16, FP32_FP8 = 0, 1, 2, 3, 4
# stage-1 tile shapes by id (L1 tile, response tile), numbered as the library does:
# id = L1 shape index * 3 + response shape index
_L1_SHAPES = ((64, 16), (32, 16), (32, 32), (64, 8))
_RESP_SHAPES = ((16, 16), (32, 8), (16, 8))
TILE_SHAPES = ["L1 %dx%d, responses %dx%d" % (a + r) for a in _L1_SHAPES for r in _RESP_SHAPES]


class CParam(ctypes.Structure):
    _fields_ = [("max_image_width", _I), ("max_image_height", _I), ("min_face_height", _I),
                ("max_face_height", _I), ("scale_factor", _F), ("min_neighbors", _I),
                ("detect_precision", _I), ("treshold", _F * 3), ("drop_detect", _I), ("equalize", _I),
                ("reflection", _I), ("pipeline", _I), ("detect_mode", _I), ("num_threads", _I),
                ("models", ctypes.c_char_p * 3), ("index_output", _I * 3), ("device_info", _I),
                ("cuda_device_id", _I), ("cl_platform_id", _I), ("cl_device_id", _I),
                ("facial_analysis", _I)]


class CFace(ctypes.Structure):
    _fields_ = [("x", _I), ("y", _I), ("width", _I), ("height", _I), ("score", _F),
                ("gender", _I), ("smile", _I), ("glasses", _I)]


class CImage(ctypes.Structure):
    _fields_ = [("cols", _I), ("rows", _I), ("channels", _I), ("data", _P), ("step", ctypes.c_size_t)]


# ccnn_image as a numpy record (the same 32 bytes as CImage), for building many at once
CIMAGE_DTYPE = np.dtype({"names": ["cols", "rows", "channels", "data", "step"],
                         "formats": [np.int32, np.int32, np.int32, np.uint64, np.uint64],
                         "offsets": [0, 4, 8, 16, 24], "itemsize": 32})

FACE_DTYPE = np.dtype([("x", np.int32), ("y", np.int32), ("width", np.int32), ("height", np.int32),
                       ("score", np.float32), ("gender", np.int32), ("smile", np.int32), ("glasses", np.int32)])

_lib = None


def load():
    """Loads libccnn.so (built by compactcnncascade_amd.build); raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libccnn.so is not built: run `python -m compactcnncascade_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    L.ccnn_param_default.argtypes = [ctypes.POINTER(CParam)]
    L.ccnn_create.restype = _P
    L.ccnn_destroy.argtypes = [_P]
    L.ccnn_init.argtypes = [_P, ctypes.POINTER(CParam)]
    L.ccnn_detect.argtypes = [_P, _P, _I, ctypes.POINTER(CImage)]
    L.ccnn_detect_device.argtypes = [_P, _P, _I, ctypes.POINTER(CImage)]
    L.ccnn_detect_batch.argtypes = [_P, _I, _P, _I, _P, _I, _P]
    L.ccnn_phase_stamps.argtypes = [_I, _P, _I]
    L.ccnn_clear.argtypes = [_P]
    L.ccnn_is_empty.argtypes = [_P]
    L.ccnn_get_param.argtypes = [_P, ctypes.POINTER(CParam)]
    L.ccnn_set_param.argtypes = [_P, ctypes.POINTER(CParam)]
    L.ccnn_get_max_image_size.argtypes = [_P, ctypes.POINTER(_I), ctypes.POINTER(_I)]
    L.ccnn_set_max_image_size.argtypes = [_P, _I, _I]
    L.ccnn_get_min_object_height.argtypes = [_P]
    L.ccnn_set_min_object_height.argtypes = [_P, _I]
    L.ccnn_get_max_object_height.argtypes = [_P]
    L.ccnn_set_max_object_height.argtypes = [_P, _I]
    L.ccnn_get_num_threads.argtypes = [_P]
    L.ccnn_set_num_threads.argtypes = [_P, _I]
    L.ccnn_get_gray_image.argtypes = [_P, ctypes.POINTER(CImage), _I]
    L.ccnn_stage_gray.argtypes = [ctypes.POINTER(CImage), _P]
    L.ccnn_stage_level.argtypes = [_P, _I, _I, _I, _I, _I, _P]
    L.ccnn_stage1_size.argtypes = [_I, _I, ctypes.POINTER(_I), ctypes.POINTER(_I)]
    L.ccnn_stage1.argtypes = [_P, _I, _I, _P]
    L.ccnn_stage_check.argtypes = [ctypes.POINTER(CParam), _P, _I, _I, _I, _P, _P, _P, _P]
    L.ccnn_last_trace.argtypes = [_P, ctypes.POINTER(_I), _P, _I, _P, _P]
    L.ccnn_last_facial.argtypes = [_P, _P, _I]
    L.ccnn_last_responses.argtypes = [_P, _P, ctypes.c_size_t]
    L.ccnn_last_responses.restype = ctypes.c_long
    L.ccnn_stage23.argtypes = [_I, _P, _I, _P]
    L.ccnn_set_batch_mode.argtypes = [_P, _I]
    L.ccnn_batch_stats.argtypes = [_P, ctypes.POINTER(_I), ctypes.POINTER(_I)]
    L.ccnn_set_tile_shape.argtypes = [_P, _I]
    L.ccnn_set_compute_precision.argtypes = [_P, _I]
    L.ccnn_last_status.argtypes = [_P, _P, _I]
    L.ccnn_set_skipnet.argtypes = [_P, _P, _F, _I, _F]
    L.ccnn_skip_stats.argtypes = [_P, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong)]
    L.ccnn_set_skip_window.argtypes = [_P, _I, _I]
    L.ccnn_set_skip_slot_thresholds.argtypes = [_P, _P]
    L.ccnn_set_check_grid.argtypes = [_P, _I, _I]
    L.ccnn_get_compute_precision.argtypes = [_P]
    L.ccnn_stage1_ex.argtypes = [_P, _I, _I, _I, _P]
    L.ccnn_stage23_ex.argtypes = [_I, _P, _I, _I, _P]
    L.ccnn_get_tile_shape.argtypes = [_P] + [ctypes.POINTER(_I)] * 4
    L.ccnn_autotune.argtypes = [_P, _I, _P, _I, _I, _P, _P]
    L.ccnn_stage_landmarks.argtypes = [_P, _P]
    L.ccnn_set_profiling.argtypes = [_P, _I]
    L.ccnn_last_error.argtypes = [_I]
    L.ccnn_last_timing.argtypes = [_P, _P]
    L.ccnn_last_batch_timing.argtypes = [_P, _P]
    _lib = L
    return L


def _ptr(a):
    return a.ctypes.data_as(_P)


def _image(img):
    """ImageData over a uint8 array; a view whose rows are dense keeps its row step
    (a crop of a wider frame), anything else is copied to a dense array first."""
    img = np.asarray(img)
    ch = 1 if img.ndim == 2 else img.shape[2]
    rows_dense = img.dtype == np.uint8 and img.strides[1] == ch and (img.ndim == 2 or img.strides[2] == 1)
    if not rows_dense or img.strides[0] < img.shape[1] * ch:
        img = np.ascontiguousarray(img, np.uint8)
    h, w = img.shape[:2]
    return img, CImage(w, h, ch, _ptr(img), img.strides[0])


PHASE_SLOTS = 24


def phase_stamps(kernel, nblocks):
    """The phase timestamps of kernel 0 (k_check) or 1 (k_stage1) for the first nblocks
    workgroups, uint64 [nblocks, 24] (100 MHz clock; 0 where not reached), then cleared;
    None when the loaded library was built without them (ccnn_phase_stamps)."""
    out = np.zeros((nblocks, PHASE_SLOTS), np.uint64)
    if load().ccnn_phase_stamps(int(kernel), _ptr(out), int(nblocks)) != 0:
        return None
    return out


class Param:
    """FaceDetector::Param with the reference's defaults (CompactCNNLib.h:98-165)."""

    def __init__(self, **kw):
        c = CParam()
        load().ccnn_param_default(ctypes.byref(c))
        self.max_image_size = (c.max_image_width, c.max_image_height)
        self.min_face_height = c.min_face_height
        self.max_face_height = c.max_face_height
        self.scale_factor = c.scale_factor
        self.min_neighbors = c.min_neighbors
        self.detect_precision = c.detect_precision
        self.treshold = [c.treshold[0], c.treshold[1], c.treshold[2]]
        self.drop_detect = bool(c.drop_detect)
        self.equalize = bool(c.equalize)
        self.reflection = bool(c.reflection)
        self.pipeline = c.pipeline
        self.detect_mode = c.detect_mode
        self.num_threads = c.num_threads
        self.index_output = [1, 0, 0]
        self.device_info = False
        self.cuda_device_id = 0
        self.facial_analysis = False
        for k, v in kw.items():
            if not hasattr(self, k):
                raise AttributeError(k)
            setattr(self, k, v)

    def to_c(self):
        c = CParam()
        load().ccnn_param_default(ctypes.byref(c))
        c.max_image_width, c.max_image_height = self.max_image_size
        c.min_face_height = self.min_face_height
        c.max_face_height = self.max_face_height
        c.scale_factor = self.scale_factor
        c.min_neighbors = self.min_neighbors
        c.detect_precision = self.detect_precision
        for i in range(3):
            c.treshold[i] = self.treshold[i]
            c.index_output[i] = self.index_output[i]
        c.drop_detect = int(self.drop_detect)
        c.equalize = int(self.equalize)
        c.reflection = int(self.reflection)
        c.pipeline = self.pipeline
        c.detect_mode = self.detect_mode
        c.num_threads = self.num_threads
        c.device_info = int(self.device_info)
        c.cuda_device_id = self.cuda_device_id
        c.facial_analysis = int(self.facial_analysis)
        return c

    @classmethod
    def from_c(cls, c):
        p = cls()
        p.max_image_size = (c.max_image_width, c.max_image_height)
        p.min_face_height = c.min_face_height
        p.max_face_height = c.max_face_height
        p.scale_factor = c.scale_factor
        p.min_neighbors = c.min_neighbors
        p.detect_precision = c.detect_precision
        p.treshold = [c.treshold[i] for i in range(3)]
        p.drop_detect = bool(c.drop_detect)
        p.equalize = bool(c.equalize)
        p.reflection = bool(c.reflection)
        p.pipeline = c.pipeline
        p.detect_mode = c.detect_mode
        p.num_threads = c.num_threads
        p.index_output = [c.index_output[i] for i in range(3)]
        p.device_info = bool(c.device_info)
        p.cuda_device_id = c.cuda_device_id
        p.facial_analysis = bool(c.facial_analysis)
        return p


class FaceDetector:
    """CompactCNNLib::FaceDetector over libccnn.so (one detector per object)."""

    def __init__(self):
        self._L = load()
        self._h = self._L.ccnn_create()
        if not self._h:
            raise MemoryError("ccnn_create failed")

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h and getattr(self, "_L", None) is not None:
            self._L.ccnn_destroy(h)

    def Init(self, param):
        c = param.to_c()
        return self._L.ccnn_init(self._h, ctypes.byref(c))

    def Detect(self, img, capacity=-1):
        """Returns (count, faces structured array); count is -1 when not initialised."""
        img, ci = _image(img)
        cap = 4096
        while True:
            # the library writes the first min(n, cap) records; only those are returned
            faces = np.empty(cap, FACE_DTYPE)
            n = self._L.ccnn_detect(self._h, _ptr(faces), cap, ctypes.byref(ci))
            if n <= cap:
                break
            cap = n
        if n < 0:
            return n, np.zeros(0, FACE_DTYPE)
        if capacity >= 0:
            return n, faces[:min(n, capacity)]
        return n, faces[:n]

    def DetectDevice(self, dev_ptr, cols, rows, channels, step):
        """Detection on a frame already resident in GPU memory (device pointer)."""
        ci = CImage(cols, rows, channels, dev_ptr, step)
        cap = 4096
        while True:
            faces = np.empty(cap, FACE_DTYPE)
            n = self._L.ccnn_detect_device(self._h, _ptr(faces), cap, ctypes.byref(ci))
            if n <= cap:
                break
            cap = n              # more faces than the buffer holds: detect again into a bigger one
        return n, faces[:max(0, n)]

    def DetectBatch(self, images=None, device_frames=None, capacity=256):
        """Pipelined detection of several frames. images: list of host arrays, or
        device_frames: list of (device pointer, cols, rows, channels, step).
        Returns a list of (count, faces) per frame."""
        if images is not None:
            keep = [_image(im) for im in images]
            cimgs = (CImage * len(keep))(*[ci for _, ci in keep])
            device = 0
        else:
            try:
                # integer pointers: the records in one pass
                a = np.asarray(device_frames, dtype=np.uint64).reshape(len(device_frames), 5)
                cimgs = np.empty(len(device_frames), CIMAGE_DTYPE)
                cimgs["data"], cimgs["cols"], cimgs["rows"] = a[:, 0], a[:, 1], a[:, 2]
                cimgs["channels"], cimgs["step"] = a[:, 3], a[:, 4]
            except (TypeError, ValueError):
                cimgs = (CImage * len(device_frames))(*[CImage(c, r, ch, p, st)
                                                        for (p, c, r, ch, st) in device_frames])
            device = 1
        n = len(cimgs)
        # the library writes each frame's first min(count, capacity) records; only those
        # are returned
        faces = np.empty(max(1, n * capacity), FACE_DTYPE)
        counts = np.zeros(max(1, n), np.int32)
        ip = cimgs.ctypes.data_as(_P) if isinstance(cimgs, np.ndarray) else cimgs
        if self._L.ccnn_detect_batch(self._h, n, ip, device, _ptr(faces), capacity, _ptr(counts)) != 0:
            raise RuntimeError("ccnn_detect_batch failed")
        return [(c, faces[i * capacity:i * capacity + min(c, capacity)]) for i, c in enumerate(counts[:n].tolist())]

    def set_batch_mode(self, mode):
        """1: DetectBatch runs each batch as one graph launch (default); 0: per-frame
        launches pipelined over three streams."""
        if self._L.ccnn_set_batch_mode(self._h, int(mode)) != 0:
            raise ValueError("bad batch mode")

    def set_compute_precision(self, p):
        """FP32 (exact, default), FP16_BF16, FP16_FP8, FP32_BF16 or FP32_FP8 (stage 1, then
        the stage-2/3 operands)."""
        if self._L.ccnn_set_compute_precision(self._h, int(p)) != 0:
            raise ValueError("bad compute precision")

    def get_compute_precision(self):
        return self._L.ccnn_get_compute_precision(self._h)

    def set_skipnet(self, params, threshold=0.05, min_face=20, scale_factor=1.15):
        """Learned tile skipping on the batch path with a trained network (None: off)."""
        if params is None:
            rc = self._L.ccnn_set_skipnet(self._h, None, 0.5, 0, 0.0)
        else:
            self._skip = np.ascontiguousarray(params, np.float32)
            rc = self._L.ccnn_set_skipnet(self._h, _ptr(self._skip), float(threshold), int(min_face),
                                          float(np.float32(scale_factor)))
        if rc != 0:
            raise ValueError("ccnn_set_skipnet failed")

    def set_check_grid(self, frame_per_cu=0, batch_per_cu=4):
        """Grid of the check kernel: per-frame path frame_per_cu persistent workgroups per
        CU (0: one per item), batch path batch_per_cu."""
        if self._L.ccnn_set_check_grid(self._h, int(frame_per_cu), int(batch_per_cu)) != 0:
            raise ValueError("bad check grid")

    def set_skip_window(self, slot_radius=0, cell_margin=1):
        """Keep window of the tile mask: the scale slots within slot_radius of a level's
        own and the cells within cell_margin of its windows' centres."""
        if self._L.ccnn_set_skip_window(self._h, int(slot_radius), int(cell_margin)) != 0:
            raise ValueError("bad skip window")

    def set_skip_slot_thresholds(self, thresholds):
        """Keep threshold per scale slot (24 probabilities in (0, 1))."""
        self._skip_thr = np.ascontiguousarray(thresholds, np.float32)
        if self._skip_thr.size != 24 or self._L.ccnn_set_skip_slot_thresholds(self._h, _ptr(self._skip_thr)) != 0:
            raise ValueError("bad slot thresholds")

    def skip_stats(self):
        """(stage-1 response tiles of the last batch, tiles run)"""
        a, b = ctypes.c_longlong(), ctypes.c_longlong()
        self._L.ccnn_skip_stats(self._h, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def last_status(self):
        """Replay status of the last frame's candidates (ccnn_last_status): 0 scan-dropped,
        1 dropped before its check, 2 checked, 3 raw window."""
        n = self._L.ccnn_last_status(self._h, None, 0)
        out = np.zeros(max(n, 0), np.int32)
        if n > 0:
            self._L.ccnn_last_status(self._h, _ptr(out), n)
        return out

    def set_tile_shape(self, shape):
        """Stage-1 tile shape id (0 .. 11) of the frames planned from now on."""
        if self._L.ccnn_set_tile_shape(self._h, int(shape)) != 0:
            raise ValueError("bad tile shape")

    def get_tile_shape(self):
        """(id, (ax, ay, tx, ty))"""
        v = [_I() for _ in range(4)]
        i = self._L.ccnn_get_tile_shape(self._h, *[ctypes.byref(x) for x in v])
        return i, tuple(x.value for x in v)

    def autotune(self, images=None, device_frames=None, reps=3):
        """Autotunes the batch path on these frames: returns (chosen shape id, per-shape
        best ms [12], per-grid best ms [3] for 2/3/4 check workgroups per CU)."""
        if images is not None:
            keep = [_image(im) for im in images]
            cimgs = (CImage * len(keep))(*[ci for _, ci in keep])
            device = 0
        else:
            cimgs = (CImage * len(device_frames))(*[CImage(c, r, ch, p, st) for (p, c, r, ch, st) in device_frames])
            device = 1
        sms = np.full(12, np.nan, np.float32)
        gms = np.full(3, np.nan, np.float32)
        best = self._L.ccnn_autotune(self._h, len(cimgs), cimgs, device, int(reps), _ptr(sms), _ptr(gms))
        if best < 0:
            raise RuntimeError("ccnn_autotune failed")
        return best, sms, gms

    def batch_stats(self):
        """(kernel launches in the last batch graph, graph captures so far)"""
        a, b = _I(), _I()
        self._L.ccnn_batch_stats(self._h, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def Clear(self):
        return self._L.ccnn_clear(self._h)

    def isEmpty(self):
        return bool(self._L.ccnn_is_empty(self._h))

    def getParam(self):
        c = CParam()
        self._L.ccnn_get_param(self._h, ctypes.byref(c))
        return Param.from_c(c)

    def setParam(self, param):
        c = param.to_c()
        self._L.ccnn_set_param(self._h, ctypes.byref(c))

    def getMaxImageSize(self):
        w, h = _I(), _I()
        self._L.ccnn_get_max_image_size(self._h, ctypes.byref(w), ctypes.byref(h))
        return w.value, h.value

    def setMaxImageSize(self, size):
        self._L.ccnn_set_max_image_size(self._h, int(size[0]), int(size[1]))

    def getMinObjectHeight(self):
        return self._L.ccnn_get_min_object_height(self._h)

    def setMinObjectHeight(self, v):
        self._L.ccnn_set_min_object_height(self._h, int(v))

    def getMaxObjectHeight(self):
        return self._L.ccnn_get_max_object_height(self._h)

    def setMaxObjectHeight(self, v):
        self._L.ccnn_set_max_object_height(self._h, int(v))

    def getNumThreads(self):
        return self._L.ccnn_get_num_threads(self._h)

    def setNumThreads(self, n):
        self._L.ccnn_set_num_threads(self._h, int(n))

    def getGrayImage32F(self, pipeline=PIPELINE_CPU):
        """Host copy (PIPELINE_CPU) of the last frame's FP32 gray image, or None."""
        ci = CImage()
        if self._L.ccnn_get_gray_image(self._h, ctypes.byref(ci), pipeline) != 0:
            return None
        if pipeline != PIPELINE_CPU:
            return ci
        buf = (ctypes.c_float * (ci.rows * ci.step)).from_address(ci.data)
        return np.frombuffer(buf, np.float32).reshape(ci.rows, ci.step)[:, :ci.cols].copy()

    # -- diagnostics -------------------------------------------------------------
    def set_profiling(self, on=True):
        self._L.ccnn_set_profiling(self._h, int(on))

    def last_timing(self):
        """Per-kernel milliseconds of the last frame, and totals since set_profiling: gray,
        stage1_l1 (k_stage1_l1), stage1 (k_stage1), gap (k_stage1's end to k_check's start),
        check (k_check), after_check (facial analysis, or the read-back without checks),
        total."""
        ms = np.zeros(15, np.float32)
        self._L.ccnn_last_timing(self._h, _ptr(ms))
        keys = ("gray", "stage1_l1", "stage1", "gap", "check", "after_check", "total")
        last = {k: float(ms[i]) for i, k in enumerate(keys)}
        totals = {k: float(ms[7 + i]) for i, k in enumerate(keys)}
        return dict(last=last, totals=totals, frames=int(ms[14]))

    def last_batch_timing(self):
        """Milliseconds per kernel group of the last batch run with profiling on (the
        kernels then run without the graph): gray, skipnet, stage1_l1, stage1, compact,
        check, total, and the batch's frame count."""
        ms = np.zeros(8, np.float32)
        self._L.ccnn_last_batch_timing(self._h, _ptr(ms))
        keys = ("gray", "skipnet", "stage1_l1", "stage1", "compact", "check", "total")
        out = {k: float(ms[i]) for i, k in enumerate(keys)}
        out["frames"] = int(ms[7])
        return out

    def last_facial(self):
        """Facial analysis of the last frame's candidates: int array [n, 4] of
        (ok, gender, smile, glasses), in candidate order."""
        n = self._L.ccnn_last_facial(self._h, None, 0)
        out = np.zeros((max(n, 0), 4), np.int32)
        if n > 0:
            self._L.ccnn_last_facial(self._h, _ptr(out), n)
        return out

    def last_trace(self):
        """Levels run and candidate windows of the last Detect, with their check results."""
        nl = _I()
        n = self._L.ccnn_last_trace(self._h, ctypes.byref(nl), None, 0, None, None)
        levels = np.zeros((max(1, nl.value), 6), np.int32)
        cand = np.zeros((max(1, n), 4), np.float32)
        res = np.zeros((max(1, n), 8), np.float32)
        self._L.ccnn_last_trace(self._h, ctypes.byref(nl), _ptr(levels), n, _ptr(cand), _ptr(res))
        return dict(levels=levels[:nl.value], cand=cand[:n], res=res[:n])


    def last_responses(self):
        """Stage-1 response maps of the frame last_trace describes: a list of float32
        arrays [rh, rw], one per level in last_trace's level order."""
        tr = self.last_trace()
        n = self._L.ccnn_last_responses(self._h, None, 0)
        if n < 0:
            raise RuntimeError("no response maps recorded")
        buf = np.zeros(max(1, n), np.float32)
        if self._L.ccnn_last_responses(self._h, _ptr(buf), n) != n:
            raise RuntimeError("ccnn_last_responses failed")
        out, off = [], 0
        for (_, _, _, rw, rh, _) in tr["levels"]:
            out.append(buf[off:off + rw * rh].reshape(rh, rw).copy())
            off += rw * rh
        return out


def last_error(reset=True):
    """Last HIP error code met by the library on this thread (0 if none)."""
    return load().ccnn_last_error(int(reset))


# -- stage-level entry points --------------------------------------------------------
def stage_gray(img):
    img, ci = _image(img)
    out = np.zeros((ci.rows, ci.cols), np.float32)
    if load().ccnn_stage_gray(ctypes.byref(ci), _ptr(out)) != 0:
        raise RuntimeError("ccnn_stage_gray failed")
    return out


def stage_level(gray, lw, lh, mode):
    gray = np.ascontiguousarray(gray, np.float32)
    gh, gw = gray.shape
    out = np.zeros((lh, lw), np.float32)
    if load().ccnn_stage_level(_ptr(gray), gw, gh, lw, lh, mode, _ptr(out)) != 0:
        raise RuntimeError("ccnn_stage_level failed")
    return out


def stage1(level, precision=FP32):
    level = np.ascontiguousarray(level, np.float32)
    lh, lw = level.shape
    rw, rh = _I(), _I()
    L = load()
    if L.ccnn_stage1_size(lw, lh, ctypes.byref(rw), ctypes.byref(rh)) != 0:
        raise ValueError("level too small")
    out = np.zeros((rh.value, rw.value), np.float32)
    if precision == FP32:
        rc = L.ccnn_stage1(_ptr(level), lw, lh, _ptr(out))
    else:
        rc = L.ccnn_stage1_ex(_ptr(level), lw, lh, int(precision), _ptr(out))
    if rc != 0:
        raise RuntimeError("ccnn_stage1 failed")
    return out


def stage_landmarks(crop):
    """The facial-analysis network on one 51x49 FP32 crop: its 13 outputs."""
    crop = np.ascontiguousarray(crop, np.float32)
    if crop.shape != (51, 49):
        raise ValueError("landmarks input is 49x51")
    out = np.zeros(13, np.float32)
    if load().ccnn_stage_landmarks(_ptr(crop), _ptr(out)) != 0:
        raise RuntimeError("ccnn_stage_landmarks failed")
    return out


def stage_check(param, gray, scales, xs, ys):
    gray = np.ascontiguousarray(gray, np.float32)
    gh, gw = gray.shape
    scales = np.ascontiguousarray(scales, np.float32)
    xs = np.ascontiguousarray(xs, np.int32)
    ys = np.ascontiguousarray(ys, np.int32)
    n = len(scales)
    res = np.zeros((max(1, n), 8), np.float32)
    c = param.to_c()
    if load().ccnn_stage_check(ctypes.byref(c), _ptr(gray), gw, gh, n, _ptr(scales), _ptr(xs), _ptr(ys),
                               _ptr(res)) != 0:
        raise RuntimeError("ccnn_stage_check failed")
    return res[:n]


def stage23(stage, inputs, precision=FP32):
    """The stage-2 (stage=1) or stage-3 (stage=2) network on float32 inputs [n, 51, 47]:
    their 4x4 response maps [n, 4, 4], in a compute precision of this library."""
    x = np.ascontiguousarray(inputs, np.float32).reshape(-1, 51, 47)
    out = np.zeros((max(1, len(x)), 4, 4), np.float32)
    if load().ccnn_stage23_ex(stage, _ptr(x), len(x), int(precision), _ptr(out)) != 0:
        raise RuntimeError("ccnn_stage23_ex failed")
    return out[:len(x)]
