"""ctypes bindings for the CPU checkers built by oracle/Makefile.

TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module, and only as the checker or the timed CPU
baseline; the product never does.

  Oracle  -- the project's C restatement (oracle/ccnn_oracle.c)
  Ref     -- the reference's own CPU path compiled from its sources
             (oracle/_ref/libccnn_ref_O0.so: CMake-default flags, golden producer;
              oracle/_ref/libccnn_ref_O3.so: -O3, the timed baseline)
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")
MODEL_DIR = os.path.join(os.path.dirname(HERE), "compactcnncascade_amd", "models")
MODEL_FILES = ["cnn4face1_cntk.bin", "cnn4face2_cntk.bin", "cnn4face3_cntk.bin", "cnn4landmarks_cntk.bin"]
# FaceDetector::Param::index_output defaults; the landmarks network computes every output
# (AdvancedParam::index_output[3] = -1, cnn_detector_v3.h:170-173)
INDEX_OUTPUT = [1, 0, 0, -1]

_P = ctypes.c_void_p
_I = ctypes.c_int
_F = ctypes.c_float


def _ptr(a):
    return a.ctypes.data_as(_P)


class OcModel(ctypes.Structure):
    _fields_ = [("min_w", _I), ("min_h", _I), ("maps", _I * 3), ("kw", _I * 3), ("kh", _I * 3),
                ("kern", _P * 3), ("af_scale", _F), ("conv_b", _P * 3), ("lrelu_w1", _P * 3),
                ("lrelu_w2", _P * 3), ("bn_w", _P * 3), ("bn_b", _P * 3),
                ("full_connect", _I), ("hl", _I), ("connect", _I), ("hl_scale", _I),
                ("hl_w", _P), ("hl_nin", _I), ("hl_b", _P), ("hl_tanh_w", _P), ("hl_bn_w", _P),
                ("hl_bn_b", _P), ("ol", _I), ("ol_w", _P), ("ol_b", _P), ("ol_tanh_w", _F),
                ("index_output", _I)]


class OcParams(ctypes.Structure):
    _fields_ = [("max_w", _I), ("max_h", _I), ("min_face", _I), ("max_face", _I),
                ("min_neighbors", _I), ("precision", _I), ("drop_detect", _I), ("equalize", _I),
                ("reflection", _I), ("detect_mode", _I), ("scale_factor", _F),
                ("t1", _F), ("t2", _F), ("t3", _F), ("facial_analysis", _I)]


class OcTrace(ctypes.Structure):
    _fields_ = [("n_scales", _I), ("scales", _F * 512), ("n_levels", _I),
                ("lvl_scl", _I * 512), ("lvl_w", _I * 512), ("lvl_h", _I * 512),
                ("lvl_rw", _I * 512), ("lvl_rh", _I * 512), ("lvl_mode", _I * 512),
                ("cap", _I), ("n_cand", _I),
                ("c_scl", _P), ("c_x", _P), ("c_y", _P), ("c_score", _P), ("c_status", _P),
                ("c_mod", _P), ("c_k1", _P), ("c_m1", _P), ("c_k2", _P), ("c_m2", _P),
                ("n_raw", _I)]


class OcState(ctypes.Structure):
    _fields_ = [("max_w", _I), ("max_h", _I), ("num_scales", _I), ("gray_only", _I)]


def default_params(w, h, **kw):
    """FaceDetector::Param defaults (prebuild/include/CompactCNNLib.h:98-165) with the
    frame size as max_image_size."""
    p = dict(max_w=w, max_h=h, min_face=80, max_face=0, min_neighbors=2, precision=2,
             drop_detect=1, equalize=1, reflection=1, detect_mode=2, scale_factor=1.15,
             t1=-1.5, t2=-1.25, t3=-1.0, facial_analysis=0)
    p.update(kw)
    return p


def _oc_params(p):
    return OcParams(**{k: p[k] for k, _ in OcParams._fields_})


def model_bytes(i):
    with open(os.path.join(MODEL_DIR, MODEL_FILES[i]), "rb") as f:
        return f.read()


class Oracle:
    def __init__(self, path=None):
        self.lib = ctypes.CDLL(path or os.path.join(REF_DIR, "libccnn_oracle.so"))
        L = self.lib
        L.oc_model_load.argtypes = [ctypes.POINTER(OcModel), ctypes.c_char_p, ctypes.c_long, _I]
        L.oc_cnn_forward.argtypes = [ctypes.POINTER(OcModel), _P, _I, _I, _I, _P]
        L.oc_output_size.argtypes = [ctypes.POINTER(OcModel), _I, _I, ctypes.POINTER(_I), ctypes.POINTER(_I)]
        L.oc_gray.argtypes = [_P, _I, _I, _I, _I, _P, _I]
        L.oc_resize_f32.argtypes = [_P, _I, _I, _I, _P, _I, _I, _I, _I]
        L.oc_resize_u8.argtypes = [_P, _I, _I, _I, _P, _I, _I, _I, _I]
        L.oc_equalize.argtypes = [_P, _I, _I, _I]
        L.oc_check_prep.argtypes = [_P, _I, _I, _I, _I, _I, _P, _P, _P, _P]
        L.oc_sort_desc.argtypes = [_P, _P, _I]
        L.oc_detect.argtypes = [ctypes.POINTER(OcModel)] * 4 + [ctypes.POINTER(OcParams), _P, _I, _I, _I, _I,
                                                                _P, _I, ctypes.POINTER(OcTrace)]
        L.oc_detect_state.argtypes = [ctypes.POINTER(OcModel)] * 4 + [ctypes.POINTER(OcParams),
                                                                      ctypes.POINTER(OcState), _P, _I, _I, _I, _I,
                                                                      _P, _I, ctypes.POINTER(OcTrace)]
        L.oc_state_init.argtypes = [ctypes.POINTER(OcState), ctypes.POINTER(OcParams)]
        L.oc_check_candidate.argtypes = [ctypes.POINTER(OcModel)] * 2 + [ctypes.POINTER(OcParams), _P, _I, _I, _I,
                                                                         _F, _I, _I, _P]
        self.models = []
        for i in range(4):
            m = OcModel()
            b = model_bytes(i)
            if L.oc_model_load(ctypes.byref(m), b, len(b), INDEX_OUTPUT[i]) != 0:
                raise RuntimeError("oracle: model %d failed to load" % i)
            self.models.append(m)

    def output_size(self, stage, w, h):
        ow, oh = _I(), _I()
        self.lib.oc_output_size(ctypes.byref(self.models[stage]), w, h, ctypes.byref(ow), ctypes.byref(oh))
        return ow.value, oh.value

    def forward(self, stage, img):
        """Output map [oh, ow]; for the landmarks network (stage 3) all outputs, [ol, oh, ow]."""
        img = np.ascontiguousarray(img, np.float32)
        h, w = img.shape
        ow, oh = self.output_size(stage, w, h)
        m = self.models[stage]
        out = np.zeros((m.ol, oh, ow) if m.index_output < 0 else (oh, ow), np.float32)
        if self.lib.oc_cnn_forward(ctypes.byref(self.models[stage]), _ptr(img), w, h, w, _ptr(out)) != 0:
            raise ValueError("image below the model minimum")
        return out

    def gray(self, img):
        img = np.ascontiguousarray(img, np.uint8)
        h, w = img.shape[:2]
        ch = 1 if img.ndim == 2 else img.shape[2]
        out = np.zeros((h, w), np.float32)
        rc = self.lib.oc_gray(_ptr(img), w, h, ch, w * ch, _ptr(out), w)
        if rc != 0:
            raise ValueError("unsupported channel count")
        return out

    def resize(self, src, dw, dh, mode):
        src = np.ascontiguousarray(src)
        sh, sw = src.shape
        if src.dtype == np.uint8:
            out = np.zeros((dh, dw), np.uint8)
            self.lib.oc_resize_u8(_ptr(src), sw, sh, sw, _ptr(out), dw, dh, dw, mode)
        else:
            src = src.astype(np.float32)
            out = np.zeros((dh, dw), np.float32)
            self.lib.oc_resize_f32(_ptr(src), sw, sh, sw, _ptr(out), dw, dh, dw, mode)
        return out

    def equalize(self, img):
        img = np.array(img, np.uint8, copy=True)
        h, w = img.shape
        self.lib.oc_equalize(_ptr(img), w, h, w)
        return img

    def check_prep(self, crop, equalize=1, reflection=1):
        crop = np.ascontiguousarray(crop, np.float32)
        h, w = crop.shape
        u8 = np.zeros((51, 49), np.uint8)
        s2 = np.zeros((51, 47), np.float32)
        s3 = np.zeros((51, 47), np.float32)
        f32 = np.zeros((51, 49), np.float32)
        self.lib.oc_check_prep(_ptr(crop), w, h, w, equalize, reflection, _ptr(u8), _ptr(s2), _ptr(s3), _ptr(f32))
        return u8, s2, s3, f32

    def sort_desc(self, keys):
        keys = np.ascontiguousarray(keys, np.float32)
        perm = np.arange(len(keys), dtype=np.int32)
        self.lib.oc_sort_desc(_ptr(perm), _ptr(keys), len(keys))
        return perm

    def check_candidate(self, params, gray, scale, px, py):
        gray = np.ascontiguousarray(gray, np.float32)
        gh, gw = gray.shape
        res = np.zeros(10, np.float32)
        p = _oc_params(params)
        self.lib.oc_check_candidate(ctypes.byref(self.models[1]), ctypes.byref(self.models[2]), ctypes.byref(p),
                                    _ptr(gray), gw, gh, gw, float(scale), int(px), int(py), _ptr(res))
        return res

    def detect(self, img, params, trace_cap=0):
        """Returns (faces [n,8] float32 x,y,w,h,score,gender,smile,glasses, trace dict or None)."""
        img = np.ascontiguousarray(img, np.uint8)
        h, w = img.shape[:2]
        ch = 1 if img.ndim == 2 else img.shape[2]
        p = _oc_params(params)
        cap = 100000
        raw = np.zeros(cap * 8, np.int32)
        tr = None
        arrays = None
        if trace_cap > 0:
            tr = OcTrace()
            tr.cap = trace_cap
            arrays = {k: np.zeros(trace_cap, np.float32 if k in ("c_score", "c_m1", "c_m2") else np.int32)
                      for k in ("c_scl", "c_x", "c_y", "c_score", "c_status", "c_mod", "c_k1", "c_m1", "c_k2", "c_m2")}
            for k, a in arrays.items():
                setattr(tr, k, _ptr(a))
        n = self.lib.oc_detect(*(ctypes.byref(m) for m in self.models), ctypes.byref(p), _ptr(img), w, h, ch,
                               w * ch, _ptr(raw), cap, ctypes.byref(tr) if tr is not None else None)
        if n > cap:
            # more faces than the buffer holds (raw boxes without Merger): run again
            cap = n
            raw = np.zeros(cap * 8, np.int32)
            n = self.lib.oc_detect(*(ctypes.byref(m) for m in self.models), ctypes.byref(p), _ptr(img), w, h, ch,
                                   w * ch, _ptr(raw), cap, ctypes.byref(tr) if tr is not None else None)
        if n < 0:
            return n, None
        rec = raw[:8 * n].reshape(n, 8)
        faces = np.zeros((n, 8), np.float32)
        faces[:, :4] = rec[:, :4]
        faces[:, 4] = rec[:, 4].view(np.float32)
        faces[:, 5:] = rec[:, 5:]
        trace = None
        if tr is not None:
            nc = tr.n_cand
            trace = {k: a[:nc].copy() for k, a in arrays.items()}
            trace["scales"] = np.array(tr.scales[:tr.n_scales], np.float32)
            nl = tr.n_levels
            for k in ("lvl_scl", "lvl_w", "lvl_h", "lvl_rw", "lvl_rh", "lvl_mode"):
                trace[k] = np.array(getattr(tr, k)[:nl], np.int32)
            trace["n_raw"] = tr.n_raw
        return faces, trace


    def new_state(self, params):
        """Detector state after FaceDetector::Init with params (an OcState)."""
        st = OcState()
        self.lib.oc_state_init(ctypes.byref(st), ctypes.byref(_oc_params(params)))
        return st

    def detect_state(self, img, params, state):
        """Detection of one frame on a detector in `state` (updated in place); faces
        [n,8] as detect."""
        p = _oc_params(params)
        img = np.ascontiguousarray(img, np.uint8)
        h, w = img.shape[:2]
        ch = 1 if img.ndim == 2 else img.shape[2]
        cap = 100000
        raw = np.zeros(cap * 8, np.int32)
        n = self.lib.oc_detect_state(*(ctypes.byref(m) for m in self.models), ctypes.byref(p), ctypes.byref(state),
                                     _ptr(img), w, h, ch, w * ch, _ptr(raw), cap, None)
        if n < 0 or n > cap:
            raise RuntimeError("oracle detect_state failed (%d)" % n)
        rec = raw[:8 * n].reshape(n, 8)
        faces = np.zeros((n, 8), np.float32)
        faces[:, :4] = rec[:, :4]
        faces[:, 4] = rec[:, 4].view(np.float32)
        faces[:, 5:] = rec[:, 5:]
        return faces

    def detect_seq(self, frames, params, min_face_sets=None):
        """Detection of each frame in turn on ONE detector (its max_image_size, num_scales
        and gray_image_only carry from call to call, as CNNDetector's do). min_face_sets[i]
        > 0: setMinObjectHeight before frame i (the detector is Cleared and Inits at the
        next Detect with gray_image_only as it stands). Returns a list of faces arrays
        [n,8] as detect."""
        params = dict(params)
        st = self.new_state(params)
        out = []
        for i, img in enumerate(frames):
            if min_face_sets is not None and min_face_sets[i] > 0:
                params["min_face"] = int(min_face_sets[i])
                st.num_scales = -1
            out.append(self.detect_state(img, params, st))
        return out


class Ref:
    """The reference CPU path compiled from its own sources (oracle/Makefile)."""

    def __init__(self, flavour="O0"):
        self.lib = ctypes.CDLL(os.path.join(REF_DIR, "libccnn_ref_%s.so" % flavour))
        L = self.lib
        L.ref_gray.argtypes = [_P, _I, _I, _I, _I, _P, _I]
        L.ref_resize.argtypes = [_P, _I, _I, _I, _P, _I, _I, _I, _I, _I, _I]
        L.ref_resize_u8.argtypes = [_P, _I, _I, _I, _P, _I, _I, _I, _I]
        L.ref_cnn_create.argtypes = [ctypes.c_char_p, _I]
        L.ref_cnn_create.restype = _P
        L.ref_cnn_destroy.argtypes = [_P]
        L.ref_cnn_forward.argtypes = [_P, _P, _I, _I, _I, _I, _I, _P, _I, ctypes.POINTER(_I), ctypes.POINTER(_I)]
        L.ref_check_prep.argtypes = [_P, _I, _I, _I, _I, _I, _I, _I, _P, _P, _P]
        L.ref_equalize.argtypes = [_P, _I, _I, _I]
        L.ref_detect.argtypes = [_P, _P, _P, _I, _I, _I, _I, _P, _I]
        L.ref_detect_ex.argtypes = [_P, _P, _P, _I, _I, _I, _I, _P, _I]
        L.ref_detect_loop.argtypes = [_P, _P, _P, _I, _I, _I, _I, _I, _I]
        L.ref_detect_loop.restype = ctypes.c_long
        L.ref_detect_seq.argtypes = [_P, _P, _P, _P, _I, _P, _I, _P]
        self.cnn = [L.ref_cnn_create(os.path.join(MODEL_DIR, MODEL_FILES[i]).encode(), INDEX_OUTPUT[i])
                    for i in range(4)]
        if not all(self.cnn):
            raise RuntimeError("reference: model load failed")

    def gray(self, img):
        img = np.ascontiguousarray(img, np.uint8)
        h, w = img.shape[:2]
        ch = 1 if img.ndim == 2 else img.shape[2]
        out = np.zeros((h, w), np.float32)
        self.lib.ref_gray(_ptr(img), w, h, ch, w * ch, _ptr(out), w)
        return out

    def resize(self, src, dw, dh, mode):
        src = np.ascontiguousarray(src)
        sh, sw = src.shape
        if src.dtype == np.uint8:
            out = np.zeros((dh, dw), np.uint8)
            self.lib.ref_resize_u8(_ptr(src), sw, sh, sw, _ptr(out), dw, dh, dw, mode)
        else:
            src = src.astype(np.float32)
            out = np.zeros((dh, dw), np.float32)
            self.lib.ref_resize(_ptr(src), sw, sh, sw, _ptr(out), dw, dh, dw, mode, sw, sh)
        return out

    def forward(self, stage, img, alloc=None):
        img = np.ascontiguousarray(img, np.float32)
        h, w = img.shape
        aw, ah = alloc or (w, h)
        cap = max(1, w * h)
        out = np.zeros(cap, np.float32)
        ow, oh = _I(), _I()
        rc = self.lib.ref_cnn_forward(self.cnn[stage], _ptr(img), w, h, w, aw, ah, _ptr(out), cap,
                                      ctypes.byref(ow), ctypes.byref(oh))
        if rc != 0:
            raise RuntimeError("reference forward failed")
        return out[:ow.value * oh.value].reshape(oh.value, ow.value).copy()

    def check_prep(self, crop, equalize=1, reflection=1):
        crop = np.ascontiguousarray(crop, np.float32)
        h, w = crop.shape
        u8 = np.zeros((51, 49), np.uint8)
        s2 = np.zeros((51, 47), np.float32)
        s3 = np.zeros((51, 47), np.float32)
        self.lib.ref_check_prep(_ptr(crop), w, h, w, w, h, equalize, reflection, _ptr(u8), _ptr(s2), _ptr(s3))
        return u8, s2, s3

    def equalize(self, img):
        img = np.array(img, np.uint8, copy=True)
        h, w = img.shape
        self.lib.ref_equalize(_ptr(img), w, h, w)
        return img

    @staticmethod
    def _pack(p):
        ip = np.array([p["max_w"], p["max_h"], p["min_face"], p["max_face"], p["min_neighbors"], p["precision"],
                       p["drop_detect"], p["equalize"], p["reflection"], p["detect_mode"],
                       p.get("facial_analysis", 0)], np.int32)
        fp = np.array([p["scale_factor"], p["t1"], p["t2"], p["t3"]], np.float32)
        return ip, fp

    def detect(self, img, params):
        """faces [n,8] float32: x, y, w, h, score, gender, smile, glasses."""
        img = np.ascontiguousarray(img, np.uint8)
        h, w = img.shape[:2]
        ch = 1 if img.ndim == 2 else img.shape[2]
        ip, fp = self._pack(params)
        cap = 100000
        out = np.zeros((cap, 8), np.float32)
        n = self.lib.ref_detect_ex(_ptr(ip), _ptr(fp), _ptr(img), w, h, ch, w * ch, _ptr(out), cap)
        if n < 0:
            return n
        return out[:n].copy()

    def detect_loop(self, frames, params, reps=1):
        frames = np.ascontiguousarray(frames, np.uint8)
        nf, h, w = frames.shape[:3]
        ch = 1 if frames.ndim == 3 else frames.shape[3]
        ip, fp = self._pack(params)
        return self.lib.ref_detect_loop(_ptr(ip), _ptr(fp), _ptr(frames), nf, w, h, ch, w * ch, reps)

    def detect_seq(self, frames, params, min_face_sets=None, cap=20000):
        """Detection of each frame in turn on ONE reference detector (state carried
        across calls; min_face_sets as Oracle.detect_seq). Returns a list of faces arrays
        [n,8] as detect."""
        frames = [np.ascontiguousarray(f, np.uint8) for f in frames]
        sets = min_face_sets if min_face_sets is not None else [0] * len(frames)
        meta = np.array([[f.shape[1], f.shape[0], 1 if f.ndim == 2 else f.shape[2],
                          f.shape[1] * (1 if f.ndim == 2 else f.shape[2]), int(m)] for f, m in zip(frames, sets)],
                        np.int32)
        buf = np.ascontiguousarray(np.concatenate([f.reshape(-1) for f in frames]))
        ip, fp = self._pack(params)
        out = np.zeros((len(frames), cap, 8), np.float32)
        counts = np.zeros(len(frames), np.int32)
        rc = self.lib.ref_detect_seq(_ptr(ip), _ptr(fp), _ptr(buf), _ptr(meta), len(frames), _ptr(out), cap,
                                     _ptr(counts))
        if rc < 0:
            raise RuntimeError("reference detect_seq failed (%d)" % rc)
        if (counts > cap).any():
            raise RuntimeError("reference detect_seq: more faces than cap")
        return [out[i, :counts[i]].copy() for i in range(len(frames))]
