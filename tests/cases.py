"""Parity cases shared by the golden-fixture script and the tests.

Each case is (name, frame builder, parameter overrides). Frames come from the seeded
generator (compactcnncascade_amd/synth.py) or from tests/golden/photo3.jpg (the
reference's own test_images/3.jpg), so the same inputs are rebuilt on every machine.
"""
import os

import numpy as np

from compactcnncascade_amd import synth

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def photo3():
    from PIL import Image
    rgb = np.array(Image.open(os.path.join(GOLDEN, "photo3.jpg")).convert("RGB"))
    return np.ascontiguousarray(rgb[:, :, ::-1])


def vga(seed, lam=6, smin=40, smax=200):
    return synth.synth_frame(seed, 640, 480, lam, smin, smax)[0]


def fhd(idx):
    return synth.workload_frame("fhd_min40_max240", idx)[0]


def _bgra(f):
    return np.ascontiguousarray(np.concatenate([f, np.full(f.shape[:2] + (1,), 255, np.uint8)], axis=2))


def _gray(f):
    return np.ascontiguousarray(f[:, :, 1])


# name -> (builder, params); params are overrides of the FaceDetector::Param defaults
# with max_image_size = frame size.
CASES = {
    "vga_s0_min40": (lambda: vga(0), dict(min_face=40)),
    "vga_s2_min40": (lambda: vga(2), dict(min_face=40)),
    "vga_s3_min40": (lambda: vga(3), dict(min_face=40)),
    "vga_s5_min40": (lambda: vga(5), dict(min_face=40)),
    "vga_s11_min20": (lambda: vga(11, smin=20), dict(min_face=20)),
    "vga_s11_low": (lambda: vga(11, smin=20), dict(min_face=20, precision=1)),
    "vga_s11_high": (lambda: vga(11, smin=20), dict(min_face=20, precision=3)),
    "vga_s11_ultra": (lambda: vga(11, smin=20), dict(min_face=20, precision=4)),
    "vga_s11_def": (lambda: vga(11, smin=20), dict(min_face=20, precision=0)),
    "vga_s11_disable": (lambda: vga(11, smin=20), dict(min_face=40, detect_mode=0)),
    "vga_s11_sync": (lambda: vga(11, smin=20), dict(min_face=20, detect_mode=1)),
    "vga_s11_noeq_norefl_nodrop": (lambda: vga(11, smin=20), dict(min_face=20, equalize=0, reflection=0,
                                                                  drop_detect=0)),
    "vga_s11_bgra_min30": (lambda: _bgra(vga(11, smin=20)), dict(min_face=30)),
    "vga_s11_gray_min24": (lambda: _gray(vga(11, smin=20)), dict(min_face=24)),
    "vga_s7_sf105_max120": (lambda: vga(7, lam=10, smin=25, smax=150), dict(min_face=25, max_face=120,
                                                                            scale_factor=1.05)),
    "ragged_333x211": (lambda: synth.synth_frame(21, 333, 211, 4, 30, 120)[0], dict(min_face=30)),
    "min_image_47x51": (lambda: synth.synth_frame(22, 47, 51, 0)[0], dict(min_face=20)),
    "below_min_46x51": (lambda: synth.synth_frame(23, 46, 51, 0)[0], dict(min_face=20)),
    "photo3_min20": (photo3, dict(min_face=20)),
    "photo3_min40_max200": (photo3, dict(min_face=40, max_face=200)),
    "fhd_w0_min40_max240": (lambda: fhd(0), dict(min_face=40, max_face=240)),
    "fhd_w1_min40_max240": (lambda: fhd(1), dict(min_face=40, max_face=240)),
    # facial analysis: the landmarks network on accepted windows, which may drop them,
    # and the gender / smile / glasses attributes of each face
    "vga_s3_min40_facial": (lambda: vga(3), dict(min_face=40, facial_analysis=1)),
    "vga_s11_min20_facial": (lambda: vga(11, smin=20), dict(min_face=20, facial_analysis=1)),
    "photo3_min20_facial": (photo3, dict(min_face=20, facial_analysis=1)),
    "photo3_min40_max200_facial_high": (photo3, dict(min_face=40, max_face=200, precision=3, facial_analysis=1)),
    "vga_s11_gray_min24_facial": (lambda: _gray(vga(11, smin=20)), dict(min_face=24, facial_analysis=1)),
    # full-size frames of the mixed-size batch and the deep pyramid (goldens from the
    # compiled reference; too slow for the oracle in the CPU suite)
    "big_hd_mixed0_min20": (lambda: synth.workload_frame("mixed_hd", 0)[0], dict(min_face=20)),
    "big_uhd_mixed0_min20": (lambda: synth.workload_frame("mixed_uhd", 0)[0], dict(min_face=20)),
    "big_deep_4800x2400_sf105_min15": (lambda: synth.workload_frame("deep_4800x2400_sf105_min15", 0)[0],
                                       dict(min_face=15, scale_factor=1.05)),
}

# 1-channel frames with min face >= 34 take the gray_image_only pre-resize path
# (cnn_detector_v3.cpp:299-306, 2531-2539). The compiled reference cannot run it on
# Linux (posix_memalign with alignment 1 for the pre-resize buffer, image.cpp:40), so
# these expectations come from the oracle alone.
ORACLE_ONLY_CASES = {
    "vga_s11_gray_min40_preresize": (lambda: _gray(vga(11, smin=20)), dict(min_face=40)),
    "photo3_gray_min48_preresize": (lambda: _gray(photo3()), dict(min_face=48)),
}


def params_for(frame, overrides):
    from oracle.bind import default_params
    h, w = frame.shape[:2]
    return default_params(w, h, **overrides)


def api_param(frame, overrides):
    """The same settings as a compactcnncascade_amd.api.Param."""
    return api_param_of(params_for(frame, overrides))


def api_param_of(p):
    """compactcnncascade_amd.api.Param of an oracle parameter dict."""
    from compactcnncascade_amd.api import Param
    return Param(max_image_size=(p["max_w"], p["max_h"]), min_face_height=p["min_face"],
                 max_face_height=p["max_face"], min_neighbors=p["min_neighbors"],
                 detect_precision=p["precision"], drop_detect=bool(p["drop_detect"]),
                 equalize=bool(p["equalize"]), reflection=bool(p["reflection"]), detect_mode=p["detect_mode"],
                 scale_factor=p["scale_factor"], treshold=[p["t1"], p["t2"], p["t3"]],
                 facial_analysis=bool(p["facial_analysis"]))


def frame_digest(frame):
    import hashlib
    return hashlib.sha256(np.ascontiguousarray(frame).tobytes()).hexdigest()[:16]


# Sequences of frames detected in turn on ONE detector, whose state carries from call to
# call: max_image_size grows with a larger frame, and num_scales (the scales RunCPUDetect
# visits) is lowered by each frame's skip rule and restored only when the frame size
# differs from max_image_size (cnn_detector_v3.cpp:603, 2083-2088, 2499-2518). A frame
# of the maximum size after a smaller one therefore skips the scales the smaller one
# could not run. name -> (frame builders, parameter overrides, max_image_size[,
# min_face set with setMinObjectHeight before each frame, 0 for none]).
def _big_faces(seed):
    return synth.synth_frame(seed, 640, 480, 4, 150, 260)[0]


def _small(seed):
    return synth.synth_frame(seed + 100, 400, 300, 2, 40, 120)[0]


SEQ_CASES = {
    "seq_max_small_max_s45": ([lambda: _big_faces(45), lambda: _small(45), lambda: _big_faces(45),
                               lambda: _big_faces(49)], dict(min_face=40), (640, 480)),
    "seq_max_small_max_s41_min20": ([lambda: _big_faces(41), lambda: _small(41), lambda: _big_faces(41)],
                                    dict(min_face=20), (640, 480)),
    # the first frame is smaller than max_image_size, the second larger (max grows)
    "seq_grow_s41": ([lambda: _small(41), lambda: _big_faces(41), lambda: _small(41), lambda: _big_faces(41)],
                     dict(min_face=40), (400, 300)),
    # a colour frame turns gray_image_only off; setMinObjectHeight keeps it off, so the
    # following 1-channel frame is not pre-resized (cnn_detector_v3.h:417-436)
    "seq_setter_gray_after_colour": ([lambda: vga(11, smin=20), lambda: _gray(vga(11, smin=20))],
                                     dict(min_face=40), (640, 480), [0, 36]),
}

# The benchmark workloads as frame sequences on one detector (the order DetectBatch and
# bench.py use): name -> ([(workload, count)], parameter overrides, max_image_size).
WORKLOAD_SEQS = {
    # Full HD drop-in: 64 frames 1920x1080, faces 40..240
    "fhd_min40_max240_x64": ([("fhd_min40_max240", 64)], dict(min_face=40, max_face=240), (1920, 1080)),
    # the mixed-size batch: 8 each of 640x480, 1280x720, 1920x1080, 3840x2160, faces >= 20,
    # one detector initialised for the largest frame
    "mixed_batch_x32": ([("mixed_vga", 8), ("mixed_hd", 8), ("mixed_fhd", 8), ("mixed_uhd", 8)],
                        dict(min_face=20), (3840, 2160)),
}


def seq_params(overrides, max_size):
    from oracle.bind import default_params
    return default_params(max_size[0], max_size[1], **overrides)


def workload_frames(spec):
    out = []
    for name, n in spec:
        out.extend(synth.workload_frame(name, i)[0] for i in range(n))
    return out
