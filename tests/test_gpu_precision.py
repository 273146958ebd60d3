"""Reduced compute precisions (ccnn_set_compute_precision): FP16 stage 1 and BF16 / FP8
matrix-core stages 2/3.

Each reduced-precision map is compared with the FP64 oracle of the network
(oracle/net64.py) and its largest absolute error is stated; the exact FP32 path is
measured against the same oracle for scale. Detection sets are judged by the recall
judge (compactcnncascade_amd/recall.py: pairing at IoU >= 0.99 with Rect::overlap)
against the exact FP32 path on the mixed-size batch and the 4800x2400 deep pyramid.
Tolerances are written in each test.
"""
import numpy as np
import pytest

from cases import photo3, params_for

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def api():
    from compactcnncascade_amd import api as A
    A.load()
    return A


def check_inputs(oracle, gray, scale, px, py):
    """The stage-2 and stage-3 inputs of one candidate's first crop (CPUCheckDetect's crop
    geometry, cnn_detector_v3.cpp:1491-1556, then the oracle's crop preparation)."""
    gh, gw = gray.shape
    inv = np.float32(1.0) / np.float32(scale)
    x = int(np.float32(px - 6) * inv)
    y = int(np.float32(py - 5) * inv)
    w = int(np.float32(36.0) * inv)
    h = int(np.float32(38.0) * inv)
    rx, ry = min(max(x, 0), gw), min(max(y, 0), gh)
    rc, rr = min(max(x + w, 0), gw) - rx, min(max(y + h, 0), gh) - ry
    _, s2, s3, _ = oracle.check_prep(gray[ry:ry + rr, rx:rx + rc])
    return s2, s3


def test_stage1_fp16_error_vs_fp64(api, oracle):
    from oracle.net64 import forward64, max_abs_error
    rng = np.random.default_rng(3)
    g = oracle.gray(photo3())
    levels = [(rng.random((120, 140)) * 255).astype(np.float32), oracle.resize(g, 432, 324, 1),
              oracle.resize(g, 1620, 1215, 0), g]
    for lvl in levels:
        e64 = forward64(oracle.models[0], lvl)
        e32 = max_abs_error(api.stage1(lvl, precision=api.FP32), e64)
        e16 = max_abs_error(api.stage1(lvl, precision=api.FP16_BF16), e64)
        print("stage 1 level %s: max |error| vs FP64: FP32 %.3g, FP16 %.3g" % (lvl.shape, e32, e16), flush=True)
        # responses lie in [-1.7159, 1.7159]; FP16 L1 on centred windows measures 0.013
        # at most here (0.023 with the raw 8-bit-scale samples in FP16)
        assert e32 < 2e-5
        assert e16 < 0.02


def test_stage1_fp16_centring_on_smooth_levels(api, oracle):
    # the FP16 L1 rounds each 5x5 window less its centre sample: on a flat level the
    # differences are zero and on a smooth ramp they stay small, so the FP16 stage-1 error
    # there is far below the 8-bit-scale rounding of raw samples (0.02 on photo levels)
    from oracle.net64 import forward64, max_abs_error
    yy, xx = np.mgrid[0:180, 0:240].astype(np.float32)
    levels = {"flat 255": np.full((180, 240), 255.0, np.float32),
              "flat 37": np.full((180, 240), 37.0, np.float32),
              "ramp": (120.0 + 0.5 * xx - 0.25 * yy).astype(np.float32)}
    for name, lvl in levels.items():
        e64 = forward64(oracle.models[0], lvl)
        e16 = max_abs_error(api.stage1(lvl, precision=api.FP16_BF16), e64)
        print("stage 1 %s level: max |error| of FP16 vs FP64 %.3g" % (name, e16), flush=True)
        assert e16 < 2e-3, name


def test_stage23_bf16_fp8_error_vs_fp64(api, oracle):
    from oracle.net64 import forward64
    # the check inputs of real candidates (faces and background windows) and random ones
    f = photo3()
    p = params_for(f, dict(min_face=20))
    _, tr = oracle.detect(f, p, trace_cap=200000)
    g = oracle.gray(f)
    sc = tr["scales"][tr["c_scl"]]
    xs2, xs3 = [], []
    for i in range(0, len(sc), max(1, len(sc) // 48)):
        a, b = check_inputs(oracle, g, sc[i], tr["c_x"][i] * 4, tr["c_y"][i] * 4)
        xs2.append(a)
        xs3.append(b)
    rng = np.random.default_rng(8)
    xs2 += list(np.round(rng.random((8, 51, 47)) * 255).astype(np.float32))
    xs3 += list(np.round(rng.random((8, 51, 47)) * 255).astype(np.float32))
    # responses lie in [-1.7159, 1.7159]; BF16 / FP8 L1 and L2 move the 4x4 maps of real
    # windows by up to about 0.12 / 0.3 (measured), well inside the threshold margins
    bounds = {api.FP32: 2e-5, api.FP16_BF16: 1e-3, api.FP16_FP8: 0.3}
    for stage, xs in ((1, np.stack(xs2)), (2, np.stack(xs3))):
        e64 = np.stack([forward64(oracle.models[stage], x) for x in xs])
        for prec, bound in bounds.items():
            got = api.stage23(stage, xs, precision=prec).astype(np.float64)
            err = np.abs(got - e64).reshape(len(xs), -1).max(1)
            print("stage %d network, precision %d: per-map max |error| vs FP64 over %d maps: max %.3g, median %.3g"
                  % (stage + 1, prec, len(xs), err.max(), np.median(err)), flush=True)
            assert err.max() < bound


@pytest.mark.parametrize("which", ["mixed_batch", "deep_4800x2400"])
def test_reduced_precision_recall_against_fp32(api, which):
    from compactcnncascade_amd import recall, synth
    if which == "mixed_batch":
        frames = [synth.workload_frame(n, i)[0] for n in ("mixed_vga", "mixed_hd", "mixed_fhd", "mixed_uhd")
                  for i in range(8)]
        W, H, mf, sf = 3840, 2160, 20, 1.15
    else:
        frames = [synth.workload_frame("deep_4800x2400_sf105_min15", 0)[0]]
        W, H, mf, sf = 4800, 2400, 15, 1.05
    res = {}
    for prec in (api.FP32, api.FP16_BF16, api.FP16_FP8, api.FP32_BF16, api.FP32_FP8):
        d = api.FaceDetector()
        assert d.Init(api.Param(max_image_size=(W, H), min_face_height=mf, scale_factor=sf)) == 0
        d.set_compute_precision(prec)
        res[prec] = [f for _, f in d.DetectBatch(images=frames, capacity=20000)]
    # a reduced precision moves scores near the thresholds, and a window that crosses one
    # changes which boxes Merger averages: the floors sit below the measured recall
    # the BF16 stages 2/3 keep every FP32 box here (measured 1.0 with either stage 1)
    floors = {api.FP16_BF16: 0.995, api.FP16_FP8: 0.9, api.FP32_BF16: 0.995, api.FP32_FP8: 0.9}
    for prec, floor in floors.items():
        rc, paired, nref, ntest = recall.recall(res[api.FP32], res[prec])
        print("%s precision %d: recall %.4f (%d of %d FP32 detections paired at IoU >= 0.99; %d detections)"
              % (which, prec, rc, paired, nref, ntest), flush=True)
        assert nref > 10
        assert rc >= floor
