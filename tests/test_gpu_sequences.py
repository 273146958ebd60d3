"""GPU parity over frame sequences on one detector: the detector state that the
reference carries from one Detect call to the next (max_image_size, num_scales,
gray_image_only), and every frame of the benchmark workloads through DetectBatch.

Expected detections: tests/golden/sequences.json, written by the compiled reference
(one CNNDetector per sequence, tests/golden/make_workload_golden.py); boxes and score
bits exact.
"""
import json
import os

import numpy as np
import pytest

from cases import GOLDEN, SEQ_CASES, WORKLOAD_SEQS, api_param_of, frame_digest, seq_params, workload_frames

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def api():
    from compactcnncascade_amd import api as A
    A.load()
    return A


@pytest.fixture(scope="module")
def seqgold():
    with open(os.path.join(GOLDEN, "sequences.json")) as f:
        return json.load(f)["cases"]


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def check_frame(faces, g, i, tag):
    exp = g["faces"][i]
    got = [[int(r["x"]), int(r["y"]), int(r["width"]), int(r["height"])] for r in faces]
    assert got == [e[:4] for e in exp], (tag, i)
    assert [int(v) for v in bits(faces["score"])] == g["score_bits"][i], (tag, i)


@pytest.mark.parametrize("name", list(SEQ_CASES))
def test_sequence_on_one_detector_matches_reference(api, seqgold, name):
    builders, over, mx = SEQ_CASES[name][:3]
    sets = SEQ_CASES[name][3] if len(SEQ_CASES[name]) > 3 else [0] * len(builders)
    frames = [b() for b in builders]
    g = seqgold[name]
    assert [frame_digest(f) for f in frames] == g["digests"]
    param = api_param_of(seq_params(over, mx))
    print("sequence", name, [f.shape for f in frames], flush=True)
    d = api.FaceDetector()
    assert d.Init(param) == 0
    for i, f in enumerate(frames):
        if sets[i] > 0:
            d.setMinObjectHeight(sets[i])
        n, faces = d.Detect(f)
        assert n == len(g["faces"][i]), (name, i)
        check_frame(faces, g, i, "detect")
    if any(sets):
        return
    d2 = api.FaceDetector()
    assert d2.Init(param) == 0
    for i, (n, faces) in enumerate(d2.DetectBatch(images=frames)):
        assert n == len(g["faces"][i]), (name, i)
        check_frame(faces, g, i, "batch")


@pytest.fixture(scope="module")
def workloads():
    return {name: workload_frames(spec) for name, (spec, _, _) in WORKLOAD_SEQS.items()}


@pytest.mark.parametrize("mode", [1, 0])
@pytest.mark.parametrize("name", list(WORKLOAD_SEQS))
def test_workload_batch_matches_reference(api, seqgold, workloads, name, mode):
    # every frame of the benchmark workload (64 Full-HD frames; the 32-frame mixed-size
    # batch), in the benchmark's order, through one DetectBatch call: the graph batch
    # (mode 1) and the per-frame pipelined batch (mode 0)
    spec, over, mx = WORKLOAD_SEQS[name]
    frames = workloads[name]
    g = seqgold[name]
    assert [frame_digest(f) for f in frames] == g["digests"]
    print("workload", name, len(frames), "mode", mode, flush=True)
    d = api.FaceDetector()
    assert d.Init(api_param_of(seq_params(over, mx))) == 0
    d.set_batch_mode(mode)
    api.last_error(reset=True)
    out = d.DetectBatch(images=frames, capacity=4096)
    assert api.last_error() == 0
    if mode == 1:
        launches, captures = d.batch_stats()
        print("graph launches", launches, "captures", captures, flush=True)
        assert captures >= 1 and launches >= 6
    total = 0
    for i, (n, faces) in enumerate(out):
        assert n == len(g["faces"][i]), (name, i)
        check_frame(faces, g, i, "batch")
        total += n
    assert total > len(frames) // 2


def test_single_field_setter_keeps_gray_image_only_off(api, oracle):
    # Init -> Detect(colour frame, min face >= 34: gray_image_only goes off for good)
    # -> setMinObjectHeight -> Detect(1-channel frame): the reference's setter only
    # Clears and changes the field (cnn_detector_v3.h:417-436), so the gray frame is not
    # pre-resized; setParam forces gray_image_only on again (CompactCNNLibAPI_v2.cpp:131)
    from cases import vga
    colour = vga(11, smin=20)
    gray = np.ascontiguousarray(colour[:, :, 1])
    p = seq_params(dict(min_face=40), (640, 480))
    st = oracle.new_state(p)
    exp1 = oracle.detect_state(colour, p, st)
    assert st.gray_only == 0
    p2 = dict(p, min_face=36)
    st.num_scales = -1                 # the setter's Clear; the next Detect Inits
    exp2 = oracle.detect_state(gray, p2, st)
    d = api.FaceDetector()
    assert d.Init(api_param_of(p)) == 0
    n1, f1 = d.Detect(colour)
    assert n1 == len(exp1) and [int(v) for v in bits(f1["score"])] == [int(v) for v in bits(exp1[:, 4])]
    d.setMinObjectHeight(36)
    n2, f2 = d.Detect(gray)
    assert n2 == len(exp2)
    assert [[int(r["x"]), int(r["y"]), int(r["width"]), int(r["height"])] for r in f2] == \
        exp2[:, :4].astype(int).tolist()
    assert [int(v) for v in bits(f2["score"])] == [int(v) for v in bits(exp2[:, 4])]
    # a fresh detector pre-resizes the gray frame: a different computation
    fresh = oracle.detect(gray, p2)[0]
    assert not (len(fresh) == len(exp2) and np.array_equal(fresh, exp2))
    # setParam restores gray_image_only: the gray frame is pre-resized again
    d.setParam(api_param_of(p2))
    n3, f3 = d.Detect(gray)
    assert n3 == len(fresh) and [int(v) for v in bits(f3["score"])] == [int(v) for v in bits(fresh[:, 4])]


def test_batch_graph_is_replayed_and_matches_per_frame_detect(api):
    # the same batch twice: one capture, the same detections; and the same as Detect calls
    # in turn on one detector, for BGR, BGRA and ragged frames of several sizes
    from cases import _bgra, vga
    from compactcnncascade_amd import synth
    frames = [vga(0), _bgra(vga(3)), synth.synth_frame(21, 333, 211, 4, 30, 120)[0], vga(5),
              synth.synth_frame(7, 1280, 720, 8, 20, 200)[0]]
    p = api_param_of(seq_params(dict(min_face=30), (1280, 720)))
    d = api.FaceDetector()
    assert d.Init(p) == 0
    a = d.DetectBatch(images=frames)
    _, c1 = d.batch_stats()
    b = d.DetectBatch(images=frames)
    _, c2 = d.batch_stats()
    assert c2 == c1
    ref = api.FaceDetector()
    assert ref.Init(p) == 0
    singles = [ref.Detect(f) for f in frames]
    for (na, fa), (nb, fb), (ns, fs) in zip(a, b, singles):
        assert na == nb == ns and np.array_equal(fa, fb) and np.array_equal(fa, fs)
    assert sum(n for n, _ in a) > 0


def test_batch_graph_pre_resized_gray_frames(api, oracle):
    # 1-channel frames with min face >= 34 are pre-resized inside the batch graph
    # (gray_image_only); each frame as the oracle's detector in turn
    from cases import _gray, vga
    frames = [_gray(vga(s, smin=30)) for s in (11, 3, 0)]
    p = seq_params(dict(min_face=40), (640, 480))
    exp = oracle.detect_seq(frames, p)
    d = api.FaceDetector()
    assert d.Init(api_param_of(p)) == 0
    out = d.DetectBatch(images=frames)
    for (n, faces), e in zip(out, exp):
        assert n == len(e)
        assert [[int(r["x"]), int(r["y"]), int(r["width"]), int(r["height"])] for r in faces] == \
            e[:, :4].astype(int).tolist()
        assert [int(v) for v in bits(faces["score"])] == [int(v) for v in bits(e[:, 4])]


def test_batch_candidate_overflow_takes_the_per_frame_path(api):
    # detect_mode disable with t1 = -10 makes every window a candidate: 217,273 per
    # 1080p frame, more than the batch's candidate buffers hold (65,536 at this size)
    from compactcnncascade_amd import synth
    frames = [synth.workload_frame("fhd_min40_max240", i)[0] for i in range(2)]
    p = seq_params(dict(min_face=40, max_face=240, detect_mode=0, t1=-10.0), (1920, 1080))
    d = api.FaceDetector()
    assert d.Init(api_param_of(p)) == 0
    out = d.DetectBatch(images=frames, capacity=300000)
    ref = api.FaceDetector()
    assert ref.Init(api_param_of(p)) == 0
    for (n, faces), f in zip(out, frames):
        ns, fs = ref.Detect(f)
        assert n == ns > 100000 and np.array_equal(faces, fs)


def test_every_tile_shape_gives_the_same_detections(api, seqgold, workloads):
    # the stage-1 kernels' tile shapes change only how the work is split: every shape gives
    # the reference's detections, on the graph batch and the per-frame path, and the same
    # response maps bit for bit
    name = "fhd_min40_max240_x64"
    spec, over, mx = WORKLOAD_SEQS[name]
    frames = workloads[name][:4]
    g = seqgold[name]
    p = api_param_of(seq_params(over, mx))
    base = None
    for shape in range(12):
        d = api.FaceDetector()
        assert d.Init(p) == 0
        d.set_tile_shape(shape)
        print("tile shape", shape, d.get_tile_shape(), flush=True)
        for i, (n, faces) in enumerate(d.DetectBatch(images=frames)):
            assert n == len(g["faces"][i]), (shape, i)
            check_frame(faces, g, i, "shape %d batch" % shape)
        n, faces = d.Detect(frames[0])
        check_frame(faces, g, 0, "shape %d detect" % shape)
        maps = d.last_responses()
        if base is None:
            base = maps
        else:
            assert all(np.array_equal(bits(a), bits(b)) for a, b in zip(maps, base)), shape


def test_autotune_picks_a_shape_and_keeps_results(api, seqgold, workloads):
    name = "mixed_batch_x32"
    spec, over, mx = WORKLOAD_SEQS[name]
    frames = workloads[name][::4]            # two frames of each size
    g = seqgold[name]
    d = api.FaceDetector()
    assert d.Init(api_param_of(seq_params(over, mx))) == 0
    best, sms, gms = d.autotune(images=frames, reps=1)
    print("autotune", best, sms.tolist(), gms.tolist(), flush=True)
    assert 0 <= best < 12 and np.isfinite(sms).all() and np.isfinite(gms).all()
    assert d.get_tile_shape()[0] == best
    d2 = api.FaceDetector()
    assert d2.Init(api_param_of(seq_params(over, mx))) == 0
    d2.set_tile_shape(best)
    full = workloads[name]
    for i, (n, faces) in enumerate(d2.DetectBatch(images=full)):
        assert n == len(g["faces"][i]), i
        check_frame(faces, g, i, "tuned")


def test_check_grid_gives_the_same_detections(api, seqgold, workloads):
    # the check kernel's grid changes only which workgroup evaluates which (candidate, crop
    # variant): one workgroup per item (per-frame default), persistent workgroups taking
    # several items each from a queue (per-frame 1 or 3 per CU; batch 1 or 8 per CU, 8 being
    # more than fit on a CU at once); the 32 mixed-size frames (640x480 .. 3840x2160)
    # through the graph batch and the per-frame pipelined batch, in the exact precision
    # (against the reference's detections) and a matrix-core one (against itself)
    name = "mixed_batch_x32"
    spec, over, mx = WORKLOAD_SEQS[name]
    frames = workloads[name]
    g = seqgold[name]
    p = api_param_of(seq_params(over, mx))
    ref16 = {}
    for frame_per_cu, batch_per_cu in ((0, 4), (1, 1), (3, 8)):
        for prec in (api.FP32, api.FP16_BF16):
            for mode in (1, 0):
                d = api.FaceDetector()
                assert d.Init(p) == 0
                d.set_compute_precision(prec)
                d.set_check_grid(frame_per_cu, batch_per_cu)
                d.set_batch_mode(mode)
                tag = "grid %d/%d precision %d mode %d" % (frame_per_cu, batch_per_cu, prec, mode)
                print(tag, flush=True)
                out = d.DetectBatch(images=frames)
                if prec == api.FP32:
                    for i, (n, faces) in enumerate(out):
                        assert n == len(g["faces"][i]), (tag, i)
                        check_frame(faces, g, i, tag)
                else:
                    got = [faces for _, faces in out]
                    if mode not in ref16:
                        ref16[mode] = got
                    assert all(np.array_equal(a, b) for a, b in zip(got, ref16[mode])), tag
