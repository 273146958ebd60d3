"""Learned tile skipping (compactcnncascade_amd/csrc/ccnn_skipnet.hip): the training
kernels' gradients against the FP64 oracle (oracle/skipnet64.py), a training run that
learns, and the ablation on held-out frames: the committed network against the plain
cascade it falls back on and that judges its recall."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def api():
    from compactcnncascade_amd import api as A
    A.load()
    return A


def test_skipnet_gradients_match_fp64_oracle(api):
    from compactcnncascade_amd import skipnet
    from oracle import skipnet64 as sn64
    rng = np.random.default_rng(11)
    for shape in [(3, 12, 20), (2, 17, 9)]:
        p = skipnet.init_params(int(rng.integers(1 << 30)))
        x = (rng.random(shape) - 0.5).astype(np.float32)
        y = (rng.random((shape[0], skipnet.SLOTS) + shape[1:]) < 0.15).astype(np.uint8)
        loss, g = skipnet.grad(api, p, x, y, 5.0)
        l64, g64 = sn64.grad64(p.astype(np.float64), x.astype(np.float64), y, 5.0)
        err = np.abs(g.astype(np.float64) - g64)
        print("shape %s: loss %.6f vs %.6f, max |grad error| %.3g (max |grad| %.3g)"
              % (shape, loss, l64, err.max(), np.abs(g64).max()), flush=True)
        assert abs(loss - l64) <= 1e-5 * abs(l64)
        assert (err <= 1e-4 * np.abs(g64).max() + 1e-3 * np.abs(g64)).all()
        z = skipnet.forward(api, p, x)
        z64 = sn64.forward64(p.astype(np.float64), x.astype(np.float64))
        assert np.abs(z - z64).max() <= 1e-5 * (1 + np.abs(z64).max())


def test_skipnet_training_learns(api):
    # labels that are a function of the input (slot s marks cells brighter than a
    # threshold that grows with s): the loss must fall well below its starting value
    from compactcnncascade_amd import skipnet
    rng = np.random.default_rng(5)
    x = (rng.random((64, 16, 24)) - 0.5).astype(np.float32)
    thr = np.linspace(-0.4, 0.4, skipnet.SLOTS)[None, :, None, None]
    y = (x[:, None] > thr).astype(np.uint8)
    p0 = skipnet.init_params(2)
    l0, _ = skipnet.grad(api, p0, x, y, 1.0)
    p, losses = skipnet.train(api, x, y, epochs=40, batch=16, lr=1e-2, pos_weight=1.0, params=p0)
    l1, _ = skipnet.grad(api, p, x, y, 1.0)
    print("loss %.4f -> %.4f, per epoch %s" % (l0, l1, np.round(losses[::8], 4).tolist()), flush=True)
    assert l1 < 0.5 * l0


def _model():
    from compactcnncascade_amd import skipnet
    if not os.path.exists(skipnet.DEFAULT_MODEL):
        pytest.skip("no trained network committed")
    params, meta = skipnet.load()
    assert meta["operating_points"]["recall"] and meta["operating_points"]["fast"]
    return params, meta


@pytest.mark.parametrize("size", [(1920, 1080, 8), (3840, 2160, 4)])
def test_skip_ablation_on_held_out_frames(api, size):
    # held-out seeded frames (seeds from 900000; training used seeds from 500000,
    # validation 800000/810000), faces >= 20, scale 1.15: the plain FP32 batch against the
    # same batch with tile skipping at the network's two operating points
    from compactcnncascade_amd import recall, skipnet, synth
    params, meta = _model()
    W, H, n = size
    frames = [synth.synth_frame(900000 + i, W, H, 10.0, 20, 300)[0] for i in range(n)]
    out, stats = {}, {}
    for key in ("plain", "recall", "fast"):
        d = api.FaceDetector()
        assert d.Init(api.Param(max_image_size=(W, H), min_face_height=20, scale_factor=1.15)) == 0
        if key != "plain":
            skipnet.apply_operating_point(d, params, meta, key)
        out[key] = [f for _, f in d.DetectBatch(images=frames, capacity=4096)]
        stats[key] = d.skip_stats()
    # the recall point's claim (>= 0.995) is measured over 256 held-out frames per size
    # (skipnet evaluate); on these few frames it may lose one box, or 3% of them
    for key in ("recall", "fast"):
        tiles, kept = stats[key]
        rc, paired, nref, ntest = recall.recall(out["plain"], out[key])
        print("%dx%d %s point %s: tiles run %d of %d (%.1f%%), recall %.4f (%d of %d), %d detections"
              % (W, H, key, meta["operating_points"][key], kept, tiles, 100.0 * kept / tiles, rc, paired, nref,
                 ntest), flush=True)
        assert nref > 0 and kept < tiles
        if key == "recall":
            assert nref - paired <= max(1, int(0.03 * nref))
        else:
            assert rc >= 0.85


def test_kept_tiles_are_bit_identical_to_the_plain_cascade(api):
    # skipping only removes work: every response of the last frame of a batch is either
    # the plain cascade's, bit for bit, or -inf (a skipped tile)
    from compactcnncascade_amd import skipnet, synth
    params, meta = _model()
    frames = [synth.synth_frame(900000 + i, 1920, 1080, 10.0, 20, 300)[0] for i in range(3)]
    maps = {}
    for skip in (False, True):
        d = api.FaceDetector()
        assert d.Init(api.Param(max_image_size=(1920, 1080), min_face_height=20, scale_factor=1.15)) == 0
        if skip:
            skipnet.apply_operating_point(d, params, meta, "fast")
        d.DetectBatch(images=frames, capacity=4096)
        maps[skip] = d.last_responses()
    skipped = total = 0
    for a, b in zip(maps[False], maps[True]):
        assert a.shape == b.shape and np.isfinite(a).all()
        off = np.isneginf(b)
        assert np.array_equal(a[~off].view(np.uint32), b[~off].view(np.uint32))
        skipped += int(off.sum())
        total += a.size
    print("responses skipped: %d of %d" % (skipped, total), flush=True)
    assert 0 < skipped < total


def test_skip_is_off_for_other_parameters(api):
    # a network trained for faces >= 20 at scale 1.15 does not apply to a detector with
    # faces >= 40: its batches run plain
    from compactcnncascade_amd import skipnet, synth
    params, meta = _model()
    frames = [synth.workload_frame("fhd_min40_max240", i)[0] for i in range(3)]
    res = []
    for skip in (False, True):
        d = api.FaceDetector()
        assert d.Init(api.Param(max_image_size=(1920, 1080), min_face_height=40, max_face_height=240)) == 0
        if skip:
            skipnet.apply_operating_point(d, params, meta, "fast")
        res.append(d.DetectBatch(images=frames))
        if skip:
            tiles, kept = d.skip_stats()
            assert kept == tiles
    for (na, fa), (nb, fb) in zip(*res):
        assert na == nb and np.array_equal(fa, fb)
