"""GPU parity: every stage of the HIP path, called through the C ABI, against the
oracle (oracle/ccnn_oracle.c) and the golden fixtures of the compiled reference.

Bar: bit-exact everywhere (gray, pyramid, response maps, crops, stage-2/3 responses,
final boxes and scores), since the HIP kernels evaluate the reference's operation order
without contraction.
"""
import json
import os

import numpy as np
import pytest

from cases import CASES, GOLDEN, ORACLE_ONLY_CASES, api_param, frame_digest, params_for

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def api():
    from compactcnncascade_amd import api as A
    A.load()
    return A


@pytest.fixture(scope="module")
def golden():
    with open(os.path.join(GOLDEN, "detections.json")) as f:
        return json.load(f)["cases"]


@pytest.fixture(scope="module")
def vectors():
    return dict(np.load(os.path.join(GOLDEN, "vectors.npz")))


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def test_gray_bitexact(api, oracle):
    rng = np.random.default_rng(1)
    # 4-pixel-vector path (4-byte aligned rows; 16-byte for BGRA) with full and partial
    # last quads, and the byte path for other row steps
    for shape in [(61, 77, 3), (33, 45, 4), (29, 31), (1080, 1920, 3), (37, 84, 3), (37, 83, 3), (21, 80, 4),
                  (21, 81, 4), (21, 84, 4), (19, 64, 1), (19, 66, 1), (5, 3, 3), (7, 16, 3), (9, 12, 4)]:
        img = rng.integers(0, 256, size=shape, dtype=np.uint8)
        print("gray", shape, flush=True)
        assert np.array_equal(bits(api.stage_gray(img)), bits(oracle.gray(img)))
    # a padded row step (a view into a wider frame)
    wide = rng.integers(0, 256, size=(40, 100, 3), dtype=np.uint8)
    wide16 = rng.integers(0, 256, size=(40, 160, 3), dtype=np.uint8)    # 480-byte rows
    for sub in [wide[:, :70], wide[3:, 5:90], wide16[:, :70], wide16[2:, 16:101], wide16[:, 32:157]]:
        assert np.array_equal(bits(api.stage_gray(sub)), bits(oracle.gray(np.ascontiguousarray(sub))))


def test_gray_all_bgr_triples(api, oracle):
    # every (B, G, R) triple: 4096 x 4096 image of 2^24 pixels
    v = np.arange(1 << 24, dtype=np.uint32)
    img = np.stack([(v & 255), (v >> 8) & 255, (v >> 16) & 255], axis=1).astype(np.uint8).reshape(4096, 4096, 3)
    print("gray 2^24 triples", flush=True)
    assert np.array_equal(bits(api.stage_gray(img)), bits(oracle.gray(img)))


def test_pyramid_levels_bitexact(api, oracle):
    from cases import photo3
    g = oracle.gray(photo3())
    gh, gw = g.shape
    for (lw, lh, mode) in [(gw, gh, 2), (int(gw * 0.8), int(gh * 0.8), 0), (int(gw * 1.35), int(gh * 1.35), 0),
                           (int(gw * 0.5), int(gh * 0.5), 1), (int(gw * 0.3), int(gh * 0.3), 1), (49, 51, 0),
                           (47, 51, 1)]:
        print("level", lw, lh, mode, flush=True)
        a = api.stage_level(g, lw, lh, mode)
        b = oracle.resize(g, lw, lh, mode) if mode != 2 else g
        assert np.array_equal(bits(a), bits(b)), (lw, lh, mode)


def test_stage1_bitexact(api, oracle):
    from cases import photo3
    rng = np.random.default_rng(5)
    g = oracle.gray(photo3())
    levels = [(rng.random((51, 47)) * 255).astype(np.float32),
              (rng.random((200, 333)) * 255).astype(np.float32),
              oracle.resize(g, 432, 324, 1),
              g]
    for lvl in levels:
        print("stage1", lvl.shape, flush=True)
        a = api.stage1(lvl)
        b = oracle.forward(0, lvl)
        assert a.shape == b.shape
        assert np.array_equal(bits(a), bits(b))


RESPONSE_CASES = {
    "vga_s0_min40": CASES["vga_s0_min40"],
    "fhd_w0_min40_max240": CASES["fhd_w0_min40_max240"],
    "photo3_min20": CASES["photo3_min20"],
    # min face 27: the first scale is exactly 1, a copy level
    "photo3_min27": (CASES["photo3_min20"][0], dict(min_face=27)),
}


@pytest.mark.parametrize("name", list(RESPONSE_CASES))
def test_detect_response_maps_bitexact(api, oracle, name):
    # the response map of every level as a real Detect computes it: the fused sampler of
    # k_stage1_l1 (copy, nearest-neighbour and bilinear levels) feeding both stage-1
    # kernels, against the oracle's resize + forward of the same level
    build, over = RESPONSE_CASES[name]
    f = build()
    d = api.FaceDetector()
    assert d.Init(api_param(f, over)) == 0
    d.Detect(f)
    levels = d.last_trace()["levels"]
    maps = d.last_responses()
    g = oracle.gray(f)
    modes = set()
    for (scl, lw, lh, rw, rh, mode), m in zip(levels, maps):
        print(name, "level", scl, lw, lh, mode, flush=True)
        lvl = g if mode == 2 else oracle.resize(g, int(lw), int(lh), int(mode))
        exp = oracle.forward(0, lvl)
        assert m.shape == exp.shape == (rh, rw)
        assert np.array_equal(bits(m), bits(exp)), (name, scl, mode)
        modes.add(int(mode))
    assert len(levels) > 5
    if name == "photo3_min20":
        assert modes == {0, 1}
    if name == "photo3_min27":
        assert modes == {0, 1, 2}


def test_stage23_maps_bitexact(api, oracle, vectors):
    # the full 4x4 response maps of the stage-2 and stage-3 networks: the reference's
    # vectors and random inputs (u8-valued, as the check feeds them) against the oracle
    rng = np.random.default_rng(77)
    x = np.concatenate([vectors["s2_in"][None], np.round(rng.random((15, 51, 47)) * 255).astype(np.float32)])
    for stage in (1, 2):
        got = api.stage23(stage, x)
        for i in range(len(x)):
            exp = oracle.forward(stage, x[i])
            assert np.array_equal(bits(got[i]), bits(exp)), (stage, i)
    assert np.array_equal(bits(api.stage23(1, vectors["s2_in"][None])[0]), bits(vectors["s2_out"]))
    assert np.array_equal(bits(api.stage23(2, vectors["s2_in"][None])[0]), bits(vectors["s3_out"]))


def test_stage1_shift_invariance(api):
    # moving the level by 4 pixels moves the response map by one cell
    rng = np.random.default_rng(9)
    lvl = (rng.random((120, 140)) * 255).astype(np.float32)
    a = api.stage1(lvl)
    b = api.stage1(lvl[4:, 4:])
    assert np.array_equal(bits(a[1:, 1:])[:b.shape[0], :b.shape[1]], bits(b)[:a.shape[0] - 1, :a.shape[1] - 1])


def test_check_stage_bitexact(api, oracle):
    from cases import vga
    f = vga(3)
    over = dict(min_face=40)
    p = params_for(f, over)
    _, tr = oracle.detect(f, p, trace_cap=100000)
    scales = tr["scales"]
    g = oracle.gray(f)
    sc = scales[tr["c_scl"]]
    res = api.stage_check(api_param(f, over), g, sc, tr["c_x"], tr["c_y"])
    n = 0
    for i in range(len(sc)):
        r = oracle.check_candidate(p, g, sc[i], tr["c_x"][i] * 4, tr["c_y"][i] * 4)
        assert int(res[i, 0]) == int(r[0]) and int(res[i, 1]) == int(r[1]), i
        assert int(res[i, 2]) == int(r[2]) and int(res[i, 3]) == int(r[3]), i
        assert bits(res[i, 4:6]).tolist() == bits(r[4:6]).tolist(), i
        n += 1
    assert n > 10


@pytest.mark.parametrize("name", list(CASES) + list(ORACLE_ONLY_CASES))
def test_detect_matches_reference(api, golden, name):
    build, over = {**CASES, **ORACLE_ONLY_CASES}[name]
    f = build()
    g = golden[name]
    assert frame_digest(f) == g["digest"], "input frame differs from the fixture's"
    print("detect", name, f.shape, flush=True)
    d = api.FaceDetector()
    api.last_error(reset=True)
    assert d.Init(api_param(f, over)) == 0
    n, faces = d.Detect(f)
    assert api.last_error() == 0
    exp = g["faces"]
    assert n == len(exp)
    got = [[int(r["x"]), int(r["y"]), int(r["width"]), int(r["height"])] for r in faces]
    assert got == [e[:4] for e in exp]
    assert [int(v) for v in bits(faces["score"])] == g["score_bits"]
    # facial attributes (-1 without facial analysis)
    assert [[int(r["gender"]), int(r["smile"]), int(r["glasses"])] for r in faces] == [e[5:8] for e in exp]


def random_case(seed):
    """A seeded random frame (size, channels, faces) and parameter set, over the
    parameters FaceDetector::Param exposes."""
    from cases import _bgra, _gray
    from compactcnncascade_amd import synth
    rng = np.random.default_rng(seed)
    w, h = int(rng.integers(64, 561)), int(rng.integers(64, 421))
    f = synth.synth_frame(1000 + seed, w, h, float(rng.uniform(2, 8)), 20, 160)[0]
    ch = int(rng.choice([1, 3, 3, 4]))
    f = _gray(f) if ch == 1 else _bgra(f) if ch == 4 else f
    min_face = int(rng.integers(20, 61))
    over = dict(min_face=min_face, max_face=int(rng.choice([0, min_face + int(rng.integers(10, 150))])),
                scale_factor=float(rng.choice([1.05, 1.1, 1.15, 1.2, 1.3])), precision=int(rng.integers(0, 5)),
                min_neighbors=int(rng.integers(1, 4)), equalize=int(rng.integers(0, 2)),
                reflection=int(rng.integers(0, 2)), drop_detect=int(rng.integers(0, 2)),
                detect_mode=int(rng.choice([0, 1, 2, 2])), t1=float(rng.uniform(-2.0, -1.0)),
                t2=float(rng.uniform(-1.5, -0.75)), t3=float(rng.uniform(-1.25, -0.5)),
                facial_analysis=int(rng.integers(0, 4) == 0))
    return f, over


# seeds of the randomized differential test; CCNN_RANDOM_CASES widens the sweep
RANDOM_CASES = int(os.environ.get("CCNN_RANDOM_CASES", "16"))


@pytest.mark.parametrize("seed", range(RANDOM_CASES))
def test_random_frames_and_parameters_match_oracle(api, oracle, seed):
    # differential test over random frame sizes, channel counts and parameter sets: the
    # GPU path against the oracle computed on the CPU for the same inputs (boxes and
    # score bits exact, facial attributes equal)
    f, over = random_case(seed)
    print("random case", seed, f.shape, over, flush=True)
    exp, _ = oracle.detect(f, params_for(f, over))
    d = api.FaceDetector()
    assert d.Init(api_param(f, over)) == 0
    n, faces = d.Detect(f)
    assert n == len(exp)
    got = [[int(r["x"]), int(r["y"]), int(r["width"]), int(r["height"])] for r in faces]
    assert got == [[int(v) for v in e[:4]] for e in exp]
    assert bits(faces["score"]).tolist() == bits(np.asarray(exp[:, 4], np.float32)).tolist()
    assert [[int(r["gender"]), int(r["smile"]), int(r["glasses"])] for r in faces] == \
        [[int(v) for v in e[5:8]] for e in exp]
    # the same frame twice through DetectBatch, as one graph launch and pipelined per frame;
    # in a reduced compute precision too, against Detect in that precision (the same
    # kernels on both paths)
    for prec in (api.FP32, api.FP16_BF16):
        if prec != api.FP32:
            d = api.FaceDetector()
            assert d.Init(api_param(f, over)) == 0
            d.set_compute_precision(prec)
            n, faces = d.Detect(f)
        for mode in (1, 0):
            d = api.FaceDetector()
            assert d.Init(api_param(f, over)) == 0
            d.set_compute_precision(prec)
            d.set_batch_mode(mode)
            for nb, fb in d.DetectBatch(images=[f, f], capacity=4096):
                # the count is always reported; the records up to the capacity
                assert nb == n and len(fb) == min(n, 4096), (seed, prec, mode)
                assert np.array_equal(fb, faces[:len(fb)]), (seed, prec, mode)


def test_trace_matches_oracle(api, oracle):
    from cases import photo3
    f = photo3()
    over = dict(min_face=20)
    p = params_for(f, over)
    _, tr = oracle.detect(f, p, trace_cap=200000)
    d = api.FaceDetector()
    assert d.Init(api_param(f, over)) == 0
    d.Detect(f)
    t = d.last_trace()
    # levels in compute order
    assert t["levels"][:, 0].tolist() == tr["lvl_scl"].tolist()
    assert t["levels"][:, 1:5].tolist() == np.stack([tr["lvl_w"], tr["lvl_h"], tr["lvl_rw"], tr["lvl_rh"]], 1).tolist()
    # candidate windows: same list, same order, same stage-1 scores
    cand = t["cand"]
    assert len(cand) == len(tr["c_x"])
    assert cand[:, 1].astype(int).tolist() == tr["c_x"].tolist()
    assert cand[:, 2].astype(int).tolist() == tr["c_y"].tolist()
    assert bits(cand[:, 3]).tolist() == bits(tr["c_score"]).tolist()
    # check outcome of every candidate the reference evaluated
    ev = np.isin(tr["c_status"], [2, 3])
    res = t["res"][ev]
    assert (res[:, 0].astype(int) == (tr["c_status"][ev] == 2)).all()
    assert res[:, 1].astype(int).tolist() == tr["c_mod"][ev].tolist()
    assert res[:, 2].astype(int).tolist() == tr["c_k1"][ev].tolist()
    assert bits(res[:, 4]).tolist() == bits(tr["c_m1"][ev]).tolist()


def test_exact_reciprocal_matches_ieee_division_on_every_float(api):
    # the hidden/output layers' rational tanh (cnnpp_cplusplus.cpp:662-666) divides by a
    # denominator >= 1; the kernels take 1/den from the hardware estimate and one fused
    # Newton step (0 beyond 2^25, where 1 - 1/d rounds to 1). Exhaustive over all 2^32
    # bit patterns: the tanh is bit-identical to the one with IEEE division, and so is the
    # reciprocal itself for every d in [1, 2^25).
    import ctypes
    L = api.load()
    L.ccnn_selftest_tanh.argtypes = [ctypes.c_uint, ctypes.c_uint, ctypes.POINTER(ctypes.c_ulonglong)]
    out = (ctypes.c_ulonglong * 3)()
    assert L.ccnn_selftest_tanh(0, 0xFFFFFFFF, out) == 0
    assert (out[0], out[1]) == (0, 0), "mismatches: tanh %d, reciprocal %d, last pattern %#x" % tuple(out)


def test_landmarks_network_bitexact(api, oracle, vectors):
    rng = np.random.default_rng(12)
    inputs = list(vectors["fa_in"]) + [(rng.random((51, 49)) * 255).astype(np.float32) for _ in range(4)]
    for x in inputs:
        a = api.stage_landmarks(x)
        b = oracle.forward(3, x)[:, 0, 0]
        assert np.array_equal(bits(a), bits(b))
    for x, y in zip(vectors["fa_in"], vectors["fa_out"]):
        assert np.array_equal(bits(api.stage_landmarks(x)), bits(y))


def test_facial_analysis_per_candidate_matches_oracle(api, oracle):
    # every window the checks accept: the landmarks' verdict and the attributes
    from cases import photo3
    f = photo3()
    over = dict(min_face=20, facial_analysis=1)
    p = params_for(f, over)
    _, tr = oracle.detect(f, p, trace_cap=200000)
    d = api.FaceDetector()
    assert d.Init(api_param(f, over)) == 0
    d.Detect(f)
    t = d.last_trace()
    fa = d.last_facial()
    assert len(fa) == len(tr["c_x"])
    ev = np.isin(tr["c_status"], [2, 4])          # accepted by the checks
    assert (t["res"][ev, 0] == 1).all()
    assert (fa[ev, 0] == (tr["c_status"][ev] == 2)).all()
    assert ev.sum() > 5 and (tr["c_status"] == 4).sum() > 0


def test_facial_analysis_on_a_crowded_frame_matches_oracle(api, oracle):
    # more candidates than one facial workgroup decides at a time (512) and more accepted
    # windows than facial workgroups (64): every window analysed once, verdicts and the
    # faces' attributes as the oracle's
    from compactcnncascade_amd import synth
    f = synth.synth_frame(5, 1280, 720, 80.0, 24, 90)[0]
    over = dict(min_face=20, facial_analysis=1)
    exp, tr = oracle.detect(f, params_for(f, over), trace_cap=200000)
    ev = np.isin(tr["c_status"], [2, 4])
    assert len(ev) > 512 and ev.sum() > 64
    d = api.FaceDetector()
    assert d.Init(api_param(f, over)) == 0
    n, faces = d.Detect(f)
    t = d.last_trace()
    fa = d.last_facial()
    assert len(fa) == len(ev)
    assert np.array_equal(t["res"][:, 0] == 1, ev)
    assert np.array_equal(fa[ev, 0] == 1, tr["c_status"][ev] == 2)
    assert n == len(exp)
    got = [[int(r["x"]), int(r["y"]), int(r["width"]), int(r["height"]), int(r["gender"]), int(r["smile"]),
            int(r["glasses"])] for r in faces]
    assert got == [[int(v) for v in e[:4]] + [int(v) for v in e[5:8]] for e in exp]
    assert bits(faces["score"]).tolist() == bits(np.asarray(exp[:, 4], np.float32)).tolist()


def test_repeated_detect_is_stable(api):
    from cases import vga
    f = vga(0)
    d = api.FaceDetector()
    assert d.Init(api_param(f, dict(min_face=40))) == 0
    first = d.Detect(f)[1]
    for _ in range(3):
        again = d.Detect(f)[1]
        assert np.array_equal(first, again)


def test_detect_capacity_truncates_but_reports_count(api):
    # ccnn_detect writes min(count, capacity) faces and returns the full count; slots past
    # the capacity stay untouched
    import ctypes
    from cases import vga
    f = vga(11, smin=20)
    d = api.FaceDetector()
    assert d.Init(api_param(f, dict(min_face=20))) == 0
    n, full = d.Detect(f)
    assert n >= 2
    img, ci = api._image(f)
    L = api.load()
    for cap in (0, 1, n - 1):
        out = np.zeros(n + 1, api.FACE_DTYPE)
        out["x"] = -7
        m = L.ccnn_detect(d._h, api._ptr(out), cap, ctypes.byref(ci))
        assert m == n
        assert np.array_equal(out[:cap], full[:cap])
        assert (out["x"][cap:] == -7).all()


def test_detect_on_row_strided_view(api):
    # ImageData with a row step wider than cols * channels (a crop of a bigger frame)
    from cases import vga
    big = np.zeros((520, 700, 3), np.uint8)
    big[17:497, 31:671] = vga(3)
    view = big[17:497, 31:671]
    assert view.strides[0] == 2100
    d = api.FaceDetector()
    assert d.Init(api_param(vga(3), dict(min_face=40))) == 0
    a = d.Detect(view)
    b = d.Detect(np.ascontiguousarray(view))
    assert a[0] == b[0] and a[0] > 0 and np.array_equal(a[1], b[1])


def test_batch_matches_single_frame_results(api, golden):
    # pipelined batch over frames of different sizes and channel counts: every frame's
    # result equals its single-call golden result
    names = ["vga_s0_min40", "vga_s3_min40", "photo3_min40_max200", "vga_s5_min40", "ragged_333x211",
             "vga_s2_min40"]
    frames = [CASES[n][0]() for n in names]
    d = api.FaceDetector()
    api.last_error(reset=True)
    # one parameter set for the batch: the largest frame bounds max_image_size; the
    # golden of each case was produced with max_image_size = its own size, which only
    # matters through min_scale when max_face is 0, so compare cases with equal settings
    for n, f in zip(names, frames):
        assert d.Init(api_param(f, CASES[n][1])) == 0
        out = d.DetectBatch(images=[f, f, f])
        assert api.last_error() == 0
        for cnt, faces in out:
            exp = golden[n]["faces"]
            assert cnt == len(exp)
            assert [[int(r["x"]), int(r["y"]), int(r["width"]), int(r["height"])] for r in faces] == \
                [e[:4] for e in exp]
    # frames of several sizes through one detector in one batch: the same as Detect calls
    # in turn on one detector (whose max_image_size and num_scales carry from frame to frame)
    d2 = api.FaceDetector()
    assert d2.Init(api_param(frames[0], dict(min_face=40, max_face=200))) == 0
    batch = d2.DetectBatch(images=frames)
    d3 = api.FaceDetector()
    assert d3.Init(api_param(frames[0], dict(min_face=40, max_face=200))) == 0
    singles = [d3.Detect(f) for f in frames]
    for (cb, fb), (cs, fs) in zip(batch, singles):
        assert cb == cs and np.array_equal(fb, fs)
    # the batch gray kernel's 16-pixel groups: BGRA rows of 16-byte multiples whose width
    # ends in a partial group, and a BGR view with a padded (16-byte multiple) row step,
    # against one Detect at a time (whose gray kernel takes 4 pixels per thread)
    rng = np.random.default_rng(7)
    bgra = np.ascontiguousarray(np.concatenate([frames[0][:300, :292], rng.integers(0, 256, (300, 292, 1),
                                                                                     dtype=np.uint8)], axis=2))
    wide = np.zeros((360, 512, 3), np.uint8)
    wide[:, :470] = frames[0][:360, :470]
    view = wide[:, :470]
    extra = [bgra, view, bgra]
    d4 = api.FaceDetector()
    assert d4.Init(api_param(frames[0], dict(min_face=40, max_face=200))) == 0
    batch = d4.DetectBatch(images=extra)
    d5 = api.FaceDetector()
    assert d5.Init(api_param(frames[0], dict(min_face=40, max_face=200))) == 0
    singles = [d5.Detect(f) for f in extra]
    for (cb, fb), (cs, fs) in zip(batch, singles):
        assert cb == cs and np.array_equal(fb, fs)


def test_device_frames_match_host_frames(api):
    # frames already in GPU memory (ccnn_detect_device, ccnn_detect_batch with device = 1)
    # give the same results as the same frames from host memory: BGR, BGRA and gray
    # frames, a blank frame (no candidates), one below the 47x51 minimum (no faces) and a
    # device view whose row step exceeds cols * channels; one Detect at a time, and whole
    # batches in both batch modes; empty batches return nothing
    import torch
    from cases import vga
    rng = np.random.default_rng(11)
    f0, f1 = vga(0), vga(4)
    frames = [f0, np.ascontiguousarray(np.concatenate([f1, rng.integers(0, 256, f1.shape[:2] + (1,),
                                                                        dtype=np.uint8)], axis=2)),
              np.ascontiguousarray(f0[:, :, 1]), np.zeros((300, 400, 3), np.uint8),
              np.ascontiguousarray(f1[:40, :44]), vga(2)]
    dev = [torch.from_numpy(f).to("cuda:0") for f in frames]
    descs = [(t.data_ptr(), f.shape[1], f.shape[0], 1 if f.ndim == 2 else f.shape[2], f.strides[0])
             for t, f in zip(dev, frames)]
    # a device view: the last frame inside a wider buffer, rows 2,048 bytes apart
    f = frames[-1]
    wide = torch.zeros((f.shape[0], 2048), dtype=torch.uint8, device="cuda:0")
    wide[:, :f.shape[1] * 3] = torch.from_numpy(np.ascontiguousarray(f).reshape(f.shape[0], -1)).to("cuda:0")
    frames.append(f)
    descs.append((wide.data_ptr(), f.shape[1], f.shape[0], 3, 2048))
    torch.cuda.synchronize()

    def fresh():
        d = api.FaceDetector()
        assert d.Init(api_param(frames[0], dict(min_face=40, max_face=200))) == 0
        return d

    d_host, d_dev = fresh(), fresh()
    host = [d_host.Detect(f) for f in frames]
    for (p, w, h, c, st), (nh, fh) in zip(descs, host):
        nd, fd = d_dev.DetectDevice(p, w, h, c, st)
        assert nd == nh and np.array_equal(fd, fh), (w, h, c)
    assert host[3][0] == 0 and host[4][0] == 0 and host[0][0] > 0
    for mode in (1, 0):
        d_host, d_dev = fresh(), fresh()
        d_host.set_batch_mode(mode)
        d_dev.set_batch_mode(mode)
        bh = d_host.DetectBatch(images=frames)
        bd = d_dev.DetectBatch(device_frames=descs)
        assert len(bh) == len(bd) == len(frames)
        for (nh, fh), (nd, fd) in zip(bh, bd):
            assert nd == nh and np.array_equal(fd, fh), mode
        assert d_host.DetectBatch(images=[]) == [] and d_dev.DetectBatch(device_frames=[]) == []


def test_parameter_accessors_and_profiling(api):
    # FaceDetector's getters and setters (CompactCNNLib.h:172-200): each setter gives the
    # detections of a detector initialised with the changed parameter; profiling
    # (per-kernel HIP events, the batch then runs without its graph) changes no result
    from cases import vga
    frames = [vga(0), vga(3), vga(5)]

    def fresh(**kw):
        d = api.FaceDetector()
        assert d.Init(api_param(frames[0], dict(dict(min_face=40, max_face=200), **kw))) == 0
        return d

    def same(a, b):
        return a[0] == b[0] and np.array_equal(a[1], b[1])

    d = fresh()
    p = d.getParam()
    assert p.max_image_size == (640, 480) and p.min_face_height == 40 and p.max_face_height == 200
    assert d.getMaxImageSize() == (640, 480) and d.getMaxObjectHeight() == 200
    # setParam re-initialises with the new parameters
    p.min_face_height = 60
    d.setParam(p)
    q = api.FaceDetector()
    assert q.Init(p) == 0
    assert same(d.Detect(frames[1]), q.Detect(frames[1]))
    # max object height and max image size
    d = fresh()
    d.setMaxObjectHeight(120)
    assert d.getMaxObjectHeight() == 120
    q = fresh(max_face=120)
    assert all(same(d.Detect(f), q.Detect(f)) for f in frames)
    d = fresh()
    d.setMaxImageSize((800, 600))
    assert d.getMaxImageSize() == (800, 600)
    pq = api_param(frames[0], dict(min_face=40, max_face=200))
    pq.max_image_size = (800, 600)
    q = api.FaceDetector()
    assert q.Init(pq) == 0
    assert all(same(d.Detect(f), q.Detect(f)) for f in frames)
    # the thread count is recorded (the GPU path has no host worker threads to size)
    d.setNumThreads(3)
    assert d.getNumThreads() == 3
    d.set_compute_precision(api.FP32_BF16)
    assert d.get_compute_precision() == api.FP32_BF16
    # profiling on: per-frame kernel times, and the batch's per-group times
    plain = fresh()
    base = [plain.Detect(f) for f in frames]
    batch = fresh().DetectBatch(images=frames)
    prof = fresh()
    prof.set_profiling(True)
    got = [prof.Detect(f) for f in frames]
    t = prof.last_timing()
    assert all(same(a, b) for a, b in zip(got, base))
    assert t["frames"] == len(frames)
    assert all(t["last"][k] > 0 for k in ("gray", "stage1_l1", "stage1", "check", "total"))
    assert t["totals"]["total"] >= t["last"]["total"]
    prof = fresh()
    prof.set_profiling(True)
    got = prof.DetectBatch(images=frames)
    bt = prof.last_batch_timing()
    assert all(same(a, b) for a, b in zip(got, batch))
    assert bt["frames"] == len(frames)
    assert all(bt[k] > 0 for k in ("gray", "stage1_l1", "stage1", "compact", "check", "total"))
    assert bt["total"] >= bt["stage1"] + bt["check"]


def test_gray_image_of_last_frame(api, oracle):
    # getGrayImage32F (CompactCNNLibAPI_v2.cpp:462-507): the FP32 gray image the last
    # Detect worked on, as a host copy (Pipeline::CPU) and as a device pointer
    from cases import vga
    f = vga(2)
    d = api.FaceDetector()
    assert d.Init(api_param(f, dict(min_face=40))) == 0
    d.Detect(f)
    g = d.getGrayImage32F(api.PIPELINE_CPU)
    assert g is not None and g.shape == f.shape[:2]
    assert np.array_equal(bits(g), bits(oracle.gray(f)))
    dev = d.getGrayImage32F(api.PIPELINE_GPU)
    assert dev is not None and dev.cols == f.shape[1] and dev.rows == f.shape[0] and dev.data
    # a 1-channel frame with min face >= 34 is pre-resized: the gray image is the resized one
    gray_frame = np.ascontiguousarray(f[:, :, 1])
    d2 = api.FaceDetector()
    assert d2.Init(api_param(gray_frame, dict(min_face=40))) == 0
    d2.Detect(gray_frame)
    g2 = d2.getGrayImage32F(api.PIPELINE_CPU)
    sw, sh = int(640 * np.float32(27) / np.float32(40)), int(480 * np.float32(27) / np.float32(40))
    assert g2.shape == (sh, sw)
    exp = oracle.resize(gray_frame, sw, sh, 1).astype(np.float32)
    assert np.array_equal(bits(g2), bits(exp))


def test_parameter_setters_and_clear(api):
    from cases import vga
    f = vga(0)
    d = api.FaceDetector()
    assert d.isEmpty()
    assert d.Init(api_param(f, dict(min_face=40))) == 0
    assert not d.isEmpty()
    n40, faces40 = d.Detect(f)
    # setMinObjectHeight re-initialises the detector like FaceDetector::setMinObjectHeight
    d.setMinObjectHeight(80)
    assert d.getMinObjectHeight() == 80
    n80, _ = d.Detect(f)
    d.setMinObjectHeight(40)
    again = d.Detect(f)
    assert again[0] == n40 and np.array_equal(again[1], faces40)
    assert d.Clear() == 0 and d.isEmpty()
    assert d.Clear() == -1
    assert d.Detect(f)[0] == -1
