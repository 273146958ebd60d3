"""The oracle (oracle/ccnn_oracle.c) pinned against the reference and against what the
mathematics fixes. CPU only.

- golden fixtures produced by the compiled reference (tests/golden/make_golden.py)
- the compiled reference itself when oracle/_ref is built
- closed forms and textbook restatements in float64 (numpy), chosen so that a dropped
  term, a wrong index or a transposed operand fails
- the reference's own equalize property test (test_accuracy.cpp:497-559)
- libstdc++ std::sort's permutation on tied keys (a g++-built helper)
"""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from cases import CASES, GOLDEN, ORACLE_ONLY_CASES, frame_digest, params_for

REF_O0 = os.path.join(os.path.dirname(GOLDEN), "..", "oracle", "_ref", "libccnn_ref_O0.so")


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.fixture(scope="module")
def golden():
    with open(os.path.join(GOLDEN, "detections.json")) as f:
        return json.load(f)["cases"]


@pytest.fixture(scope="module")
def vectors():
    return dict(np.load(os.path.join(GOLDEN, "vectors.npz")))


# ---------------------------------------------------------------------------------
# golden fixtures
# ---------------------------------------------------------------------------------
@pytest.mark.parametrize("name", list(CASES) + list(ORACLE_ONLY_CASES))
def test_oracle_detections_match_golden(oracle, golden, name):
    build, over = {**CASES, **ORACLE_ONLY_CASES}[name]
    if name.startswith("fhd") or name.startswith("big_"):
        pytest.skip("large frames: covered by the GPU parity tests against the compiled reference's goldens")
    f = build()
    g = golden[name]
    assert frame_digest(f) == g["digest"]
    faces, _ = oracle.detect(f, params_for(f, over))
    assert faces[:, :4].astype(int).tolist() == [e[:4] for e in g["faces"]]
    assert [int(v) for v in bits(faces[:, 4])] == g["score_bits"]
    # gender / smile / glasses (-1 without facial analysis)
    assert faces[:, 5:8].astype(int).tolist() == [e[5:8] for e in g["faces"]]


def test_oracle_small_vectors(oracle, vectors):
    v = vectors
    for k in ("3", "4", "1"):
        assert np.array_equal(bits(oracle.gray(v["gray_in" + k])), bits(v["gray_out" + k]))
    for i in range(7):
        dw, dh, mode = v["rs_meta%d" % i]
        assert np.array_equal(bits(oracle.resize(v["rs_src"], int(dw), int(dh), int(mode))), bits(v["rs_out%d" % i]))
    assert np.array_equal(oracle.resize(v["rs8_src"], 40, 30, 1), v["rs8_out"])
    assert np.array_equal(bits(oracle.forward(0, v["s1_in"])), bits(v["s1_out"]))
    assert np.array_equal(bits(oracle.forward(1, v["s2_in"])), bits(v["s2_out"]))
    assert np.array_equal(bits(oracle.forward(2, v["s2_in"])), bits(v["s3_out"]))
    u8, s2, s3, _ = oracle.check_prep(v["prep_crop"], 1, 1)
    assert np.array_equal(u8, v["prep_u8"])
    assert np.array_equal(s2, v["prep_s2"]) and np.array_equal(s3, v["prep_s3"])
    for x, y in zip(v["fa_in"], v["fa_out"]):
        assert np.array_equal(bits(oracle.forward(3, x)[:, 0, 0]), bits(y))


@pytest.mark.skipif(not os.path.exists(REF_O0), reason="compiled reference not built (make -C oracle ref)")
def test_oracle_matches_compiled_reference_random(oracle):
    from oracle.bind import Ref, default_params
    from compactcnncascade_amd import synth
    ref = Ref("O0")
    rng = np.random.default_rng(77)
    for trial in range(3):
        w, h = int(rng.integers(60, 300)), int(rng.integers(60, 260))
        img = rng.integers(0, 256, size=(h, w, 3), dtype=np.uint8)
        assert np.array_equal(bits(oracle.gray(img)), bits(ref.gray(img)))
        g = oracle.gray(img)
        for mode in (0, 1):
            dw, dh = int(rng.integers(30, 400)), int(rng.integers(30, 400))
            assert np.array_equal(bits(oracle.resize(g, dw, dh, mode)), bits(ref.resize(g, dw, dh, mode)))
        lvl = (rng.random((int(rng.integers(51, 140)), int(rng.integers(47, 160)))) * 255).astype(np.float32)
        assert np.array_equal(bits(oracle.forward(0, lvl)), bits(ref.forward(0, lvl)))
    f = synth.synth_frame(31, 320, 240, 5, 25, 120)[0]
    p = default_params(320, 240, min_face=25)
    a, _ = oracle.detect(f, p)
    b = ref.detect(f, p)
    assert np.array_equal(a[:, :4], b[:, :4]) and np.array_equal(bits(a[:, 4]), bits(b[:, 4]))


# ---------------------------------------------------------------------------------
# closed forms
# ---------------------------------------------------------------------------------
def test_gray_closed_form_all_triples(oracle):
    v = np.arange(1 << 24, dtype=np.uint32)
    B, G, R = (v & 255), (v >> 8) & 255, (v >> 16) & 255
    img = np.stack([B, G, R], axis=1).astype(np.uint8).reshape(4096, 4096, 3)
    f32 = np.float32
    expect = (f32(0.114) * B.astype(f32) + f32(0.587) * G.astype(f32)) + f32(0.299) * R.astype(f32)
    assert np.array_equal(bits(oracle.gray(img)).ravel(), bits(expect))
    # BGRA ignores alpha; one channel is a plain conversion
    bgra = np.concatenate([img[:64], np.full((64, 4096, 1), 7, np.uint8)], axis=2)
    assert np.array_equal(bits(oracle.gray(bgra)), bits(oracle.gray(img[:64])))
    g1 = img[:3, :, 0]
    assert np.array_equal(oracle.gray(g1), g1.astype(np.float32))


def _lut(sw, sh, dw, dh):
    sclx = ((sw << 12) // dw) + 1
    scly = ((sh << 12) // dh) + 1
    x = np.arange(dw, dtype=np.int64) * sclx
    y = np.arange(dh, dtype=np.int64) * scly
    px, py = x >> 12, y >> 12
    return x, y, px, py


def test_resize_nearest_closed_form(oracle):
    rng = np.random.default_rng(3)
    for (sw, sh, dw, dh) in [(40, 30, 17, 23), (40, 30, 90, 70), (13, 11, 49, 51), (100, 57, 33, 80)]:
        src = (rng.random((sh, sw)) * 255).astype(np.float32)
        _, _, px, py = _lut(sw, sh, dw, dh)
        px = np.minimum(px, sw - 2)          # the x index never reaches the last column
        py = np.minimum(py, sh - 1)
        assert np.array_equal(oracle.resize(src, dw, dh, 0), src[py][:, px])
    src = (rng.random((21, 19)) * 255).astype(np.float32)
    assert np.array_equal(oracle.resize(src, 19, 21, 0), src)           # equal sizes copy


def test_resize_bilinear_closed_form(oracle):
    rng = np.random.default_rng(4)
    f32 = np.float32
    for (sw, sh, dw, dh) in [(40, 30, 17, 23), (64, 48, 100, 70), (30, 50, 29, 49)]:
        src = (rng.random((sh, sw)) * 255).astype(np.float32)
        x, y, px, py = _lut(sw, sh, dw, dh)
        fx = (x - (px << 12)).astype(f32)
        cx = f32(4096) - fx
        fy = (y - (py << 12)).astype(f32) / f32(2 ** 24)
        cy = (f32(4096) - (y - (py << 12)).astype(f32)) / f32(2 ** 24)
        px = np.minimum(px, sw - 2)
        py = np.minimum(py, sh - 1)
        py = np.where(py + 1 >= sh, py - 1, py)
        p0 = src[py][:, px]
        p1 = src[py][:, px + 1]
        p2 = src[py + 1][:, px]
        p3 = src[py + 1][:, px + 1]
        exp = (p0 * cx + p1 * fx) * cy[:, None] + (p2 * cx + p3 * fx) * fy[:, None]
        assert np.array_equal(bits(oracle.resize(src, dw, dh, 1)), bits(exp.astype(f32)))


@pytest.mark.parametrize("stage,shape", [(0, (51, 47)), (0, (90, 75)), (1, (51, 47)), (2, (51, 47)), (3, (51, 49))])
def test_networks_match_float64_restatement(oracle, stage, shape):
    rng = np.random.default_rng(10 + stage)
    x = (rng.random(shape) * 255).round().astype(np.float32)
    got = oracle.forward(stage, x).astype(np.float64)
    from oracle.net64 import forward64
    exp = forward64(oracle.models[stage], x)
    assert got.shape == exp.shape
    assert np.max(np.abs(got - exp)) < 2e-5 * 1.7159
    if stage == 3:
        # the landmarks network: 13 output maps of one value each
        assert got.shape == (13, 1, 1)


def test_stage1_output_size_rule(oracle):
    for (w, h) in [(47, 51), (48, 52), (23, 27), (1920, 1080), (333, 211)]:
        rw, rh = oracle.output_size(0, w, h)
        assert rw == ((((w - 3) >> 1) - 2) >> 1) - 3 and rh == ((((h - 3) >> 1) - 2) >> 1) - 4
    assert oracle.output_size(1, 47, 51) == (4, 4)


def test_equalize_closed_form_and_reference_property(oracle):
    rng = np.random.default_rng(6)
    img = rng.integers(0, 256, size=(51, 49)).astype(np.uint8)
    out = oracle.equalize(img)
    hist = np.bincount(img.ravel(), minlength=256)
    cf = np.cumsum(hist).astype(np.float32)
    delt = np.float32(255) / (cf[255] - cf[0])
    lut = (((cf - cf[0]) * delt) + np.float32(0.5)).astype(np.int32)
    assert np.array_equal(out, lut[img].astype(np.uint8))
    # the reference's own property test (test_accuracy.cpp:497-559): a 1024x1024 image of
    # four quadrant values, written with its pointer-advance pattern over a zeroed image,
    # must equalize to levels reaching the top of the range (256 - L_eq <= 2)
    n = 1024
    q = np.zeros((n, n), np.uint8)
    for j in range(n):
        row = []
        for i in range(n):
            if j > n // 2 and i < n // 2: row.append(0)
            if j < n // 2 and i > n // 2: row.append(1)
            if j > n // 2 and i > n // 2: row.append(2)
            if j < n // 2 and i < n // 2: row.append(3)
        q[j, :len(row)] = row
    e = oracle.equalize(q)
    occupied = np.nonzero(np.bincount(e.ravel(), minlength=256))[0]
    l_eq = int(occupied.max())            # sum of the gaps between occupied bins from 0
    assert 256 - l_eq <= 2
    # b_min is the count of zeros, not the first occupied bin: only an all-zero image is
    # left as it is; a uniform non-zero image maps to 255
    zero = np.zeros((51, 49), np.uint8)
    assert np.array_equal(oracle.equalize(zero), zero)
    assert (oracle.equalize(np.full((51, 49), 17, np.uint8)) == 255).all()


def test_check_prep_closed_form(oracle):
    rng = np.random.default_rng(8)
    crop = (rng.random((49 * 2, 51)) * 255).astype(np.float32)[:51, :49]
    u8, s2, s3, f32 = oracle.check_prep(crop, equalize=0, reflection=1)
    assert np.array_equal(u8, (np.ascontiguousarray(crop) + np.float32(0.5)).astype(np.uint8))   # 49x51: a copy
    assert np.array_equal(s2, u8[:, 2:49].astype(np.float32))
    assert np.array_equal(s3, u8[:, 46::-1].astype(np.float32))
    _, _, s3b, _ = oracle.check_prep(crop, equalize=0, reflection=0)
    assert np.array_equal(s3b, u8[:, :47].astype(np.float32))


def test_sort_matches_libstdcxx(oracle, tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    src = tmp_path / "s.cpp"
    src.write_text(r'''
#include <algorithm>
#include <cstdio>
#include <vector>
int main() {
    int n; std::vector<int> perm; std::vector<float> key;
    while (std::scanf("%d", &n) == 1) {
        key.resize(n); perm.resize(n);
        for (int i = 0; i < n; ++i) { std::scanf("%f", &key[i]); perm[i] = i; }
        std::sort(perm.begin(), perm.end(), [&](int a, int b) { return key[b] < key[a]; });
        for (int i = 0; i < n; ++i) std::printf("%d ", perm[i]);
        std::printf("\n");
    }
}''')
    exe = tmp_path / "s"
    subprocess.run([gxx, "-O1", "-o", str(exe), str(src)], check=True)
    rng = np.random.default_rng(12)
    lines, cases = [], []
    for n in list(range(1, 40)) + [64, 100, 257, 1000]:
        k = rng.integers(0, max(1, n // 4 + 1), size=n).astype(np.float32) / 7
        cases.append(k)
        lines.append(" ".join([str(n)] + ["%.9g" % v for v in k]))
    out = subprocess.run([str(exe)], input="\n".join(lines) + "\n", capture_output=True, text=True, check=True)
    got = out.stdout.strip("\n").split("\n")
    for k, line in zip(cases, got):
        exp = [int(t) for t in line.split()]
        assert oracle.sort_desc(k).tolist() == exp


def test_model_loader_rejects_bad_files(oracle):
    import ctypes
    from oracle.bind import OcModel, model_bytes
    b = model_bytes(0)
    m = OcModel()
    assert oracle.lib.oc_model_load(ctypes.byref(m), b, len(b), 1) == 0
    assert oracle.lib.oc_model_load(ctypes.byref(OcModel()), b[:100], 100, 1) == -1      # truncated
    bad = bytearray(b)
    bad[0:4] = np.float32(2.0).tobytes()                                                # version 2.0
    assert oracle.lib.oc_model_load(ctypes.byref(OcModel()), bytes(bad), len(bad), 1) == -1


# ---------------------------------------------------------------------------------
# detector state across calls (max_image_size, num_scales, gray_image_only)
# ---------------------------------------------------------------------------------
@pytest.mark.parametrize("name", ["seq_max_small_max_s45", "seq_max_small_max_s41_min20", "seq_grow_s41",
                                  "seq_setter_gray_after_colour"])
def test_oracle_sequences_match_reference_goldens(oracle, name):
    from cases import SEQ_CASES, seq_params
    with open(os.path.join(GOLDEN, "sequences.json")) as f:
        g = json.load(f)["cases"][name]
    builders, over, mx = SEQ_CASES[name][:3]
    sets = SEQ_CASES[name][3] if len(SEQ_CASES[name]) > 3 else None
    frames = [b() for b in builders]
    assert [frame_digest(f) for f in frames] == g["digests"]
    out = oracle.detect_seq(frames, seq_params(over, mx), min_face_sets=sets)
    for i, faces in enumerate(out):
        assert faces[:, :4].astype(int).tolist() == [e[:4] for e in g["faces"][i]], (name, i)
        assert [int(v) for v in bits(faces[:, 4])] == g["score_bits"][i], (name, i)
    # the carried state matters here: a fresh detector per frame gives other boxes
    if name != "seq_setter_gray_after_colour":
        fresh = [oracle.detect(f, seq_params(over, mx))[0] for f in frames]
        assert any(len(a) != len(b) or not np.array_equal(a, b) for a, b in zip(out, fresh))


# ---------------------------------------------------------------------------------
# the FP64 oracle of the tile-skipping network
# ---------------------------------------------------------------------------------
def test_skipnet_oracle_gradient_matches_finite_differences():
    from oracle import skipnet64 as sn
    rng = np.random.default_rng(4)
    p = sn.init_params(9)
    x = rng.random((2, 6, 8)) - 0.5
    y = rng.random((2, sn.S, 6, 8)) < 0.2
    loss, g = sn.grad64(p, x, y, 4.0)
    assert abs(loss - sn.loss64(p, x, y, 4.0)) < 1e-12
    # every bias and a sample of the weights, central differences
    idx = list(range(sn.NPAR - sn.S, sn.NPAR)) + list(rng.choice(sn.NPAR - sn.S, 40, replace=False))
    for i in idx:
        e = np.zeros_like(p)
        e[i] = 1e-6
        fd = (sn.loss64(p + e, x, y, 4.0) - sn.loss64(p - e, x, y, 4.0)) / 2e-6
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, g[i])
