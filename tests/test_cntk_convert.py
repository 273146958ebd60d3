"""The CNTK dump -> binary model converter (ccnn_cntk_dump_to_binary; the reference's
FaceDetector::CNTKDump2Binary, CompactCNNLibAPI_v2.cpp:509-519, over
CNNModelsConverter, cnn_models_converter.cpp:79-742). Host code only: no GPU.

Parity bar: byte-identical model files. The committed fixtures under tests/golden/cntk/
hold seeded synthetic dumps (tests/golden/cntk_dump.py) and the files the compiled
reference converter wrote for them (tests/golden/make_cntk_golden.py); where the
reference CPU path is built (oracle/_ref), further dumps are converted by both on the
spot. The converted models are then run through the compiled reference network and the
oracle's restatement of it, which must agree bit for bit."""
import ctypes
import filecmp
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "compactcnncascade_amd", "libccnn.so")
GOLD = os.path.join(ROOT, "tests", "golden", "cntk")
REF_O0 = os.path.join(ROOT, "oracle", "_ref", "libccnn_ref_O0.so")
ORACLE = os.path.join(ROOT, "oracle", "_ref", "libccnn_oracle.so")
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))

from cntk_dump import SHAPES, make_dump  # noqa: E402

pytestmark = pytest.mark.skipif(not os.path.exists(LIB), reason="libccnn.so not built")
needs_ref = pytest.mark.skipif(not os.path.exists(REF_O0), reason="reference CPU path not built (make -C oracle ref)")


def convert(binary, dump):
    lib = ctypes.CDLL(LIB)
    lib.ccnn_cntk_dump_to_binary.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
    return lib.ccnn_cntk_dump_to_binary(None if binary is None else str(binary).encode(),
                                        None if dump is None else str(dump).encode())


def ref_convert(binary, dump):
    lib = ctypes.CDLL(REF_O0)
    lib.ref_cntk_dump2binary.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
    lib.ref_cntk_dump2binary(str(binary).encode(), str(dump).encode())


def write_dump(path, text):
    with open(path, "w", newline="") as f:
        f.write(text)
    return path


@pytest.mark.parametrize("shape", sorted(SHAPES))
def test_converter_matches_golden_binary(tmp_path, shape):
    out = tmp_path / ("%s.bin" % shape)
    assert convert(out, os.path.join(GOLD, "%s_dump.txt" % shape)) == 0
    assert filecmp.cmp(out, os.path.join(GOLD, "%s.bin" % shape), shallow=False)


def test_golden_binaries_have_the_expected_layout():
    # format v1.1 header (cnn_simd_cntk.cpp:71-130): version, max-pool byte, input size,
    # layer count, maps per layer; af_scale 0 for more than three outputs
    for shape, (maps, ks, groups, hl_scale, connect, outs, (iw, ih), full) in SHAPES.items():
        b = open(os.path.join(GOLD, "%s.bin" % shape), "rb").read()
        assert np.frombuffer(b[:4], np.float32)[0] == np.float32(1.1)
        hdr = np.frombuffer(b[5:5 + 4 * 6], np.int32)
        assert list(hdr) == [iw, ih, 3] + list(maps)
        # the fixture for stage2 is written with average pooling
        assert b[4] == (0 if shape == "stage2" else 1)
        n_kern = sum(m * kw * kh for m, (kw, kh) in zip(maps, ks))
        n_map = 5 * sum(maps)
        hl = groups * hl_scale
        nw = maps[2] if full else connect
        size = 4 + 1 + 4 * (2 + 1 + 3) + 4 * 6 + 4 * n_kern + 4 + 4 * n_map + 1 + 4 * 3 + 4 * (
            hl * nw + hl + groups + 2 * hl_scale) + 4 + 4 * (outs * hl + outs + 1)
        assert len(b) == size


def test_binary_suffix_and_failures(tmp_path):
    dump = write_dump(tmp_path / "d.txt", make_dump("stage1", 5))
    # ".bin" is appended to a name without it (cnn_models_converter.cpp:607-610)
    assert convert(tmp_path / "model", dump) == 0
    assert (tmp_path / "model.bin").exists() and not (tmp_path / "model").exists()
    # an unreadable dump writes nothing
    assert convert(tmp_path / "x.bin", tmp_path / "missing.txt") == -1
    assert not (tmp_path / "x.bin").exists()
    # a dump without one of the networks' parameter nodes is not converted
    for drop in ["c12.W", "bn31.runVariance", "h3.b", "bn_h.scale", "z.arrayOfFunctions[0].b"]:
        d = write_dump(tmp_path / "bad.txt", make_dump("stage1", 6, drop=drop))
        assert convert(tmp_path / "bad.bin", d) == -1, drop
        assert not (tmp_path / "bad.bin").exists()
    assert convert(None, dump) == -1 and convert(tmp_path / "y.bin", None) == -1


@needs_ref
@pytest.mark.parametrize("shape", sorted(SHAPES))
def test_converter_matches_compiled_reference(tmp_path, shape):
    variants = [dict(), dict(max_pool=False), dict(sep="\r\n"), dict(dot_underscore=False),
                dict(n_out=1), dict(n_out=4), dict(n_out=2, max_pool=False)]
    for i, opts in enumerate(variants):
        dump = write_dump(tmp_path / ("%d.txt" % i), make_dump(shape, 100 + i, **opts))
        a, b = tmp_path / ("a%d.bin" % i), tmp_path / ("b%d.bin" % i)
        assert convert(a, dump) == 0
        ref_convert(b, dump)
        assert filecmp.cmp(a, b, shallow=False), (shape, opts)


@needs_ref
def test_missing_dump_writes_nothing_in_either(tmp_path):
    ref_convert(tmp_path / "r.bin", tmp_path / "missing.txt")
    assert convert(tmp_path / "o.bin", tmp_path / "missing.txt") == -1
    assert not (tmp_path / "r.bin").exists() and not (tmp_path / "o.bin").exists()


@needs_ref
@pytest.mark.skipif(not os.path.exists(ORACLE), reason="oracle not built")
@pytest.mark.parametrize("shape,index_output", [("stage1", 1), ("stage2", 0)])
def test_converted_model_runs_identically_in_reference_and_oracle(tmp_path, shape, index_output):
    # the converted file is a model both the compiled reference network
    # (ConvNeuralNetwork::Init/Forward, cnn_simd_cntk.cpp:42-404, 964) and the oracle load,
    # and their responses agree bit for bit
    from oracle.bind import OcModel, Ref, _ptr
    ref = Ref("O0")
    orc = ctypes.CDLL(ORACLE)
    m = OcModel()
    # both networks run max pooling only (cnn_simd_cntk.cpp:322-326): a model converted
    # from an average-pooling dump is refused by both
    avg = tmp_path / "avg.bin"
    assert convert(avg, write_dump(tmp_path / "avg.txt", make_dump(shape, 7, max_pool=False))) == 0
    assert not ref.lib.ref_cnn_create(str(avg).encode(), index_output)
    data = open(avg, "rb").read()
    assert orc.oc_model_load(ctypes.byref(m), data, ctypes.c_long(len(data)), index_output) != 0
    out = tmp_path / "m.bin"
    assert convert(out, write_dump(tmp_path / "m.txt", make_dump(shape, 8))) == 0
    h = ref.lib.ref_cnn_create(str(out).encode(), index_output)
    assert h
    m = OcModel()
    data = open(out, "rb").read()
    assert orc.oc_model_load(ctypes.byref(m), data, ctypes.c_long(len(data)), index_output) == 0
    rng = np.random.default_rng(3)
    iw, ih = SHAPES[shape][6]
    for w, hh in [(iw, ih), (iw + 13, ih + 9)]:
        img = np.ascontiguousarray(rng.integers(0, 256, size=(hh, w)).astype(np.float32))
        ow, oh = ctypes.c_int(), ctypes.c_int()
        orc.oc_output_size(ctypes.byref(m), w, hh, ctypes.byref(ow), ctypes.byref(oh))
        mine = np.zeros((oh.value, ow.value), np.float32)
        assert orc.oc_cnn_forward(ctypes.byref(m), _ptr(img), w, hh, w, _ptr(mine)) == 0
        theirs = np.zeros(w * hh, np.float32)
        rw, rh = ctypes.c_int(), ctypes.c_int()
        assert ref.lib.ref_cnn_forward(h, _ptr(img), w, hh, w, w, hh, _ptr(theirs), w * hh,
                                       ctypes.byref(rw), ctypes.byref(rh)) == 0
        assert (rw.value, rh.value) == (ow.value, oh.value)
        theirs = theirs[:rw.value * rh.value].reshape(rh.value, rw.value)
        assert np.array_equal(mine.view(np.uint32), theirs.view(np.uint32))
    ref.lib.ref_cnn_destroy(h)
    orc.oc_model_free(ctypes.byref(m))


def test_facade_static_member_converts(tmp_path):
    # FaceDetector::CNTKDump2Binary as a reference user calls it (prebuild/examples/main.cpp:47)
    import shutil
    import subprocess
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    src = tmp_path / "conv.cpp"
    src.write_text(r'''
#include "CompactCNNLib.h"
int main(int argc, char** argv) {
    CompactCNNLib::FaceDetector::CNTKDump2Binary(argv[1], argv[2]);
    return 0;
}
''')
    exe = tmp_path / "conv"
    subprocess.run([gxx, "-std=c++11", "-I" + os.path.join(ROOT, "include"), str(src), "-o", str(exe),
                    "-L" + os.path.dirname(LIB), "-Wl,-rpath," + os.path.dirname(LIB), "-lccnn"], check=True)
    out = tmp_path / "fc13.bin"
    subprocess.run([str(exe), str(out), os.path.join(GOLD, "fc13_dump.txt")], check=True, timeout=120)
    assert filecmp.cmp(out, os.path.join(GOLD, "fc13.bin"), shallow=False)
