"""The C ABI and the C++ facade, checked without a GPU: the library loads, exports every
function include/ccnn.h declares and every FaceDetector member of include/CompactCNNLib.h,
keeps the reference's parameter defaults, a main.cpp-style program compiles and links
against it, and on a machine without a HIP device Init fails loudly (no CPU fallback)."""
import ctypes
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "compactcnncascade_amd", "libccnn.so")
HEADER = os.path.join(ROOT, "include", "ccnn.h")

pytestmark = pytest.mark.skipif(not os.path.exists(LIB), reason="libccnn.so not built")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ccnn_[a-z0-9_]+)\s*\(", text)))


def test_exports_every_declared_function():
    lib = ctypes.CDLL(LIB)
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert missing == []


def test_exports_facade_members():
    nm = shutil.which("nm")
    if nm is None:
        pytest.skip("nm not available")
    out = subprocess.run([nm, "-D", "--defined-only", "-C", LIB], capture_output=True, text=True, check=True).stdout
    for member in ["FaceDetector::FaceDetector()", "FaceDetector::~FaceDetector()", "FaceDetector::Init(",
                   "FaceDetector::Detect(", "FaceDetector::Clear()", "FaceDetector::isEmpty() const",
                   "FaceDetector::getParam() const", "FaceDetector::setParam(", "FaceDetector::getMaxImageSize() const",
                   "FaceDetector::setMaxImageSize(", "FaceDetector::getMinObjectHeight() const",
                   "FaceDetector::setMinObjectHeight(", "FaceDetector::getMaxObjectHeight() const",
                   "FaceDetector::setMaxObjectHeight(", "FaceDetector::getNumThreads() const",
                   "FaceDetector::setNumThreads(", "FaceDetector::getGrayImage32F(", "FaceDetector::CNTKDump2Binary("]:
        assert "CompactCNNLib::" + member in out, member


def test_param_defaults_match_reference_header():
    from compactcnncascade_amd import api
    c = api.CParam()
    api.load().ccnn_param_default(ctypes.byref(c))
    # prebuild/include/CompactCNNLib.h:98-165
    assert (c.max_image_width, c.max_image_height) == (1920, 1080)
    assert c.min_face_height == 80 and c.max_face_height == 0
    assert abs(c.scale_factor - 1.15) < 1e-7 and c.min_neighbors == 2
    assert c.detect_precision == api.PRECISION_NORMAL
    assert [round(t, 5) for t in c.treshold] == [-1.5, -1.25, -1.0]
    assert (c.drop_detect, c.equalize, c.reflection) == (1, 1, 1)
    assert c.pipeline == api.PIPELINE_GPU and c.detect_mode == api.DETECT_ASYNC and c.num_threads == 2
    assert list(c.index_output) == [1, 0, 0]


def _have_gpu():
    try:
        import torch
        return torch.cuda.device_count() > 0
    except Exception:  # noqa: BLE001
        return False


@pytest.mark.skipif(_have_gpu(), reason="a HIP device is present")
def test_init_fails_loudly_without_gpu():
    from compactcnncascade_amd import api
    d = api.FaceDetector()
    assert d.Init(api.Param()) == -1
    assert d.isEmpty()
    import numpy as np
    n, _ = d.Detect(np.zeros((100, 100, 3), np.uint8))
    assert n == -1                     # not initialised: the reference's -1


def test_error_paths_without_detector():
    from compactcnncascade_amd import api
    L = api.load()
    assert L.ccnn_detect(None, None, 0, None) == -1
    assert L.ccnn_clear(None) == -1
    assert L.ccnn_is_empty(None) == 1
    assert L.ccnn_stage1_size(10, 10, ctypes.byref(ctypes.c_int()), ctypes.byref(ctypes.c_int())) == -1


def test_main_cpp_style_program_compiles_and_links(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    src = tmp_path / "main.cpp"
    # the calls of the reference's prebuild/examples/main.cpp (image and video modes)
    src.write_text(r'''
#include "CompactCNNLib.h"
#include <cstdio>
#include <vector>
int main(int argc, char** argv) {
    CompactCNNLib::FaceDetector face_detector;
    CompactCNNLib::FaceDetector::Param param;
    param.max_image_size = CompactCNNLib::FaceDetector::Size(5000, 3000);
    param.min_face_height = 20;
    param.scale_factor = 1.1f;
    param.equalize = false;
    param.reflection = false;
    param.pipeline = CompactCNNLib::FaceDetector::Pipeline::GPU;
    param.num_threads = 4;
    param.facial_analysis = false;
    if (face_detector.Init(param) < 0) return argc > 1 ? 0 : 1;
    std::vector<unsigned char> img(640 * 480 * 3, 128);
    CompactCNNLib::FaceDetector::ImageData im(640, 480, 3, img.data(), 640 * 3);
    std::vector<CompactCNNLib::FaceDetector::Face> faces(1000);
    int n = face_detector.Detect(faces.data(), im);
    std::printf("faces %d\n", n);
    face_detector.Clear();
    return 0;
}
''')
    exe = tmp_path / "main"
    cmd = [gxx, "-std=c++11", "-I" + os.path.join(ROOT, "include"), str(src), "-o", str(exe),
           "-L" + os.path.dirname(LIB), "-Wl,-rpath," + os.path.dirname(LIB), "-lccnn"]
    subprocess.run(cmd, check=True)
    assert exe.exists()


def test_image_records_match_the_c_struct():
    # DetectBatch builds device-frame descriptors as numpy records; they must lay out
    # exactly as ccnn_image (cols, rows, channels, 4 bytes of padding, data, step)
    import numpy as np
    from compactcnncascade_amd import api
    assert api.CIMAGE_DTYPE.itemsize == ctypes.sizeof(api.CImage) == 32
    for name in ("cols", "rows", "channels", "data", "step"):
        assert api.CIMAGE_DTYPE.fields[name][1] == getattr(api.CImage, name).offset, name
    frames = [(0x7f0012340000 + 977 * i, 640 + i, 480 - i, (1, 3, 4)[i % 3], 3 * (640 + i) + 64) for i in range(9)]
    rec = np.empty(len(frames), api.CIMAGE_DTYPE)
    a = np.asarray(frames, dtype=np.uint64)
    rec["data"], rec["cols"], rec["rows"], rec["channels"], rec["step"] = a[:, 0], a[:, 1], a[:, 2], a[:, 3], a[:, 4]
    ref = bytes((api.CImage * len(frames))(*[api.CImage(c, r, ch, p, st) for (p, c, r, ch, st) in frames]))
    got = rec.tobytes()
    for i in range(len(frames)):
        assert got[32 * i:32 * i + 12] == ref[32 * i:32 * i + 12]
        assert got[32 * i + 16:32 * i + 32] == ref[32 * i + 16:32 * i + 32]
