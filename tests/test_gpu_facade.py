"""The C++ FaceDetector facade (include/CompactCNNLib.h) driven the way the reference's
examples/main.cpp drives it: a program compiled against the header and linked with
libccnn.so detects faces on a golden frame and must print the golden detections."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from cases import CASES, GOLDEN

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "compactcnncascade_amd")

PROGRAM = r'''
#include "CompactCNNLib.h"
#include <cstdio>
#include <vector>
int main(int argc, char** argv) {
    int w = 0, h = 0, ch = 0, minf = 0;
    std::FILE* f = std::fopen(argv[1], "rb");
    if (!f || std::fscanf(f, "%d %d %d %d\n", &w, &h, &ch, &minf) != 4) return 2;
    std::vector<unsigned char> img((size_t)w * h * ch);
    if (std::fread(img.data(), 1, img.size(), f) != img.size()) return 3;
    std::fclose(f);
    CompactCNNLib::FaceDetector fd;
    CompactCNNLib::FaceDetector::Param param;
    param.max_image_size = CompactCNNLib::FaceDetector::Size(w, h);
    param.min_face_height = minf;
    param.pipeline = CompactCNNLib::FaceDetector::Pipeline::GPU;
    if (fd.Init(param) < 0) return 4;
    CompactCNNLib::FaceDetector::ImageData im(w, h, ch, img.data(), (size_t)w * ch);
    std::vector<CompactCNNLib::FaceDetector::Face> faces(1000);
    int n = fd.Detect(faces.data(), im);
    for (int i = 0; i < n; ++i)
        std::printf("%d %d %d %d %08x\n", faces[i].x, faces[i].y, faces[i].width, faces[i].height,
                    *reinterpret_cast<unsigned*>(&faces[i].score));
    fd.Clear();
    return n >= 0 ? 0 : 5;
}
'''


@pytest.mark.parametrize("name", ["vga_s3_min40", "photo3_min20"])
def test_cpp_facade_detects_golden_faces(tmp_path, name):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    build, over = CASES[name]
    frame = np.ascontiguousarray(build())
    h, w = frame.shape[:2]
    ch = 1 if frame.ndim == 2 else frame.shape[2]
    data = tmp_path / "frame.bin"
    with open(data, "wb") as fh:
        fh.write(b"%d %d %d %d\n" % (w, h, ch, over["min_face"]))
        fh.write(frame.tobytes())
    src = tmp_path / "main.cpp"
    src.write_text(PROGRAM)
    exe = tmp_path / "main"
    subprocess.run([gxx, "-std=c++11", "-I" + os.path.join(ROOT, "include"), str(src), "-o", str(exe),
                    "-L" + LIBDIR, "-Wl,-rpath," + LIBDIR, "-lccnn"], check=True)
    print("facade", name, flush=True)
    out = subprocess.run([str(exe), str(data)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    got = [l.split() for l in out.stdout.splitlines() if l and l[0].isdigit() or (l and l[0] == "-")]
    with open(os.path.join(GOLDEN, "detections.json")) as fh:
        g = json.load(fh)["cases"][name]
    assert [[int(v) for v in r[:4]] for r in got] == [e[:4] for e in g["faces"]]
    assert [int(r[4], 16) for r in got] == g["score_bits"]
