"""Writes the CNTK dump -> binary converter fixtures under tests/golden/cntk/.

For each network shape of cntk_dump.SHAPES it writes a seeded synthetic dump
(<shape>_dump.txt) and the binary model that the compiled reference converter makes of
it (<shape>.bin; FaceDetector::CNTKDump2Binary, CompactCNNLibAPI_v2.cpp:509-519, in
oracle/_ref/libccnn_ref_O0.so). Run where the reference checkout is present and its CPU
path is built (make -C oracle ref):

    python tests/golden/make_cntk_golden.py
"""
import ctypes
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)

from cntk_dump import SHAPES, make_dump  # noqa: E402

# per shape: seed and the dump options that vary the text (pooling kind, line ends)
FIXTURES = {
    "stage1": (11, dict()),
    "stage2": (12, dict(max_pool=False, sep="\r\n")),
    "fc13": (13, dict(dot_underscore=False)),
}


def main():
    ref = ctypes.CDLL(os.path.join(ROOT, "oracle", "_ref", "libccnn_ref_O0.so"))
    out = os.path.join(HERE, "cntk")
    os.makedirs(out, exist_ok=True)
    for shape in SHAPES:
        seed, opts = FIXTURES[shape]
        dump = os.path.join(out, "%s_dump.txt" % shape)
        with open(dump, "w", newline="") as f:
            f.write(make_dump(shape, seed, **opts))
        binary = os.path.join(out, "%s.bin" % shape)
        if os.path.exists(binary):
            os.remove(binary)
        ref.ref_cntk_dump2binary(binary.encode(), dump.encode())
        if not os.path.exists(binary):
            raise SystemExit("the reference converter wrote nothing for %s" % shape)
        print(shape, os.path.getsize(dump), os.path.getsize(binary), flush=True)


if __name__ == "__main__":
    main()
