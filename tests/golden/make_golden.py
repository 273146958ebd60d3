"""Writes tests/golden/detections.json and tests/golden/vectors.npz.

Run in a container that has the reference checkout and its CPU path built
(make -C oracle ref): the expected detections are those of the compiled reference
(oracle/_ref/libccnn_ref_O0.so, the reference sources with their CMake-default flags).
The small-vector fixtures pin the individual stages: gray conversion, resizes, the
stage-1/2/3 networks and the check-crop preparation.

    python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from cases import CASES, ORACLE_ONLY_CASES, frame_digest, params_for  # noqa: E402
from oracle.bind import Oracle, Ref  # noqa: E402


def faces_list(a):
    """x, y, w, h, score, gender, smile, glasses"""
    return [[int(v[0]), int(v[1]), int(v[2]), int(v[3]), float(v[4]), int(v[5]), int(v[6]), int(v[7])] for v in a]


def main():
    ref = Ref("O0")
    orc = Oracle()
    out = {"source": "compiled reference CPU path (CMake-default flags) unless noted", "cases": {}}
    for name, (build, over) in CASES.items():
        f = build()
        p = params_for(f, over)
        faces = ref.detect(f, p)
        if isinstance(faces, int):
            raise SystemExit("reference detect failed for %s" % name)
        out["cases"][name] = {"digest": frame_digest(f), "params": p, "faces": faces_list(faces),
                              "score_bits": [int(np.float32(v[4]).view(np.uint32)) for v in faces],
                              "producer": "reference"}
        print(name, len(faces), flush=True)
    for name, (build, over) in ORACLE_ONLY_CASES.items():
        f = build()
        p = params_for(f, over)
        faces, _ = orc.detect(f, p)
        out["cases"][name] = {"digest": frame_digest(f), "params": p, "faces": faces_list(faces),
                              "score_bits": [int(np.float32(v[4]).view(np.uint32)) for v in faces],
                              "producer": "oracle"}
        print(name, len(faces), "(oracle)", flush=True)
    with open(os.path.join(HERE, "detections.json"), "w") as fh:
        json.dump(out, fh, indent=1)

    # small-vector fixtures from the reference's public functions
    rng = np.random.default_rng(2024)
    vec = {}
    img3 = rng.integers(0, 256, size=(37, 41, 3), dtype=np.uint8)
    img4 = rng.integers(0, 256, size=(23, 29, 4), dtype=np.uint8)
    img1 = rng.integers(0, 256, size=(19, 31), dtype=np.uint8)
    vec["gray_in3"], vec["gray_out3"] = img3, ref.gray(img3)
    vec["gray_in4"], vec["gray_out4"] = img4, ref.gray(img4)
    vec["gray_in1"], vec["gray_out1"] = img1, ref.gray(img1)
    src = (rng.random((53, 67)) * 255).astype(np.float32)
    for i, (dw, dh, mode) in enumerate([(37, 41, 0), (37, 41, 1), (90, 70, 0), (90, 70, 1), (67, 53, 0),
                                        (49, 51, 0), (20, 60, 1)]):
        vec["rs_out%d" % i] = ref.resize(src, dw, dh, mode)
        vec["rs_meta%d" % i] = np.array([dw, dh, mode], np.int32)
    vec["rs_src"] = src
    src8 = rng.integers(0, 256, size=(45, 61), dtype=np.uint8)
    vec["rs8_src"] = src8
    vec["rs8_out"] = ref.resize(src8, 40, 30, 1)
    lvl = (rng.random((71, 59)) * 255).astype(np.float32)
    vec["s1_in"], vec["s1_out"] = lvl, ref.forward(0, lvl)
    chk = (rng.random((51, 47)) * 255).round().astype(np.float32)
    vec["s2_in"], vec["s2_out"], vec["s3_out"] = chk, ref.forward(1, chk), ref.forward(2, chk)
    crop = (rng.random((44, 39)) * 255).astype(np.float32)
    u8, s2, s3 = ref.check_prep(crop, 1, 1)
    vec["prep_crop"], vec["prep_u8"], vec["prep_s2"], vec["prep_s3"] = crop, u8, s2, s3
    # the landmarks network (all 13 outputs) on two 49x51 crops
    fa_in = np.stack([(rng.random((51, 49)) * 255).astype(np.float32),
                      np.round(rng.random((51, 49)) * 255).astype(np.float32)])
    vec["fa_in"] = fa_in
    vec["fa_out"] = np.stack([ref.forward(3, x)[:, 0] for x in fa_in])
    np.savez_compressed(os.path.join(HERE, "vectors.npz"), **vec)
    print("wrote", os.path.join(HERE, "detections.json"), os.path.join(HERE, "vectors.npz"))


if __name__ == "__main__":
    main()
