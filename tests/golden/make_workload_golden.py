"""Writes tests/golden/sequences.json: the detections of frame sequences run in turn on
one detector of the compiled reference (oracle/_ref/libccnn_ref_O0.so, CMake-default
flags), so that the detector state carried from call to call (max_image_size,
num_scales) is part of what is pinned. Covers the short state sequences of
tests/cases.py SEQ_CASES and every frame of the benchmark workloads (WORKLOAD_SEQS: the
64 Full-HD frames and the 32-frame mixed-size batch).

Run in a container that has the reference checkout and its CPU path built
(make -C oracle ref):

    python tests/golden/make_workload_golden.py
"""
import json
import os
import sys
from concurrent.futures import ProcessPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from cases import SEQ_CASES, WORKLOAD_SEQS, frame_digest, seq_params, workload_frames  # noqa: E402


def faces_list(a):
    """x, y, w, h, score, gender, smile, glasses"""
    return [[int(v[0]), int(v[1]), int(v[2]), int(v[3]), float(v[4]), int(v[5]), int(v[6]), int(v[7])] for v in a]


def run(item):
    from oracle.bind import Ref
    name, kind = item
    sets = None
    if kind == "seq":
        builders, over, mx = SEQ_CASES[name][:3]
        sets = SEQ_CASES[name][3] if len(SEQ_CASES[name]) > 3 else None
        frames = [b() for b in builders]
    else:
        spec, over, mx = WORKLOAD_SEQS[name]
        frames = workload_frames(spec)
    p = seq_params(over, mx)
    out = Ref("O0").detect_seq(frames, p, min_face_sets=sets)
    rec = {"params": p, "digests": [frame_digest(f) for f in frames],
           "faces": [faces_list(a) for a in out],
           "score_bits": [[int(np.float32(v[4]).view(np.uint32)) for v in a] for a in out],
           "producer": "reference"}
    print(name, [len(a) for a in out], flush=True)
    return name, rec


def main():
    items = [(k, "seq") for k in SEQ_CASES] + [(k, "workload") for k in WORKLOAD_SEQS]
    with ProcessPoolExecutor(max_workers=len(items)) as ex:
        res = dict(ex.map(run, items))
    out = {"source": "compiled reference CPU path (CMake-default flags), one detector per sequence",
           "cases": {k: res[k] for k, _ in items}}
    with open(os.path.join(HERE, "sequences.json"), "w") as fh:
        json.dump(out, fh, indent=0)
    print("wrote", os.path.join(HERE, "sequences.json"))


if __name__ == "__main__":
    main()
