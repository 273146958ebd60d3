"""Learned tile skipping (SkipNet): training data from the plain cascade, training on the
GPU with the library's own forward/backward kernels, and the model file.

The network (compactcnncascade_amd/csrc/ccnn_skipnet.hip) reads the gray frame
downsampled by 8 (each 8x8 block's mean / 255 - 0.5) and gives, per 8x8 frame cell and
per pyramid scale index s < SLOTS, the logit that the level of scale s has a window that
the cascade accepts whose centre lies in that cell. The labels come from the plain FP32
cascade itself: every candidate window the stage-2/3 checks accept (the stage-1 response
map thresholded at t1, then CPUCheckDetect; cnn_detector_v3.cpp:1485-1818, 1918-1947)
marks its centre cell and the cells within two of it.

At detection time the batch path runs the network on every frame, keeps each stage-1
response tile whose largest probability over its windows' cells (and nearby scale slots)
reaches the threshold, and runs the plain cascade on those tiles only; the levels without a slot, and every
tile the network does not rule out with that confidence, run as usual.

    python -m compactcnncascade_amd.skipnet train --frames 20000 --out models/skipnet_min20_sf115.npz
    python -m compactcnncascade_amd.skipnet validate --model M --out M   (operating points only)
    python -m compactcnncascade_amd.skipnet evaluate --model M --held-out 256

Operating points, chosen on validation frames (1920x1080 and 3840x2160, seeds apart from
training and testing) over keep windows (scale slots within a radius of a level's own,
cells within a margin of its windows' centres) and thresholds: 'recall', the fewest
tiles run at recall >= 0.995 against the plain cascade on every validation set, and
'fast', threshold 0.5 with the narrowest window.
"""
import argparse
import ctypes
import json
import os
import sys
import time
from concurrent.futures import ProcessPoolExecutor

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
if os.path.dirname(PKG) not in sys.path:
    sys.path.insert(0, os.path.dirname(PKG))

from compactcnncascade_amd import synth  # noqa: E402

SLOTS = 24
NPAR = 8 * 9 + 8 + 8 * 8 * 9 + 8 + SLOTS * 8 * 9 + SLOTS
DEFAULT_MODEL = os.path.join(PKG, "models", "skipnet_min20_sf115.npz")
TRAIN_WORKLOAD = dict(w=1280, h=720, lam=10.0, min_size=20, max_size=300)
MIN_FACE, SCALE_FACTOR = 20, 1.15


def frame_gray(frame):
    """FP32 gray of a BGR/BGRA/gray frame (image_proc.cpp:99-102, left to right)."""
    f = np.asarray(frame)
    if f.ndim == 2:
        return f.astype(np.float32)
    b, g, r = (f[:, :, k].astype(np.float32) for k in range(3))
    return (np.float32(0.114) * b + np.float32(0.587) * g) + np.float32(0.299) * r


def down8(gray):
    """The network input: the mean of each 8x8 block (partial blocks at the right and
    bottom edges over their own pixels) / 255 - 0.5, as k_sn_down8 computes it."""
    h, w = gray.shape
    h8, w8 = (h + 7) // 8, (w + 7) // 8
    pad = np.zeros((h8 * 8, w8 * 8), np.float64)
    cnt = np.zeros_like(pad)
    pad[:h, :w] = gray
    cnt[:h, :w] = 1
    s = pad.reshape(h8, 8, w8, 8).sum(axis=(1, 3))
    n = cnt.reshape(h8, 8, w8, 8).sum(axis=(1, 3))
    return (s / n / 255.0 - 0.5).astype(np.float32)


def scales_for(min_face, max_h, scale_factor, pattern_h=27):
    """InitScales' float32 recurrence (cnn_detector_v3.cpp:289-346) for max_face 0."""
    f32 = np.float32
    mx = min(f32(3.0), f32(pattern_h) / f32(min_face))
    mn = min(mx, f32(pattern_h) / f32(max_h))
    sf = f32(scale_factor)
    out, s = [], mx
    while s > mn:
        out.append(float(s))
        s = f32(s / sf)
    return out


def labels_from_trace(shape, trace, scales, dilate=2):
    """uint8 labels [SLOTS, h8, w8] from one frame's trace (levels, candidates, check
    results): accepted windows' centre cells and the cells within `dilate` of them."""
    h, w = shape
    h8, w8 = (h + 7) // 8, (w + 7) // 8
    y = np.zeros((SLOTS, h8, w8), np.uint8)
    lv = trace["levels"]
    for c, r in zip(trace["cand"], trace["res"]):
        if r[0] < 0.5:
            continue
        scl = int(lv[int(c[0]), 0])
        if scl >= SLOTS:
            continue
        s = scales[scl]
        u = int(np.floor((4.0 * c[1] + 11.5) / s / 8.0))
        v = int(np.floor((4.0 * c[2] + 13.5) / s / 8.0))
        y[scl, max(0, v - dilate):min(h8, v + dilate + 1), max(0, u - dilate):min(w8, u + dilate + 1)] = 1
    return y


def _gen(seed):
    cfg = TRAIN_WORKLOAD
    return synth.synth_frame(seed, cfg["w"], cfg["h"], cfg["lam"], cfg["min_size"], cfg["max_size"])[0]


def make_dataset(api, n, seed0=500000, workers=16, log=print):
    """n synthetic 1280x720 frames labelled by the plain GPU cascade: (x0 [n, 90, 160],
    labels [n, SLOTS, 90, 160], faces found)."""
    cfg = TRAIN_WORKLOAD
    det = api.FaceDetector()
    if det.Init(api.Param(max_image_size=(cfg["w"], cfg["h"]), min_face_height=MIN_FACE,
                          scale_factor=SCALE_FACTOR)) != 0:
        raise RuntimeError("FaceDetector.Init failed")
    scales = scales_for(MIN_FACE, cfg["h"], SCALE_FACTOR)
    h8, w8 = (cfg["h"] + 7) // 8, (cfg["w"] + 7) // 8
    xs = np.zeros((n, h8, w8), np.float32)
    ys = np.zeros((n, SLOTS, h8, w8), np.uint8)
    faces = 0
    t0 = time.time()
    chunk = 256
    with ProcessPoolExecutor(max_workers=workers) as ex:
        for c0 in range(0, n, chunk):
            frames = list(ex.map(_gen, range(seed0 + c0, seed0 + min(n, c0 + chunk))))
            for k, f in enumerate(frames):
                cnt, _ = det.Detect(f)
                faces += max(cnt, 0)
                tr = det.last_trace()
                xs[c0 + k] = down8(frame_gray(f))
                ys[c0 + k] = labels_from_trace(f.shape[:2], tr, scales)
            log("dataset %d/%d frames, %.0f s" % (min(n, c0 + chunk), n, time.time() - t0))
    return xs, ys, faces


def init_params(seed=1):
    """He-style initial weights, zero biases, output biases at -2 (mostly-skip prior)."""
    rng = np.random.default_rng(seed)
    parts = [rng.normal(0, np.sqrt(2.0 / 9), 8 * 9), np.zeros(8), rng.normal(0, np.sqrt(2.0 / 72), 8 * 8 * 9),
             np.zeros(8), rng.normal(0, np.sqrt(2.0 / 72), SLOTS * 8 * 9), np.full(SLOTS, -2.0)]
    return np.concatenate(parts).astype(np.float32)


def _lib(api):
    L = api.load()
    P = ctypes.c_void_p
    L.ccnn_skipnet_train.argtypes = [P, P, P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_float, ctypes.c_float, ctypes.c_uint, P]
    L.ccnn_skipnet_grad.argtypes = [P, P, P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float, P, P]
    L.ccnn_skipnet_forward.argtypes = [P, P, ctypes.c_int, ctypes.c_int, ctypes.c_int, P]
    return L


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def train(api, xs, ys, epochs=8, batch=64, lr=3e-3, pos_weight=500.0, seed=7, params=None):
    """Trains on the GPU (Adam); returns (params, loss per epoch)."""
    L = _lib(api)
    p = np.ascontiguousarray(init_params(seed) if params is None else params, np.float32).copy()
    xs = np.ascontiguousarray(xs, np.float32)
    ys = np.ascontiguousarray(ys, np.uint8)
    n, h8, w8 = xs.shape
    losses = np.zeros(epochs, np.float32)
    if L.ccnn_skipnet_train(_p(p), _p(xs), _p(ys), n, h8, w8, epochs, batch, lr, pos_weight, seed, _p(losses)) != 0:
        raise RuntimeError("ccnn_skipnet_train failed")
    return p, losses


def grad(api, params, xs, ys, pos_weight):
    """(loss, gradient) of one forward/backward pass on the GPU."""
    L = _lib(api)
    p = np.ascontiguousarray(params, np.float32)
    xs = np.ascontiguousarray(xs, np.float32)
    ys = np.ascontiguousarray(ys, np.uint8)
    g = np.zeros(NPAR, np.float32)
    loss = np.zeros(1, np.float32)
    n, h8, w8 = xs.shape
    if L.ccnn_skipnet_grad(_p(p), _p(xs), _p(ys), n, h8, w8, pos_weight, _p(g), _p(loss)) != 0:
        raise RuntimeError("ccnn_skipnet_grad failed")
    return float(loss[0]), g


def forward(api, params, xs):
    """Logits [n, SLOTS, h8, w8] on the GPU."""
    L = _lib(api)
    p = np.ascontiguousarray(params, np.float32)
    xs = np.ascontiguousarray(xs, np.float32)
    n, h8, w8 = xs.shape
    z = np.zeros((n, SLOTS, h8, w8), np.float32)
    if L.ccnn_skipnet_forward(_p(p), _p(xs), n, h8, w8, _p(z)) != 0:
        raise RuntimeError("ccnn_skipnet_forward failed")
    return z


def to_fp16(params):
    """The inference weights: the trained FP32 weights rounded to FP16."""
    return np.asarray(params, np.float32).astype(np.float16).astype(np.float32)


def save(path, params, meta):
    np.savez(path, params=np.asarray(params, np.float32), meta=json.dumps(meta))


def load(path=DEFAULT_MODEL):
    """(params float32 [NPAR], meta dict) of a model file."""
    with np.load(path, allow_pickle=False) as z:
        return np.array(z["params"], np.float32), json.loads(str(z["meta"]))


WINDOWS = ((0, 1), (1, 1), (2, 1), (0, 2), (1, 2), (2, 2))
THRESHOLDS = (0.5, 0.3, 0.2, 0.1, 0.05, 0.03, 0.02, 0.01)


def _gen_frame(args):
    seed, w, h = args
    return synth.synth_frame(seed, w, h, 10.0, 20, 300)[0]


def validation_frames(n_fhd=48, n_uhd=12, workers=16):
    """Validation frames: seeds 800000.. at 1920x1080 and 810000.. at 3840x2160 (training
    uses seeds from 500000, the tests and the benchmark from 900000)."""
    with ProcessPoolExecutor(max_workers=workers) as ex:
        fhd = list(ex.map(_gen_frame, [(800000 + i, 1920, 1080) for i in range(n_fhd)]))
        uhd = list(ex.map(_gen_frame, [(810000 + i, 3840, 2160) for i in range(n_uhd)]))
    return {"1920x1080": fhd, "3840x2160": uhd}


def sweep_operating_points(api, params, sets, windows=WINDOWS, thresholds=THRESHOLDS, chunk=16):
    """Recall against the plain cascade (boxes paired at IoU >= 0.99) and the fraction of
    stage-1 tiles run, per frame set, for every keep window (slot radius, cell margin)
    and threshold: a list of {window, threshold, per set {recall, run}}."""
    from compactcnncascade_amd import recall
    out = []
    dets = {}
    for name, frames in sets.items():
        H, W = frames[0].shape[:2]
        det = api.FaceDetector()
        if det.Init(api.Param(max_image_size=(W, H), min_face_height=MIN_FACE, scale_factor=SCALE_FACTOR)) != 0:
            raise RuntimeError("FaceDetector.Init failed")
        base = []
        for k in range(0, len(frames), chunk):
            base += [f for _, f in det.DetectBatch(images=frames[k:k + chunk], capacity=4096)]
        dets[name] = (det, base)
    for (srad, margin) in windows:
        for thr in thresholds:
            row = {"slot_radius": srad, "cell_margin": margin, "threshold": thr}
            for name, frames in sets.items():
                det, base = dets[name]
                det.set_skipnet(params, thr, MIN_FACE, SCALE_FACTOR)
                det.set_skip_window(srad, margin)
                got, tiles, kept = [], 0, 0
                for k in range(0, len(frames), chunk):
                    got += [f for _, f in det.DetectBatch(images=frames[k:k + chunk], capacity=4096)]
                    a, b = det.skip_stats()
                    tiles += a
                    kept += b
                rc = recall.recall(base, got)[0]
                row[name] = {"recall": round(rc, 5), "run": round(kept / max(1, tiles), 4)}
            out.append(row)
    return out


def choose_operating_points(sweep, target=0.995):
    """'recall': the point that reaches `target` on every set with the fewest tiles run
    (mean over the sets); 'fast': threshold 0.5 with the default window (0, 1)."""
    sets = [k for k in sweep[0] if isinstance(sweep[0][k], dict)]
    best = None
    for row in sweep:
        if all(row[k]["recall"] >= target for k in sets):
            run = sum(row[k]["run"] for k in sets) / len(sets)
            if best is None or run < best[0]:
                best = (run, row)
    pick = lambda r: {"threshold": r["threshold"], "slot_radius": r["slot_radius"], "cell_margin": r["cell_margin"]}
    fast = next(r for r in sweep if r["threshold"] == 0.5 and r["slot_radius"] == 0 and r["cell_margin"] == 1)
    return {"recall": pick(best[1]) if best else None, "fast": pick(fast), "target": target}


def apply_operating_point(det, params, meta, which="recall"):
    """Switches tile skipping on in a detector at one of the model's operating points (a
    threshold for every slot, or one per slot)."""
    op = meta["operating_points"][which]
    det.set_skipnet(params, op.get("threshold", 0.5), meta["min_face"], meta["scale_factor"])
    if "thresholds" in op:
        det.set_skip_slot_thresholds(op["thresholds"])
    det.set_skip_window(op["slot_radius"], op["cell_margin"])
    return op


def choose_slot_thresholds(api, params, sets, window, start, target=0.995,
                           grid=(0.9, 0.7, 0.5, 0.4, 0.3, 0.2, 0.15, 0.1, 0.07, 0.05, 0.03), slots=range(12),
                           chunk=16, log=None):
    """Per-slot keep thresholds, greedily: every slot starts at `start` (a threshold whose
    recall meets the target everywhere); then slot by slot, in the order given (the
    largest levels first), the largest threshold of `grid` that keeps the recall on every
    validation set at `target` or above is taken. Returns (thresholds, recall per set,
    tiles run per set)."""
    from compactcnncascade_amd import recall
    dets = {}
    for name, frames in sets.items():
        H, W = frames[0].shape[:2]
        det = api.FaceDetector()
        if det.Init(api.Param(max_image_size=(W, H), min_face_height=MIN_FACE, scale_factor=SCALE_FACTOR)) != 0:
            raise RuntimeError("FaceDetector.Init failed")
        base = []
        for k in range(0, len(frames), chunk):
            base += [f for _, f in det.DetectBatch(images=frames[k:k + chunk], capacity=4096)]
        det.set_skipnet(params, 0.5, MIN_FACE, SCALE_FACTOR)
        det.set_skip_window(*window)
        dets[name] = (det, base)

    def run(thr):
        rc, run_ = {}, {}
        for name, frames in sets.items():
            det, base = dets[name]
            det.set_skip_slot_thresholds(thr)
            got, tiles, kept = [], 0, 0
            for k in range(0, len(frames), chunk):
                got += [f for _, f in det.DetectBatch(images=frames[k:k + chunk], capacity=4096)]
                a, b = det.skip_stats()
                tiles += a
                kept += b
            rc[name] = recall.recall(base, got)[0]
            run_[name] = kept / max(1, tiles)
        return rc, run_

    thr = [float(start)] * SLOTS
    rc, run_ = run(thr)
    for s in slots:
        for t in grid:
            if t <= thr[s]:
                break
            trial = list(thr)
            trial[s] = t
            r2, u2 = run(trial)
            if all(v >= target for v in r2.values()):
                thr, rc, run_ = trial, r2, u2
                break
        if log:
            log("slot %d: threshold %.2f, recall %s, tiles run %s" % (s, thr[s], rc, run_))
    return thr, {k: round(v, 5) for k, v in rc.items()}, {k: round(v, 4) for k, v in run_.items()}


def validate(api, params, meta, n_fhd=48, n_uhd=12, log=None):
    """Sweeps the operating points of a trained network on the validation frames and
    records the sweep and the chosen points in its meta: 'recall' and 'fast' with a single
    threshold for all slots over keep windows, and 'recall_per_slot', per-slot thresholds
    with the window of 'recall' (greedy on the same frames: it runs fewer tiles there but
    fits those frames, so its held-out recall is lower)."""
    sets = validation_frames(n_fhd, n_uhd)
    p16 = to_fp16(params)
    sweep = sweep_operating_points(api, p16, sets)
    ops = choose_operating_points(sweep)
    if ops["recall"]:
        uni = ops["recall"]
        thr, rc, run_ = choose_slot_thresholds(api, p16, sets, (uni["slot_radius"], uni["cell_margin"]),
                                               uni["threshold"], target=ops["target"], log=log)
        row = next(r for r in sweep if r["threshold"] == uni["threshold"] and r["slot_radius"] == uni["slot_radius"]
                   and r["cell_margin"] == uni["cell_margin"])
        uni_run = sum(row[k]["run"] for k in sets) / len(sets)
        ops["recall_per_slot"] = {"thresholds": [round(t, 4) for t in thr], "slot_radius": uni["slot_radius"],
                                  "cell_margin": uni["cell_margin"], "validation_recall": rc, "validation_run": run_,
                                  "validation_run_single_threshold": round(uni_run, 4)}
    meta["operating_points"] = ops
    meta["validation"] = {"frames": {k: len(v) for k, v in sets.items()}, "seeds": {"1920x1080": 800000,
                                                                                   "3840x2160": 810000},
                          "sweep": sweep}
    meta["threshold"] = meta["operating_points"]["fast"]["threshold"]
    return meta


def evaluate(api, params, meta, n=256, chunk=32, sizes=((1920, 1080), (3840, 2160)), seed0=900000, log=print):
    """Held-out evaluation (seeds from 900000, generated chunk by chunk): per frame size,
    the plain FP32 cascade against tile skipping at each operating point -- recall
    (boxes paired at IoU >= 0.99, summed over the frames), tiles run, and the batch time
    of each (DetectBatch over `chunk` frames from host memory, upload included)."""
    from compactcnncascade_amd import recall
    out = {}
    for (W, H) in sizes:
        name = "%dx%d" % (W, H)
        dets = {}
        points = [k for k in ("recall", "fast", "recall_per_slot") if meta["operating_points"].get(k)]
        for key in ["plain"] + points:
            d = api.FaceDetector()
            if d.Init(api.Param(max_image_size=(W, H), min_face_height=MIN_FACE, scale_factor=SCALE_FACTOR)) != 0:
                raise RuntimeError("FaceDetector.Init failed")
            if key != "plain":
                apply_operating_point(d, params, meta, key)
            dets[key] = d
        acc = {k: {"paired": 0, "ref": 0, "test": 0, "tiles": 0, "run": 0, "seconds": 0.0} for k in dets}
        with ProcessPoolExecutor(max_workers=16) as ex:
            for c0 in range(0, n, chunk):
                frames = list(ex.map(_gen_frame, [(seed0 + c0 + i, W, H) for i in range(min(chunk, n - c0))]))
                res = {}
                for key, d in dets.items():
                    d.DetectBatch(images=frames, capacity=4096)          # capture / warm-up
                    t0 = time.perf_counter()
                    res[key] = [f for _, f in d.DetectBatch(images=frames, capacity=4096)]
                    acc[key]["seconds"] += time.perf_counter() - t0
                    a, b = d.skip_stats()
                    acc[key]["tiles"] += a
                    acc[key]["run"] += b
                for key in points:
                    _, paired, nref, ntest = recall.recall(res["plain"], res[key])
                    acc[key]["paired"] += paired
                    acc[key]["ref"] += nref
                    acc[key]["test"] += ntest
                log("%s: %d/%d frames" % (name, c0 + len(frames), n))
        row = {"frames": n, "plain_fps": round(n / acc["plain"]["seconds"], 1)}
        for key in points:
            a = acc[key]
            row[key] = {"recall": round(a["paired"] / max(1, a["ref"]), 5), "paired": [a["paired"], a["ref"], a["test"]],
                        "tiles_run_frac": round(a["run"] / max(1, a["tiles"]), 4),
                        "fps": round(n / a["seconds"], 1),
                        "gain_over_plain": round(acc["plain"]["seconds"] / a["seconds"], 3),
                        "point": meta["operating_points"][key]}
        out[name] = row
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("cmd", choices=["train", "validate", "evaluate"])
    ap.add_argument("--held-out", type=int, default=256, help="evaluate: frames per size")
    ap.add_argument("--frames", type=int, default=20000)
    ap.add_argument("--epochs", type=int, default=12)
    ap.add_argument("--pos-weight", type=float, default=500.0)
    ap.add_argument("--val-fhd", type=int, default=48)
    ap.add_argument("--val-uhd", type=int, default=12)
    ap.add_argument("--model", default=DEFAULT_MODEL, help="model to validate")
    ap.add_argument("--out", default=DEFAULT_MODEL)
    args = ap.parse_args()
    from compactcnncascade_amd import api
    if args.cmd == "evaluate":
        p, meta = load(args.model)
        print(json.dumps(evaluate(api, to_fp16(p), meta, args.held_out,
                                  log=lambda m: print(m, file=sys.stderr, flush=True))), flush=True)
        return
    if args.cmd == "validate":
        p, meta = load(args.model)
    else:
        t0 = time.time()
        xs, ys, faces = make_dataset(api, args.frames)
        print(json.dumps({"dataset_frames": args.frames, "faces": faces, "positive_fraction": float(ys.mean()),
                          "seconds": round(time.time() - t0, 1)}), flush=True)
        t1 = time.time()
        p, losses = train(api, xs, ys, epochs=args.epochs, pos_weight=args.pos_weight)
        train_s = round(time.time() - t1, 1)
        del xs, ys
        meta = {"min_face": MIN_FACE, "scale_factor": SCALE_FACTOR, "slots": SLOTS, "frames": args.frames,
                "workload": TRAIN_WORKLOAD, "epochs": args.epochs, "pos_weight": args.pos_weight,
                "loss": [float(v) for v in losses], "train_seconds": train_s}
    meta = validate(api, p, meta, args.val_fhd, args.val_uhd, log=lambda m: print(m, file=sys.stderr, flush=True))
    save(args.out, p, meta)
    print(json.dumps(meta["operating_points"]), flush=True)


if __name__ == "__main__":
    main()
