"""Benchmark of the MI355X compact-CNN-cascade detection path.

Default workload (one GPU): 64 seeded 1920x1080 BGR frames, scale factor 1.15, faces
40..240 px, default FaceDetector parameters, frames resident in GPU memory before the
timed region. A step is one DetectBatch pass over the 64 frames (per-frame launches,
consecutive frames pipelined over three streams); `value` is frames/s over K steps.
Beside it the line reports one synchronous Detect at a time, per-kernel rooflines from a
single-frame pass, the 4K / mixed-size / 640x480 / deep-pyramid / facial-analysis legs,
the headline frames handed over in host memory (PCIe upload included) and the
reference's CPU path on a bounded sample. Prints one JSON line.

    python bench.py [--steps K] [--warmup W] [--frames F] [--gpus 1] [--facial]
                    [--no-secondary] [--no-cpu-baseline]
"""
import argparse
import ctypes
import json
import os
import sys
import time
from concurrent.futures import ProcessPoolExecutor

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from compactcnncascade_amd import synth  # noqa: E402

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
# The exact path evaluates every multiply and add as its own rounded VALU op (no FMA,
# no packed FP32): 256 CUs x 4 SIMD x 16 lanes x 2.4 GHz = 39.3 T ops/s.
VALU_EXACT_PEAK_TOPS = 39.3


def _frame(args):
    name, idx = args
    return synth.workload_frame(name, idx)[0]


def make_frames(name, n):
    workers = min(16, max(1, (os.cpu_count() or 1)))
    with ProcessPoolExecutor(max_workers=workers) as ex:
        return np.stack(list(ex.map(_frame, [(name, i) for i in range(n)])))


def secondary_fps(api, torch, names, per_name, steps, facial=False):
    """Frames/s of DetectBatch over frames of the named workloads (all frames resident in
    GPU memory, one detector initialised for the largest frame), best of `steps` passes;
    facial: with facial analysis."""
    frames = []
    for name in names:
        frames.extend(make_frames(name, per_name))
    dev = [torch.from_numpy(f).to("cuda:0") for f in frames]
    torch.cuda.synchronize()
    cfg = synth.WORKLOADS[names[-1]]
    W = max(f.shape[1] for f in frames)
    H = max(f.shape[0] for f in frames)
    det = api.FaceDetector()
    if det.Init(api.Param(max_image_size=(W, H), min_face_height=cfg["min_face"], max_face_height=cfg["max_face"],
                          scale_factor=cfg["scale_factor"], facial_analysis=facial)) != 0:
        raise SystemExit("FaceDetector.Init failed")
    df = [(t.data_ptr(), f.shape[1], f.shape[0], 3, f.shape[1] * 3) for t, f in zip(dev, frames)]
    det.DetectBatch(device_frames=df)
    best = None
    faces = 0
    for _ in range(steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = det.DetectBatch(device_frames=df)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        faces = sum(n for n, _ in out)
        best = dt if best is None else min(best, dt)
    return {"fps": round(len(frames) / best, 2), "frames": len(frames), "faces": faces,
            "sizes": sorted({"%dx%d" % (f.shape[1], f.shape[0]) for f in frames}),
            "min_face": cfg["min_face"], "scale_factor": cfg["scale_factor"]}


def host_frame_fps(api, torch, frames, param):
    """Frames/s with the frames in host memory, the PCIe upload included (not part of
    `value`): pageable and pinned buffers, through DetectBatch (best of 3 passes) and one
    synchronous Detect at a time."""
    det = api.FaceDetector()
    if det.Init(param) != 0:
        raise SystemExit("FaceDetector.Init failed")
    pageable = [np.ascontiguousarray(f) for f in frames]
    pinned = [t.numpy() for t in (torch.from_numpy(f).pin_memory() for f in frames)]
    out = {}
    for name, imgs in (("pageable", pageable), ("pinned", pinned)):
        det.DetectBatch(images=imgs)
        best = None
        for _ in range(3):
            t0 = time.perf_counter()
            det.DetectBatch(images=imgs)
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
        t0 = time.perf_counter()
        for im in imgs:
            det.Detect(im)
        out[name] = {"batch_fps": round(len(imgs) / best, 1),
                     "sync_fps": round(len(imgs) / (time.perf_counter() - t0), 1)}
    return out


def stage1_macs(levels):
    """Algorithmic stage-1 MACs of one frame, split as the two stage-1 kernels compute
    them: (L1: 64 per conv-1 output position, i.e. 4 maps x 16 taps; rest: 72 per
    conv-2 output position (8 maps x 9 taps), 320 (L3) + 260 (hidden + output layer)
    per response)."""
    l1 = rest = 0
    for (_, lw, lh, rw, rh, _) in levels:
        c1, r1 = lw - 3, lh - 3
        c2, r2 = (c1 >> 1) - 2, (r1 >> 1) - 2
        l1 += 64 * c1 * r1
        rest += 72 * c2 * r2 + (320 + 260) * rw * rh
    return l1, rest


STAGE2_MAC = 278496
STAGE3_MAC = 276672


def check_evals(res):
    """Stage-2 and stage-3 network evaluations the check kernel performed (per-candidate
    counts recorded by the kernel)."""
    if len(res) == 0:
        return 0, 0
    return int(res[:, 6].sum()), int(res[:, 7].sum())


def cpu_baseline(frames, params, budget_s):
    """The reference's CPU path compiled from its own sources at -O3 (oracle/_ref),
    single thread, over as many of the workload's frames as fit in budget_s."""
    try:
        from oracle.bind import Ref, default_params
        ref = Ref("O3")
    except Exception as e:  # noqa: BLE001
        return {"value": None, "unit": "frames/s", "cores": 1, "kind": "reference",
                "sample": "unavailable: %s" % e}
    h, w = frames.shape[1:3]
    p = default_params(w, h, **params)
    done = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget_s:
        i = done % len(frames)
        ref.detect_loop(frames[i:i + 1], p, 1)
        done += 1
    dt = time.perf_counter() - t0
    return {"value": done / dt, "unit": "frames/s", "cores": 1, "kind": "reference",
            "sample": "%d detections over the workload's first %d %dx%d frames (cycled), one thread, %.1f s"
                      % (done, min(done, len(frames)), w, h, dt)}


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload` from the rocprofv3 FETCH_SIZE and
    WRITE_SIZE passes summarised under profiles/ (line `bytes_per_launch N`), or None."""
    path = os.path.join(ROOT, "profiles", "%s_%s_hbm_traffic.txt" % (kernel, workload))
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("bytes_per_launch"):
                    return int(float(line.split()[1]))
    except OSError:
        pass
    return None


def calibrate_copy(mib):
    """A device-to-device copy of `mib` MiB, five times: the known-size kernel whose
    FETCH_SIZE / WRITE_SIZE readings scale the counters of the detection kernels."""
    import torch
    n = mib << 18
    a = torch.arange(n, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"calibration_copy_bytes": n * 4, "launches": 5}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--frames", type=int, default=64)
    ap.add_argument("--workload", default="fhd_min40_max240")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-secondary", action="store_true",
                    help="skip the 4K and mixed-size batch measurements reported beside `value`")
    ap.add_argument("--facial", action="store_true",
                    help="run the workload with facial analysis (the line's config says so)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a device copy of MIB MiB (counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: there is no multi-GPU path here (frames would shard "
              "across GPUs as independent detectors, one process per GPU). Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return

    import torch
    from compactcnncascade_amd import api

    cfg = synth.WORKLOADS[args.workload]
    W, H = cfg["w"], cfg["h"]
    frames = make_frames(args.workload, args.frames)
    dev = torch.device("cuda:0")
    d_frames = torch.from_numpy(frames).to(dev)
    torch.cuda.synchronize()
    base = d_frames.data_ptr()
    fbytes = W * H * 3

    det = api.FaceDetector()
    param = api.Param(max_image_size=(W, H), min_face_height=cfg["min_face"], max_face_height=cfg["max_face"],
                      scale_factor=cfg["scale_factor"], facial_analysis=args.facial)
    if det.Init(param) != 0:
        raise SystemExit("FaceDetector.Init failed")

    dframes = [(base + i * fbytes, W, H, 3, W * 3) for i in range(args.frames)]

    def one_pass():
        out = det.DetectBatch(device_frames=dframes)
        return sum(n for n, _ in out)

    # record of the work per frame (untimed): candidates and stage-2/3 evaluations
    records = []
    for i in range(args.frames):
        k, _ = det.DetectDevice(ctypes.c_void_p(base + i * fbytes), W, H, 3, W * 3)
        if k < 0:
            raise SystemExit("Detect failed")
        records.append(det.last_trace())
    # single-frame pass: one Detect call at a time, so each frame's kernels have the GPU
    # to themselves; their durations (HIP events stamped by the launches) give the
    # per-kernel rooflines
    det.set_profiling(True)
    for i in range(args.frames):
        det.DetectDevice(ctypes.c_void_p(base + i * fbytes), W, H, 3, W * 3)
    tm = det.last_timing()
    det.set_profiling(False)
    # the same pass without the events, timed as sync_single_frame_fps (not part of `value`)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.frames):
        det.DetectDevice(ctypes.c_void_p(base + i * fbytes), W, H, 3, W * 3)
    sync_fps = args.frames / (time.perf_counter() - t0)

    # the timed region: K passes of pipelined batch detection over the frames
    for _ in range(args.warmup):
        one_pass()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    faces = 0
    for _ in range(args.steps):
        faces += one_pass()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    nframes = args.steps * args.frames
    fps = nframes / dt

    # algorithmic work of the recorded pass
    s1 = np.array([stage1_macs(tr["levels"]) for tr in records], np.float64).mean(0)
    s1_mac = s1.sum()
    e2 = e3 = 0
    cands = 0
    for tr in records:
        a, b = check_evals(tr["res"])
        e2 += a
        e3 += b
        cands += len(tr["cand"])
    e2 /= len(records)
    e3 /= len(records)
    per_frame_ms = {k: tm["totals"][k] / max(1, tm["frames"])
                    for k in ("gray", "stage1_l1", "stage1", "gap", "check", "after_check")}
    # algorithmic FLOP (2 per MAC) per launch of the three VALU kernels, and what their
    # event-timed launches achieve
    kflops = {"k_stage1_l1": 2.0 * s1[0], "k_stage1": 2.0 * s1[1],
              "k_check": 2.0 * (e2 * STAGE2_MAC + e3 * STAGE3_MAC)}
    kms = {"k_stage1_l1": per_frame_ms["stage1_l1"], "k_stage1": per_frame_ms["stage1"],
           "k_check": per_frame_ms["check"]}
    kernels = {k: {"ms": round(kms[k], 4), "gflop": round(kflops[k] / 1e9, 4),
                   "tflops": round(kflops[k] / (kms[k] * 1e-3) / 1e12, 3),
                   "frac_of_exact_valu": round(kflops[k] / (kms[k] * 1e-3) / 1e12 / VALU_EXACT_PEAK_TOPS, 3)}
               for k in kflops}
    dominant = max(kms, key=lambda k: kms[k])
    achieved = kflops[dominant] / (kms[dominant] * 1e-3) / 1e12
    in_bytes = W * H * 3
    line = {
        "metric": "1080p frames/sec (plain FP32 cascade, bit-exact with the reference CPU path)",
        "value": round(fps, 2),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded generator, compactcnncascade_amd/synth.py)",
        "config": {"workload": args.workload, "frames_per_step": args.frames, "frame": "%dx%d BGR u8" % (W, H),
                   "scale_factor": cfg["scale_factor"], "min_face": cfg["min_face"], "max_face": cfg["max_face"],
                   "facial_analysis": args.facial,
                   "launch": "per-frame launches, consecutive frames pipelined over three streams, "
                             "frames resident in HBM"},
        "roofline": {
            "kernel": dominant,
            "bound": "mfma",
            "achieved": round(achieved, 3),
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / FP32_PEAK_TFLOPS, 4),
            "traffic": measured_traffic(dominant, args.workload),
            "measured_in": "single-frame pass (one frame's kernels alone on the GPU), HIP start/stop events "
                           "stamped by the kernel launches",
            "note": "FP32 VALU kernel on the exact path (separate mul/add, no packed FP32); its own ceiling "
                    "is %.1f T ops/s, frac of that %.3f" % (VALU_EXACT_PEAK_TOPS, achieved / VALU_EXACT_PEAK_TOPS),
        },
        "kernel_ms_per_frame": {k: round(v, 4) for k, v in per_frame_ms.items()},
        "valu_kernels": kernels,
        "work_per_frame": {"stage1_gmac": round(s1_mac / 1e9, 4), "candidates": round(cands / len(records), 1),
                           "stage2_evals": round(e2, 1), "stage3_evals": round(e3, 1)},
        "faces_per_frame": round(faces / nframes, 2),
        "sync_single_frame_fps": round(sync_fps, 1),
        "hbm_input_gbs": round(in_bytes * fps / 1e9, 3),
    }
    if not args.no_secondary:
        # the metric's 4K leg and the mixed-size batch (not part of `value`)
        line["uhd_3840x2160"] = secondary_fps(api, torch, ["mixed_uhd"], 8, 3)
        line["mixed_batch"] = secondary_fps(api, torch, ["mixed_vga", "mixed_hd", "mixed_fhd", "mixed_uhd"], 8, 3)
        # the reference's CPU-runnable case (640x480, faces >= 40) and the deep pyramid
        # (4800x2400, scale 1.05, faces >= 15: 91 levels)
        line["vga_640x480"] = secondary_fps(api, torch, ["vga_min40"], 64, 3)
        line["deep_4800x2400"] = secondary_fps(api, torch, ["deep_4800x2400_sf105_min15"], 2, 2)
        # the headline workload with facial analysis (landmarks network on accepted windows)
        line["fhd_facial"] = secondary_fps(api, torch, ["fhd_min40_max240"], 64, 3, facial=True)
        # the headline frames handed over in host memory: the upload crosses PCIe
        line["fhd_host_frames"] = host_frame_fps(api, torch, frames, param)
    if not args.no_cpu_baseline:
        cb = cpu_baseline(frames, dict(min_face=cfg["min_face"], max_face=cfg["max_face"],
                                       scale_factor=cfg["scale_factor"]), args.cpu_seconds)
        line["cpu_baseline"] = cb
        if cb["value"]:
            line["gpu_over_cpu"] = round(fps / cb["value"], 1)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
