"""Benchmark of the MI355X compact-CNN-cascade detection path.

Headline (`value`): 64 seeded 1920x1080 BGR frames, scale factor 1.15, faces 40..240 px,
default FaceDetector parameters, frames resident in GPU memory before the timed region.
A step is one DetectBatch call over the 64 frames on the exact FP32 path (bit-identical
to the reference CPU path): the batch graph (one captured HIP graph launch per batch:
gray, stage 1 over every (frame, level, tile), ordered candidate compaction, stages 2/3
on a persistent grid) or per-frame launches pipelined over three streams, whichever is
faster on the box (both are reported). `value` = frames/s over K steps.

Beside it, the line reports: one synchronous Detect at a time; per-kernel rooflines from
a batch pass (each kernel launched once over the 64 frames, HIP events between them) and
from a single-frame pass; the configurations of the benchmark set -- the 4K leg, the 32-frame
mixed-size batch (FP32, FP16/BF16, FP16/FP8, launches per batch, recall against FP32),
the 4800x2400 deep pyramid (FP32/BF16/FP8 with recall), learned tile skipping on held-out
1920x1080 and 3840x2160 frames with faces >= 20 (plain, skip, skip + FP16/BF16, with
recall, tiles run and the gain over the plain FP32 cascade), 640x480, facial analysis,
host frames over PCIe -- and the reference's CPU path on the host's cores (one core, and
one process per core). Prints one JSON line.

    python bench.py [--steps K] [--warmup W] [--frames F] [--gpus 1] [--facial]
                    [--no-secondary] [--no-cpu-baseline] [--profile-pass N]
                    [--phase-pass check|stage1]
"""
import argparse
import ctypes
import json
import os
import subprocess
import sys
import time
from concurrent.futures import ProcessPoolExecutor

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from compactcnncascade_amd import synth  # noqa: E402

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
BF16_PEAK_TFLOPS = 2500.0
FP8_PEAK_TFLOPS = 5000.0
# The exact path evaluates every multiply and add as its own rounded VALU op (no FMA,
# no packed FP32), so it reaches at most half the FP32 peak (which counts an FMA as two
# FLOP): 256 CUs x 4 SIMD x 32 FP32 lanes per clock x 2.4 GHz = 78.6 T ops/s. (The SQ
# counters confirm the 32 lanes per clock: in a 1080p batch k_stage1_l1 issues 0.32
# wave64 VALU instructions per SIMD per clock, above the 0.25 that 16 lanes would allow.)
VALU_EXACT_PEAK_TOPS = 78.6
STAGE2_MAC = 278496
STAGE3_MAC = 276672
# matrix-core stages 2/3: per network evaluation L1 takes 33 row tiles x 2 column tiles and
# L2 7 row tiles x 3 sum pairs of 16x16x32 products, each as three MFMAs (two-term
# operands): 261 MFMAs of 16 x 16 x 32 x 2 FLOP
MFMA_FLOP_PER_EVAL = 261 * 16 * 16 * 32 * 2


def progress(msg):
    """One line on stderr per finished part of the run (the JSON line goes to stdout)."""
    print("[bench] %s" % msg, file=sys.stderr, flush=True)


def _frame(args):
    name, idx = args
    return synth.workload_frame(name, idx)[0]


def _frame_seed(args):
    seed, w, h = args
    return synth.synth_frame(seed, w, h, 10.0, 20, 300)[0]


def make_frames(name, n):
    workers = min(16, max(1, (os.cpu_count() or 1)))
    with ProcessPoolExecutor(max_workers=workers) as ex:
        return np.stack(list(ex.map(_frame, [(name, i) for i in range(n)])))


def held_out_frames(w, h, n, seed0):
    """Seeded frames outside every training set (seeds from 900000; the skip network was
    trained on seeds from 500000)."""
    workers = min(16, max(1, (os.cpu_count() or 1)))
    with ProcessPoolExecutor(max_workers=workers) as ex:
        return list(ex.map(_frame_seed, [(seed0 + i, w, h) for i in range(n)]))


def to_device(torch, frames):
    dev = [torch.from_numpy(np.ascontiguousarray(f)).to("cuda:0") for f in frames]
    torch.cuda.synchronize()
    return dev, [(t.data_ptr(), f.shape[1], f.shape[0], 3, f.shape[1] * 3) for t, f in zip(dev, frames)]


def timed_batches(torch, det, df, passes):
    """Best of `passes` DetectBatch calls (after one untimed call): (seconds, results)."""
    det.DetectBatch(device_frames=df)
    torch.cuda.synchronize()
    best, out = None, None
    for _ in range(passes):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = det.DetectBatch(device_frames=df)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        best = dt if best is None else min(best, dt)
    return best, [f for _, f in out]


def detector(api, W, H, min_face, max_face, sf, precision=0, mode=1, facial=False):
    det = api.FaceDetector()
    if det.Init(api.Param(max_image_size=(W, H), min_face_height=min_face, max_face_height=max_face,
                          scale_factor=sf, facial_analysis=facial)) != 0:
        raise SystemExit("FaceDetector.Init failed")
    det.set_compute_precision(precision)
    det.set_batch_mode(mode)
    return det


def secondary_fps(api, torch, names, per_name, steps, facial=False, precisions=(0,), mode=1):
    """Frames/s of DetectBatch over frames of the named workloads (resident in GPU memory,
    one detector initialised for the largest frame), best of `steps` passes, per compute
    precision; recall of each reduced precision against FP32."""
    from compactcnncascade_amd import recall
    frames = []
    for name in names:
        frames.extend(make_frames(name, per_name))
    dev, df = to_device(torch, frames)
    cfg = synth.WORKLOADS[names[-1]]
    W = max(f.shape[1] for f in frames)
    H = max(f.shape[0] for f in frames)
    out = {"frames": len(frames), "sizes": sorted({"%dx%d" % (f.shape[1], f.shape[0]) for f in frames}),
           "min_face": cfg["min_face"], "scale_factor": cfg["scale_factor"]}
    base = None
    evals = None
    if any(p in (3, 4) for p in precisions):
        # the stage-2/3 evaluations the check kernel executes over these frames (the exact
        # stage 1 of precisions 3/4 gives the same candidates), from per-frame traces
        det = detector(api, W, H, cfg["min_face"], cfg["max_face"], cfg["scale_factor"], 0, mode, facial)
        evals = 0
        for (ptr, w, h, c, st) in df:
            det.DetectDevice(ctypes.c_void_p(ptr), w, h, c, st)
            r = det.last_trace()["res"]
            evals += int(r[:, 6].sum() + r[:, 7].sum()) if len(r) else 0
    for p in precisions:
        det = detector(api, W, H, cfg["min_face"], cfg["max_face"], cfg["scale_factor"], p, mode, facial)
        dt, res = timed_batches(torch, det, df, steps)
        key = ["fp32", "fp16_bf16", "fp16_fp8", "fp32_bf16", "fp32_fp8"][p]
        leg = {"fps": round(len(frames) / dt, 2), "faces": int(sum(len(r) for r in res))}
        if mode == 1:
            leg["launches_per_batch"] = det.batch_stats()[0]
        if mode == 1 and p in (3, 4) and evals:
            # matrix-core roofline of the check kernel: one profiled batch (kernels without
            # the graph, HIP events between them)
            det.set_profiling(True)
            det.DetectBatch(device_frames=df)
            bt = det.last_batch_timing()
            det.set_profiling(False)
            flop = evals * MFMA_FLOP_PER_EVAL
            tf = flop / (bt["check"] * 1e-3) / 1e12
            peak = BF16_PEAK_TFLOPS if p == 3 else FP8_PEAK_TFLOPS
            leg["mfma_roofline"] = {"kernel": "k_check", "bound": "mfma", "evaluations": evals,
                                    "mfma_tflop": round(flop / 1e12, 4), "ms": round(bt["check"], 4),
                                    "achieved": round(tf, 2), "peak": peak, "unit": "TFLOP/s",
                                    "frac": round(tf / peak, 5),
                                    "dtype": "bf16" if p == 3 else "fp8 e4m3"}
        if p == 0:
            base = res
        elif base is not None:
            rc, paired, nref, ntest = recall.recall(base, res)
            leg["recall_vs_fp32"] = round(rc, 5)
            leg["paired"] = [paired, nref, ntest]
        out[key] = leg
    out["fps"] = out["fp32"]["fps"] if "fp32" in out else None
    del dev
    return out


def skip_leg(api, torch, W, H, n, steps):
    """Learned tile skipping on held-out frames (faces >= 20, scale 1.15): the plain batch
    at FP32 and FP16/BF16 against skipping with the committed network at its two
    operating points ('recall': chosen on validation frames for recall >= 0.995 against
    the plain cascade; 'fast': threshold 0.5), each at FP32 and FP16/BF16; recall of each
    against the plain FP32 cascade, tiles run."""
    from compactcnncascade_amd import recall, skipnet
    try:
        params, meta = skipnet.load()
    except OSError as e:
        return {"unavailable": str(e)}
    frames = held_out_frames(W, H, n, 900000)
    dev, df = to_device(torch, frames)
    out = {"frames": n, "frame": "%dx%d" % (W, H), "min_face": 20, "scale_factor": 1.15,
           "network": {k: meta[k] for k in ("frames", "epochs", "slots") if k in meta},
           "operating_points": meta.get("operating_points")}
    base = None
    for key, p, op in (("plain_fp32", 0, None), ("skip_recall_fp32", 0, "recall"), ("skip_fast_fp32", 0, "fast"),
                       ("plain_fp16_bf16", 1, None), ("skip_recall_fp16_bf16", 1, "recall"),
                       ("skip_fast_fp16_bf16", 1, "fast")):
        det = detector(api, W, H, 20, 0, 1.15, p, 1)
        if op is not None:
            if not meta.get("operating_points", {}).get(op):
                continue
            skipnet.apply_operating_point(det, params, meta, op)
        dt, res = timed_batches(torch, det, df, steps)
        leg = {"fps": round(n / dt, 2), "faces": int(sum(len(r) for r in res))}
        if op is not None:
            tiles, kept = det.skip_stats()
            leg.update(tiles=tiles, tiles_run=kept, tiles_run_frac=round(kept / max(1, tiles), 4))
        if base is None:
            base = res
        else:
            rc, paired, nref, ntest = recall.recall(base, res)
            leg["recall_vs_plain_fp32"] = round(rc, 5)
            leg["paired"] = [paired, nref, ntest]
        out[key] = leg
    for key in ("skip_recall_fp16_bf16", "skip_fast_fp16_bf16"):
        if key in out:
            out["gain_" + key] = round(out[key]["fps"] / out["plain_fp32"]["fps"], 3)
    del dev
    return out


def host_frame_fps(api, torch, frames, param):
    """Frames/s with the frames in host memory, the PCIe upload included (not part of
    `value`): pageable and pinned buffers, through DetectBatch in both batch modes (best
    of 3 passes) and one synchronous Detect at a time."""
    det = api.FaceDetector()
    if det.Init(param) != 0:
        raise SystemExit("FaceDetector.Init failed")
    pageable = [np.ascontiguousarray(f) for f in frames]
    pinned = [t.numpy() for t in (torch.from_numpy(f).pin_memory() for f in frames)]
    out = {}
    for name, imgs in (("pageable", pageable), ("pinned", pinned)):
        leg = {}
        for mode, key in ((1, "graph_batch_fps"), (0, "pipelined_batch_fps")):
            det.set_batch_mode(mode)
            det.DetectBatch(images=imgs)
            best = None
            for _ in range(3):
                t0 = time.perf_counter()
                det.DetectBatch(images=imgs)
                dt = time.perf_counter() - t0
                best = dt if best is None else min(best, dt)
            leg[key] = round(len(imgs) / best, 1)
        t0 = time.perf_counter()
        for im in imgs:
            det.Detect(im)
        leg["sync_fps"] = round(len(imgs) / (time.perf_counter() - t0), 1)
        out[name] = leg
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


def check_evals(res, status):
    """Stage-2 and stage-3 network evaluations: (those the reference performs -- of the
    candidates its ordered replay checks, status 2 --, all the check kernel performed)."""
    if len(res) == 0:
        return (0, 0), (0, 0)
    ev = status == 2
    return (int(res[ev, 6].sum()), int(res[ev, 7].sum())), (int(res[:, 6].sum()), int(res[:, 7].sum()))


def _cpu_worker(args):
    frames, params, budget = args
    from oracle.bind import Ref, default_params
    ref = Ref("O3")
    h, w = frames.shape[1:3]
    p = default_params(w, h, **params)
    done = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget:
        i = done % len(frames)
        ref.detect_loop(frames[i:i + 1], p, 1)
        done += 1
    return done, time.perf_counter() - t0


def cpu_model():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name"):
                return line.split(":", 1)[1].strip()
    except Exception:  # noqa: BLE001
        pass
    return None


def cpu_baseline(frames, params, budget_s):
    """The reference's CPU path compiled from its own sources at -O3 (oracle/_ref) on the
    workload's frames: one thread (latency), then one process per core (throughput)."""
    try:
        from oracle.bind import Ref  # noqa: F401
    except Exception as e:  # noqa: BLE001
        return {"value": None, "unit": "frames/s", "cores": 1, "kind": "reference", "sample": "unavailable: %s" % e}
    sub = frames[:8]
    done, dt = _cpu_worker((sub, params, budget_s))
    one = done / dt
    cores = os.cpu_count() or 1
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    workers = max(1, cores)
    with ProcessPoolExecutor(max_workers=workers) as ex:
        res = list(ex.map(_cpu_worker, [(sub[k % len(sub):k % len(sub) + 1], params, budget_s / 2)
                                        for k in range(workers)]))
    allc = sum(d / t for d, t in res)
    return {"value": round(allc, 2), "unit": "frames/s", "cores": workers, "kind": "reference",
            "one_core_fps": round(one, 2), "one_core_ms_per_frame": round(1e3 / one, 1),
            "cpu_model": cpu_model(),
            "sample": "reference CPU path (-O3 -ffast-math, single-threaded as its Linux build is) on the first %d "
                      "workload frames: %d detections in %.1f s on one core; then one process per core (%d) for "
                      "%.1f s each" % (len(sub), done, dt, workers, budget_s / 2)}


def measured(path_name):
    """Counter summaries under profiles/ (lines `key value`)."""
    path = os.path.join(ROOT, "profiles", path_name)
    out = {}
    try:
        with open(path) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 2:
                    try:
                        out[parts[0]] = float(parts[1])
                    except ValueError:
                        pass
    except OSError:
        pass
    return out


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


def pmc_summary(db, calib_db, counter, frames):
    """Per-kernel summary of one rocprofv3 counter pass (a rocpd database): dispatches,
    average counter value (KiB per dispatch) and, scaled by what the same counter reads
    for the calibration copy (bench.py --calibrate-copy 256: 256 MiB read and written per
    dispatch), bytes per dispatch and per frame of the pass."""
    import sqlite3

    def per_kernel(path):
        con = sqlite3.connect(path)
        rows = con.execute("select kernel_name, value from counters_collection where counter_name = ?",
                           (counter,)).fetchall()
        agg = {}
        for name, v in rows:
            a = agg.setdefault(name.split("(")[0], [0, 0.0])
            a[0] += 1
            a[1] += float(v)
        return agg

    cal = per_kernel(calib_db)
    copy = max(cal.items(), key=lambda kv: kv[1][1] / kv[1][0])
    copy_kib = copy[1][1] / copy[1][0]
    factor = (256 * 1024) / copy_kib
    agg = per_kernel(db)
    total = 0.0
    print("# %s per dispatch; calibration: %s reads %.0f KiB for 262144 KiB moved -> factor %.4f"
          % (counter, copy[0], copy_kib, factor))
    print("# bytes_per_frame: the detection kernels (ccnn::) only; runtime copies and fills listed apart")
    for name, (n, v) in sorted(agg.items(), key=lambda kv: -kv[1][1]):
        b = v / n * factor * 1024
        if "ccnn::" in name:
            total += b * n
        print("kernel %-50s dispatches %6d avg_kib %12.1f bytes_per_dispatch %14.0f" % (name[:50], n, v / n, b))
    print("bytes_total %.0f" % total)
    print("bytes_per_frame %.0f" % (total / max(1, frames)))


def profile_pass(n, mode, workload="fhd_min40_max240", precision=0):
    """Only the detection of the first n frames of a workload, one call per pass (for
    rocprofv3 counter runs): mode 'single' = one Detect per frame, 'batch' = one
    DetectBatch (graph) over the n frames; precision as set_compute_precision."""
    import torch
    from compactcnncascade_amd import api
    cfg = synth.WORKLOADS[workload]
    frames = make_frames(workload, n)
    dev, df = to_device(torch, frames)
    det = detector(api, cfg["w"], cfg["h"], cfg["min_face"], cfg["max_face"], cfg["scale_factor"], precision)
    if mode == "batch":
        det.DetectBatch(device_frames=df)
    else:
        for (p, w, h, c, s) in df:
            det.DetectDevice(ctypes.c_void_p(p), w, h, c, s)
    torch.cuda.synchronize()
    print(json.dumps({"profile_pass": n, "mode": mode, "workload": workload, "precision": precision}), flush=True)


def phase_pass(kind, workload="fhd_min40_max240"):
    """Per-phase latency from the kernels' phase timestamps (the profiling build
    libccnn_phases.so): 'check' = the k_check items of 8 single 1080p frames (one Detect
    each, after 16 untimed ones), 'stage1' = the k_stage1 tiles of one 64-frame batch (the
    graph path, after 3 untimed batches). Prints one JSON line of per-phase microseconds."""
    import torch
    torch.zeros(1, device="cuda:0")
    from compactcnncascade_amd import api
    api.LIB_PATH = api.PHASES_LIB_PATH
    if not os.path.exists(api.LIB_PATH):
        raise SystemExit("libccnn_phases.so is not built: python -m compactcnncascade_amd.build --phase-stamps")
    cfg = synth.WORKLOADS[workload]
    det = detector(api, cfg["w"], cfg["h"], cfg["min_face"], cfg["max_face"], cfg["scale_factor"])
    us = lambda a: np.round(a / 100.0, 3)              # 100 MHz ticks -> microseconds

    def stats(v):
        v = np.asarray(v, np.float64)
        return {"mean": round(float(v.mean()), 2), "median": round(float(np.median(v)), 2),
                "max": round(float(v.max()), 2)} if len(v) else None

    if kind == "check":
        frames = make_frames(workload, 24)
        dev, df = to_device(torch, frames)
        for (p, w, h, c, s) in df[:16]:
            det.DetectDevice(ctypes.c_void_p(p), w, h, c, s)
        torch.cuda.synchronize()
        names2 = ["crop", "stage-2 input", "stage-2 L1", "L2", "sums", "L3", "hidden", "output"]
        names3 = ["stage-3 input + L1", "L2", "sums", "L3", "hidden", "output", "to item end"]
        per2 = {n: [] for n in names2}
        per3 = {n: [] for n in names3}
        tot2, tot3, spans, items, s3n = [], [], [], [], []
        api.phase_stamps(0, 0)
        for (p, w, h, c, s) in df[16:24]:
            det.DetectDevice(ctypes.c_void_p(p), w, h, c, s)
            torch.cuda.synchronize()
            a = api.phase_stamps(0, 1024).astype(np.int64)
            a = a[(a[:, 0] > 0) & (a[:, 15] > 0)]
            t = a[:, :17] - a[:, 0].min()
            s3 = a[:, 22] == 1
            for k, n in enumerate(names2):
                per2[n] += list(us(t[:, k + 1] - t[:, k]))
            seq3 = [8, 9, 10, 11, 12, 13, 14, 15]
            for k, n in enumerate(names3):
                per3[n] += list(us(t[s3, seq3[k + 1]] - t[s3, seq3[k]]))
            tot2 += list(us(t[~s3, 15] - t[~s3, 0]))
            tot3 += list(us(t[s3, 15] - t[s3, 0]))
            spans.append(float(us(t[:, 15].max())))
            items.append(len(a))
            s3n.append(int(s3.sum()))
        out = {"phase_pass": "check", "workload": workload, "frames": 8, "items_per_frame": items,
               "stage3_items_per_frame": s3n, "stage2_phases_us": {n: stats(v) for n, v in per2.items()},
               "stage3_phases_us": {n: stats(v) for n, v in per3.items()},
               "item_us": {"stage2_only": stats(tot2), "with_stage3": stats(tot3)},
               "span_us": stats(spans)}
    else:
        frames = make_frames(workload, 64)
        dev, df = to_device(torch, frames)
        for _ in range(3):
            det.DetectBatch(device_frames=df)
        torch.cuda.synchronize()
        api.phase_stamps(1, 0)
        det.DetectBatch(device_frames=df)
        torch.cuda.synchronize()
        a = api.phase_stamps(1, 65536).astype(np.int64)
        a = a[(a[:, 0] > 0) & (a[:, 4] > 0)]
        t = a[:, :5] - a[:, 0].min()
        names = ["tile lookup + L1-sum window to LDS", "L2", "L2 sums + tail weights to LDS",
                 "L3 + hidden + output (thread 0)"]
        cus = len(np.unique(a[:, 20]))
        life = us(t[:, 4] - t[:, 0])
        out = {"phase_pass": "stage1", "workload": workload, "frames": 64, "tiles": int(len(a)),
               "phases_us": {n: stats(us(t[:, k + 1] - t[:, k])) for k, n in enumerate(names)},
               "tile_us": stats(life), "span_us": round(float(us(t[:, 4].max())), 1),
               "mean_tiles_in_flight_per_cu": round(float(life.sum() / (float(us(t[:, 4].max())) * cus)), 2)}
    print(json.dumps(out), flush=True)


def autotune_report(steps):
    """The batch path's autotuner (ccnn_autotune: every stage-1 tile shape, then 2/3/4
    persistent check workgroups per CU, each timed as whole DetectBatch passes) on the
    mixed-size batch (8 each of 640x480, 1280x720, 1920x1080, 3840x2160, faces >= 20),
    the 1080p workload (64 frames, faces 40..240) and 8 frames of 3840x2160 (faces >= 20):
    one JSON line per workload."""
    import torch
    torch.zeros(1, device="cuda:0")
    from compactcnncascade_amd import api
    from compactcnncascade_amd.api import TILE_SHAPES
    for name, names, per in (("mixed_batch", ["mixed_vga", "mixed_hd", "mixed_fhd", "mixed_uhd"], 8),
                             ("fhd_min40_max240", ["fhd_min40_max240"], 64), ("uhd_min20", ["mixed_uhd"], 8)):
        frames = []
        for n in names:
            frames.extend(make_frames(n, per))
        dev, df = to_device(torch, frames)
        cfg = synth.WORKLOADS[names[-1]]
        W = max(f.shape[1] for f in frames)
        H = max(f.shape[0] for f in frames)
        det = detector(api, W, H, cfg["min_face"], cfg["max_face"], cfg["scale_factor"])
        best, sms, gms = det.autotune(device_frames=df, reps=steps)
        print(json.dumps({"autotune": name, "frames": len(frames), "best_shape": int(best),
                          "best_shape_tiles": TILE_SHAPES[int(best)],
                          "shape_ms": {"%d %s" % (i, TILE_SHAPES[i]): round(float(v), 4) for i, v in enumerate(sms)},
                          "check_workgroups_per_cu_ms": {str(g): round(float(v), 4) for g, v in zip((2, 3, 4), gms)}}),
              flush=True)
        del dev


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # 100 timed steps of 64 frames (~0.35 s per batch mode), after 10 untimed ones that
    # also let the GPU clocks settle
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--frames", type=int, default=64)
    ap.add_argument("--workload", default="fhd_min40_max240")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-secondary", action="store_true",
                    help="skip the measurements reported beside `value`")
    ap.add_argument("--facial", action="store_true",
                    help="run the workload with facial analysis (the line's config says so)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a device copy of MIB MiB (counter calibration) and exit")
    ap.add_argument("--profile-pass", type=int, default=0, metavar="N",
                    help="only detect the first N workload frames (for rocprofv3 counter runs) and exit")
    ap.add_argument("--profile-mode", default="single", choices=["single", "batch"])
    ap.add_argument("--autotune", action="store_true",
                    help="only run the batch path's autotuner on the benchmark workloads and exit")
    ap.add_argument("--profile-precision", type=int, default=0, choices=[0, 1, 2, 3, 4])
    ap.add_argument("--phase-pass", choices=["check", "stage1"], default=None,
                    help="per-phase latency of k_check items (8 single frames) or k_stage1 tiles (one "
                         "64-frame batch) from the phase-stamp build (build.py --phase-stamps)")
    ap.add_argument("--pmc-summary", nargs=4, metavar=("DB", "CALIB_DB", "COUNTER", "FRAMES"),
                    help="summarize a rocprofv3 counter database per kernel and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: there is no multi-GPU path here (frames would shard "
              "across GPUs as independent detectors, one process per GPU). Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.pmc_summary:
        db, cal, counter, nfr = args.pmc_summary
        pmc_summary(db, cal, counter, int(nfr))
        return
    if args.autotune:
        autotune_report(max(1, args.steps))
        return
    if args.profile_pass:
        profile_pass(args.profile_pass, args.profile_mode, args.workload, args.profile_precision)
        return
    if args.phase_pass:
        phase_pass(args.phase_pass, args.workload)
        return

    import torch
    torch.zeros(1, device="cuda:0")
    from compactcnncascade_amd import api

    cfg = synth.WORKLOADS[args.workload]
    W, H = cfg["w"], cfg["h"]
    frames = make_frames(args.workload, args.frames)
    d_frames = torch.from_numpy(frames).to("cuda:0")
    torch.cuda.synchronize()
    base = d_frames.data_ptr()
    fbytes = W * H * 3
    dframes = [(base + i * fbytes, W, H, 3, W * 3) for i in range(args.frames)]
    param = api.Param(max_image_size=(W, H), min_face_height=cfg["min_face"], max_face_height=cfg["max_face"],
                      scale_factor=cfg["scale_factor"], facial_analysis=args.facial)

    det = api.FaceDetector()
    if det.Init(param) != 0:
        raise SystemExit("FaceDetector.Init failed")
    # record of the work per frame (untimed): candidates, replay status, stage-2/3 evaluations
    records = []
    for i in range(args.frames):
        k, _ = det.DetectDevice(ctypes.c_void_p(base + i * fbytes), W, H, 3, W * 3)
        if k < 0:
            raise SystemExit("Detect failed")
        tr = det.last_trace()
        tr["status"] = det.last_status()
        records.append(tr)
    progress("recorded %d frames" % args.frames)
    # single-frame pass: one Detect call at a time, so each frame's kernels have the GPU
    # to themselves; their durations (HIP events stamped by the launches) give the
    # per-kernel rooflines
    det.set_profiling(True)
    for i in range(args.frames):
        det.DetectDevice(ctypes.c_void_p(base + i * fbytes), W, H, 3, W * 3)
    tm = det.last_timing()
    det.set_profiling(False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.frames):
        det.DetectDevice(ctypes.c_void_p(base + i * fbytes), W, H, 3, W * 3)
    sync_fps = args.frames / (time.perf_counter() - t0)

    # the timed region: K DetectBatch passes over the frames, for each batch mode
    modes = {}
    step_ms = {}
    captures = {}
    for mode in (1, 0):
        det.set_batch_mode(mode)
        for _ in range(args.warmup):
            det.DetectBatch(device_frames=dframes)
        torch.cuda.synchronize()
        c0 = det.batch_stats()[1]
        t0 = time.perf_counter()
        faces = 0
        marks = [t0]
        for _ in range(args.steps):
            out = det.DetectBatch(device_frames=dframes)
            faces += sum(n for n, _ in out)
            marks.append(time.perf_counter())
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        modes[mode] = (dt, faces)
        # DetectBatch returns after the host replay, so each step's span is its own
        step_ms[mode] = [round((b - a) * 1e3, 3) for a, b in zip(marks, marks[1:])]
        captures[mode] = det.batch_stats()[1] - c0
        sm = sorted(step_ms[mode])
        progress("batch mode %d: %.1f frames/s, step ms min %.3f median %.3f max %.3f, graph captures %d"
                 % (mode, args.steps * args.frames / dt, sm[0], sm[len(sm) // 2], sm[-1], captures[mode]))
    best_mode = min(modes, key=lambda m: modes[m][0])
    dt, faces = modes[best_mode]
    nframes = args.steps * args.frames
    fps = nframes / dt

    # algorithmic work of the recorded pass
    s1 = np.array([stage1_macs(tr["levels"]) for tr in records], np.float64).mean(0)
    s1_mac = s1.sum()
    ref_e = np.zeros(2)
    all_e = np.zeros(2)
    cands = 0
    for tr in records:
        r, a = check_evals(tr["res"], tr["status"])
        ref_e += r
        all_e += a
        cands += len(tr["cand"])
    ref_e /= len(records)
    all_e /= len(records)
    per_frame_ms = {k: tm["totals"][k] / max(1, tm["frames"])
                    for k in ("gray", "stage1_l1", "stage1", "gap", "check", "after_check")}
    # algorithmic FLOP (2 per MAC) per launch of the three VALU kernels: stage 1 as the
    # reference computes it, stages 2/3 for the evaluations the reference performs
    # (the kernel also evaluates the candidates the replay drops, counted as executed)
    kflops = {"k_stage1_l1": 2.0 * s1[0], "k_stage1": 2.0 * s1[1],
              "k_check": 2.0 * (ref_e[0] * STAGE2_MAC + ref_e[1] * STAGE3_MAC)}
    executed = {"k_check": 2.0 * (all_e[0] * STAGE2_MAC + all_e[1] * STAGE3_MAC)}
    kms = {"k_stage1_l1": per_frame_ms["stage1_l1"], "k_stage1": per_frame_ms["stage1"],
           "k_check": per_frame_ms["check"]}
    kernels = {}
    for k in kflops:
        tf = kflops[k] / (kms[k] * 1e-3) / 1e12
        kernels[k] = {"ms": round(kms[k], 4), "gflop": round(kflops[k] / 1e9, 4), "tflops": round(tf, 3),
                      "frac_of_fp32_peak": round(tf / FP32_PEAK_TFLOPS, 4),
                      "frac_of_exact_valu": round(tf / VALU_EXACT_PEAK_TOPS, 3)}
        if k in executed:
            kernels[k]["gflop_executed"] = round(executed[k] / 1e9, 4)
    # batch pass: the same 64 frames as one batch whose kernels run without the graph, with
    # HIP events between the kernel groups on the batch stream (each launch covers every
    # frame: the GPU is full, as in the throughput runs)
    det.set_batch_mode(1)
    det.set_profiling(True)
    for _ in range(3):
        det.DetectBatch(device_frames=dframes)
    bt = det.last_batch_timing()
    det.set_profiling(False)
    nb_frames = bt["frames"]
    bms = {"k_stage1_l1": bt["stage1_l1"], "k_stage1": bt["stage1"], "k_check": bt["check"]}
    batch_kernels = {}
    for k in kflops:
        tf = kflops[k] * nb_frames / (bms[k] * 1e-3) / 1e12
        batch_kernels[k] = {"ms_per_launch": round(bms[k], 4), "us_per_frame": round(bms[k] * 1e3 / nb_frames, 2),
                            "tflops": round(tf, 3), "frac_of_exact_valu": round(tf / VALU_EXACT_PEAK_TOPS, 3)}
    batch_kernels["k_gray_batch"] = {"ms_per_launch": round(bt["gray"], 4),
                                     "us_per_frame": round(bt["gray"] * 1e3 / nb_frames, 2)}
    batch_kernels["compaction"] = {"ms_per_launch": round(bt["compact"], 4)}
    dominant = max(bms, key=lambda k: bms[k])
    achieved = kflops[dominant] * nb_frames / (bms[dominant] * 1e-3) / 1e12
    traffic = measured("%s_batch64_%s_hbm_traffic.txt" % (dominant, args.workload)).get("bytes_per_launch")
    frame_bytes = measured("frame_%s_hbm_traffic.txt" % args.workload)
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
                   "launch": ("batch graph: one captured HIP graph launch per 64-frame batch" if best_mode == 1 else
                              "per-frame launches pipelined over three streams") + ", frames resident in HBM"},
        "batch_modes_fps": {"graph": round(nframes / modes[1][0], 2), "pipelined": round(nframes / modes[0][0], 2)},
        "step_ms": {"graph": step_ms[1], "pipelined": step_ms[0]},
        "graph_captures_in_timed_steps": captures[1],
        "roofline": {
            "kernel": dominant,
            "bound": "valu",
            "achieved": round(achieved, 3),
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / FP32_PEAK_TFLOPS, 4),
            "frac_exact_valu": round(achieved / VALU_EXACT_PEAK_TOPS, 4),
            "exact_valu_peak": VALU_EXACT_PEAK_TOPS,
            "traffic": int(traffic) if traffic else None,
            "measured_in": "batch pass: one launch of each kernel over the %d frames (run without the graph, "
                           "HIP events between the kernels on the batch stream), the longest kernel" % nb_frames,
            "note": "FP32 VALU kernel on the exact path (every multiply and add rounded separately: no FMA, no "
                    "packed FP32), so its own ceiling is %.1f T ops/s; achieved counts 2 FLOP per MAC of the "
                    "reference's algorithm (stage 2/3: the evaluations the reference performs)" % VALU_EXACT_PEAK_TOPS,
        },
        "batch_kernels": batch_kernels,
        "kernel_ms_per_frame_single_frame_pass": {k: round(v, 4) for k, v in per_frame_ms.items()},
        "valu_kernels_single_frame_pass": kernels,
        "work_per_frame": {"stage1_gmac": round(s1_mac / 1e9, 4), "candidates": round(cands / len(records), 1),
                           "stage2_evals_reference": round(ref_e[0], 1), "stage3_evals_reference": round(ref_e[1], 1),
                           "stage2_evals_executed": round(all_e[0], 1), "stage3_evals_executed": round(all_e[1], 1)},
        "faces_per_frame": round(faces / nframes, 2),
        "sync_single_frame_fps": round(sync_fps, 1),
    }
    if frame_bytes.get("bytes_per_frame"):
        line["hbm"] = {"bytes_per_frame": int(frame_bytes["bytes_per_frame"]),
                       "achieved_gbs": round(frame_bytes["bytes_per_frame"] * fps / 1e9, 1),
                       "frac_of_peak": round(frame_bytes["bytes_per_frame"] * fps / 1e9 / HBM_PEAK_GBS, 4),
                       "source": "profiles/frame_%s_hbm_traffic.txt (rocprofv3 FETCH_SIZE + WRITE_SIZE of every "
                                 "kernel of the single-frame pass, copy-calibrated)" % args.workload}
    if not args.no_secondary:
        # the metric's 4K leg, the mixed-size batch and the deep pyramid, each in FP32 and
        # the reduced precisions (recall against FP32); the reference's CPU-runnable case;
        # learned tile skipping; facial analysis; host frames over PCIe
        legs = [
            ("uhd_3840x2160", lambda: secondary_fps(api, torch, ["mixed_uhd"], 8, 3, precisions=(0, 1, 2, 3, 4))),
            ("mixed_batch", lambda: secondary_fps(api, torch, ["mixed_vga", "mixed_hd", "mixed_fhd", "mixed_uhd"],
                                                  8, 3, precisions=(0, 1, 2, 3, 4))),
            ("deep_4800x2400", lambda: secondary_fps(api, torch, ["deep_4800x2400_sf105_min15"], 2, 2,
                                                     precisions=(0, 1, 2, 3, 4))),
            ("vga_640x480", lambda: secondary_fps(api, torch, ["vga_min40"], 64, 3)),
            ("skip_fhd_min20", lambda: skip_leg(api, torch, 1920, 1080, 32, 3)),
            ("skip_uhd_min20", lambda: skip_leg(api, torch, 3840, 2160, 8, 3)),
            ("fhd_facial", lambda: secondary_fps(api, torch, ["fhd_min40_max240"], 64, 3, facial=True)),
            ("fhd_host_frames", lambda: host_frame_fps(api, torch, frames, param)),
        ]
        for name, leg in legs:
            line[name] = leg()
            progress("leg %s done" % name)
    if not args.no_cpu_baseline:
        cb = cpu_baseline(frames, dict(min_face=cfg["min_face"], max_face=cfg["max_face"],
                                       scale_factor=cfg["scale_factor"]), args.cpu_seconds)
        line["cpu_baseline"] = cb
        progress("cpu baseline done")
        if cb["value"]:
            line["gpu_over_cpu_all_cores"] = round(fps / cb["value"], 1)
            line["gpu_over_cpu_one_core"] = round(fps / cb["one_core_fps"], 1)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
