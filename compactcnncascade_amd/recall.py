"""Recall judge: how many of the plain (exact FP32) cascade's detections a reduced form
(FP16 / BF16 / FP8 precision, learned tile skipping) still finds.

A detection of the reference set counts as found when a not-yet-paired detection of the
tested set overlaps it by at least `iou` (0.99 by default), with the overlap measure of
the reference's Rect::overlap (type.h:154-173): the intersection area over the area of
the bounding box of both rectangles (0 when they do not touch). Pairs are taken greedily
in order of decreasing overlap.
"""
import numpy as np


def rect_overlap(a, b):
    """Rect::overlap of two (x, y, w, h) boxes."""
    ax, ay, aw, ah = (int(v) for v in a[:4])
    bx, by, bw, bh = (int(v) for v in b[:4])
    ax2, ay2, bx2, by2 = ax + aw, ay + ah, bx + bw, by + bh
    if ax > bx2 or bx > ax2 or ay > by2 or by > ay2:
        return 0.0
    xs = sorted((ax, ax2, bx, bx2))
    ys = sorted((ay, ay2, by, by2))
    su = float((xs[3] - xs[0]) * (ys[3] - ys[0]))
    si = float((xs[2] - xs[1]) * (ys[2] - ys[1]))
    return si / su if su > 0 else 0.0


def boxes(faces):
    """(n, 4) int array of x, y, w, h from a FACE_DTYPE array or an (n, >=4) array."""
    f = np.asarray(faces)
    if f.dtype.names:
        return np.stack([f["x"], f["y"], f["width"], f["height"]], 1).astype(np.int64) if len(f) else \
            np.zeros((0, 4), np.int64)
    return f[:, :4].astype(np.int64) if len(f) else np.zeros((0, 4), np.int64)


def match(ref, test, iou=0.99):
    """Greedy one-to-one pairing of ref and test boxes with overlap >= iou: the number of
    ref boxes paired."""
    r, t = boxes(ref), boxes(test)
    if len(r) == 0 or len(t) == 0:
        return 0
    cand = []
    for i in range(len(r)):
        for j in range(len(t)):
            o = rect_overlap(r[i], t[j])
            if o >= iou:
                cand.append((o, i, j))
    cand.sort(key=lambda c: -c[0])
    ur, ut = set(), set()
    for o, i, j in cand:
        if i in ur or j in ut:
            continue
        ur.add(i)
        ut.add(j)
    return len(ur)


def recall(ref_frames, test_frames, iou=0.99):
    """Over a list of frames: (recall, paired, reference count, tested count). Recall is 1
    when the reference found nothing."""
    paired = nref = ntest = 0
    for r, t in zip(ref_frames, test_frames):
        paired += match(r, t, iou)
        nref += len(boxes(r))
        ntest += len(boxes(t))
    return (paired / nref if nref else 1.0), paired, nref, ntest
