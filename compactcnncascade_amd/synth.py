"""Seeded synthetic BGR frames for tests and benchmarks.

Deterministic numpy (PCG64) generator: a background of uniform u8 noise smoothed
at three Gaussian octaves and summed, plus Poisson(lam) face-like patches (a bright
ellipse with two dark eye blobs and a mouth bar), log-uniform size, rotation
U(-15, 15) degrees, contrast U(0.6, 1.4), written to B, G and R with a small
per-channel jitter so that the colour conversion is exercised.
"""
import numpy as np


def _gauss_kernel(sigma):
    r = max(1, int(3 * sigma))
    x = np.arange(-r, r + 1, dtype=np.float64)
    k = np.exp(-0.5 * (x / sigma) ** 2)
    return k / k.sum()


def _blur(img, sigma):
    k = _gauss_kernel(sigma)
    r = len(k) // 2
    p = np.pad(img, ((0, 0), (r, r)), mode="reflect")
    out = np.zeros_like(img)
    for i, kv in enumerate(k):
        out += kv * p[:, i:i + img.shape[1]]
    p = np.pad(out, ((r, r), (0, 0)), mode="reflect")
    out2 = np.zeros_like(img)
    for i, kv in enumerate(k):
        out2 += kv * p[i:i + img.shape[0], :]
    return out2


def _upsample(small, h, w):
    sh, sw = small.shape
    ys = np.linspace(0, sh - 1, h)
    xs = np.linspace(0, sw - 1, w)
    y0 = np.floor(ys).astype(np.int64)
    x0 = np.floor(xs).astype(np.int64)
    y1 = np.minimum(y0 + 1, sh - 1)
    x1 = np.minimum(x0 + 1, sw - 1)
    fy = (ys - y0)[:, None]
    fx = (xs - x0)[None, :]
    a = small[y0][:, x0] * (1 - fx) + small[y0][:, x1] * fx
    b = small[y1][:, x0] * (1 - fx) + small[y1][:, x1] * fx
    return a * (1 - fy) + b * fy


def background(rng, h, w):
    acc = np.zeros((h, w), np.float64)
    for k, (factor, sigma, weight) in enumerate(((2, 1.0, 0.35), (8, 1.5, 0.4), (32, 1.5, 0.25))):
        sh, sw = max(2, h // factor + 2), max(2, w // factor + 2)
        noise = rng.random((sh, sw)) * 255.0
        acc += weight * _upsample(_blur(noise, sigma), h, w)
    acc -= acc.mean()
    std = acc.std() + 1e-9
    return np.clip(118.0 + acc * (24.0 / std), 0, 255)


def _face_patch(rng, size):
    """Gray patch (float, 0..255 offsets) and alpha mask for one face-like object."""
    s = int(size)
    half = s // 2 + 2
    yy, xx = np.mgrid[-half:half + 1, -half:half + 1].astype(np.float64)
    ang = np.deg2rad(rng.uniform(-15.0, 15.0))
    ca, sa = np.cos(ang), np.sin(ang)
    u = ca * xx + sa * yy
    v = -sa * xx + ca * yy
    a, b = 0.40 * s, 0.50 * s
    ell = (u / a) ** 2 + (v / b) ** 2
    alpha = np.clip((1.0 - ell) * 6.0, 0.0, 1.0)
    val = np.full_like(u, 200.0)
    eye_r = 0.085 * s
    for ex in (-0.17 * s, 0.17 * s):
        d = ((u - ex) ** 2 + (v + 0.12 * s) ** 2) / (eye_r ** 2)
        val -= 140.0 * np.clip(1.5 - d, 0.0, 1.0)
    mouth = np.clip(1.0 - np.abs(u) / (0.17 * s), 0.0, 1.0) * np.clip(1.0 - np.abs(v - 0.24 * s) / (0.04 * s + 0.5), 0.0, 1.0)
    val -= 110.0 * np.clip(mouth * 2.0, 0.0, 1.0)
    nose = np.clip(1.0 - np.abs(u) / (0.05 * s + 0.5), 0.0, 1.0) * np.clip(1.0 - np.abs(v - 0.05 * s) / (0.1 * s), 0.0, 1.0)
    val -= 40.0 * nose
    return val, alpha


def synth_frame(seed, w, h, lam=6.0, min_size=40, max_size=200):
    """Returns (frame uint8 [h, w, 3] BGR, boxes int32 [n, 4] x, y, w, h)."""
    rng = np.random.default_rng(seed)
    gray = background(rng, h, w)
    n = int(rng.poisson(lam))
    boxes = []
    for _ in range(n):
        size = float(np.exp(rng.uniform(np.log(min_size), np.log(max_size))))
        size = min(size, 0.9 * min(w, h))
        val, alpha = _face_patch(rng, size)
        ph, pw = val.shape
        cx = int(rng.uniform(0, max(1, w - pw))) if w > pw else 0
        cy = int(rng.uniform(0, max(1, h - ph))) if h > ph else 0
        contrast = rng.uniform(0.6, 1.4)
        region = gray[cy:cy + ph, cx:cx + pw]
        vv = val[:region.shape[0], :region.shape[1]]
        aa = alpha[:region.shape[0], :region.shape[1]]
        face = np.clip(185.0 + (vv - 200.0) * contrast, 0, 255)
        gray[cy:cy + ph, cx:cx + pw] = region * (1 - aa) + face * aa
        boxes.append((cx + int(pw * 0.1), cy + int(ph * 0.05), int(pw * 0.8), int(ph * 0.9)))
    jitter = rng.normal(0.0, 4.0, size=(3, h, w))
    frame = np.empty((h, w, 3), np.uint8)
    for c in range(3):
        frame[:, :, c] = np.clip(np.rint(gray + jitter[c] + (c - 1) * 3.0), 0, 255).astype(np.uint8)
    return frame, np.array(boxes, np.int32).reshape(-1, 4)


# The benchmark workloads (frame size, pyramid and generator settings).
WORKLOADS = {
    "vga_min40": dict(w=640, h=480, lam=6.0, min_size=40, max_size=200, scale_factor=1.15, min_face=40, max_face=0),
    "fhd_min40_max240": dict(w=1920, h=1080, lam=10.0, min_size=20, max_size=300, scale_factor=1.15, min_face=40,
                             max_face=240),
    "deep_4800x2400_sf105_min15": dict(w=4800, h=2400, lam=40.0, min_size=15, max_size=400, scale_factor=1.05,
                                       min_face=15, max_face=0),
    # the mixed-size batch: 8 frames each of 640x480, 1280x720, 1920x1080 and 3840x2160
    "mixed_vga": dict(w=640, h=480, lam=10.0, min_size=20, max_size=300, scale_factor=1.15, min_face=20, max_face=0),
    "mixed_hd": dict(w=1280, h=720, lam=10.0, min_size=20, max_size=300, scale_factor=1.15, min_face=20, max_face=0),
    "mixed_fhd": dict(w=1920, h=1080, lam=10.0, min_size=20, max_size=300, scale_factor=1.15, min_face=20,
                      max_face=0),
    "mixed_uhd": dict(w=3840, h=2160, lam=10.0, min_size=20, max_size=300, scale_factor=1.15, min_face=20,
                      max_face=0),
}


def workload_frame(name, idx, base_seed=1234):
    cfg = WORKLOADS[name]
    seed = (base_seed * 1000003 + hash_name(name) * 7919 + idx) & 0xFFFFFFFF
    return synth_frame(seed, cfg["w"], cfg["h"], cfg["lam"], cfg["min_size"], cfg["max_size"])


def hash_name(name):
    h = 2166136261
    for ch in name.encode():
        h = ((h ^ ch) * 16777619) & 0xFFFFFFFF
    return h
