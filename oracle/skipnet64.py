"""FP64 oracle of the tile-skipping proposal network (SkipNet; the network of
compactcnncascade_amd/csrc/ccnn_skipnet.hip): forward pass, weighted binary
cross-entropy loss and its analytic gradient in float64, written from the definition:

  a1 = relu(conv3x3(x0; w1, b1))      1 -> 8 channels, zero padding
  a2 = relu(conv3x3(a1; w2, b2))      8 -> 8
  z  = conv3x3(a2; w3, b3)            8 -> S (one logit per pyramid scale slot)
  L  = mean over all outputs of  pos_weight * softplus(-z)  (label 1)
                                 softplus(z)                (label 0)

TEST INFRASTRUCTURE: the GPU training kernels' gradients are checked against grad64, and
grad64 itself against central finite differences of loss64 (tests/test_oracle.py pins it
that way; tests/test_gpu_skipnet.py compares the GPU).
"""
import numpy as np

S = 24
SHAPES = [("w1", (8, 1, 3, 3)), ("b1", (8,)), ("w2", (8, 8, 3, 3)), ("b2", (8,)), ("w3", (S, 8, 3, 3)), ("b3", (S,))]
NPAR = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(p):
    p = np.asarray(p, np.float64)
    out, o = {}, 0
    for name, shp in SHAPES:
        n = int(np.prod(shp))
        out[name] = p[o:o + n].reshape(shp)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], np.float64).reshape(-1) for name, _ in SHAPES])


def _conv(x, w, b):
    """x [N, Ci, H, W], w [Co, Ci, 3, 3] -> [N, Co, H, W], zero padding."""
    n, ci, h, wd = x.shape
    xp = np.zeros((n, ci, h + 2, wd + 2))
    xp[:, :, 1:-1, 1:-1] = x
    out = np.zeros((n, w.shape[0], h, wd)) + b[None, :, None, None]
    for ky in range(3):
        for kx in range(3):
            out += np.einsum("nchw,oc->nohw", xp[:, :, ky:ky + h, kx:kx + wd], w[:, :, ky, kx])
    return out


def _conv_bwd(x, w, g):
    """Gradients of sum(g * conv(x; w, b)) w.r.t. x, w and b."""
    n, ci, h, wd = x.shape
    xp = np.zeros((n, ci, h + 2, wd + 2))
    xp[:, :, 1:-1, 1:-1] = x
    gx = np.zeros_like(xp)
    gw = np.zeros_like(w)
    for ky in range(3):
        for kx in range(3):
            gw[:, :, ky, kx] = np.einsum("nohw,nchw->oc", g, xp[:, :, ky:ky + h, kx:kx + wd])
            gx[:, :, ky:ky + h, kx:kx + wd] += np.einsum("nohw,oc->nchw", g, w[:, :, ky, kx])
    return gx[:, :, 1:-1, 1:-1], gw, g.sum(axis=(0, 2, 3))


def forward64(p, x0):
    """Logits [N, S, H8, W8] of inputs x0 [N, H8, W8]."""
    d = unpack(p)
    a1 = np.maximum(_conv(x0[:, None], d["w1"], d["b1"]), 0)
    a2 = np.maximum(_conv(a1, d["w2"], d["b2"]), 0)
    return _conv(a2, d["w3"], d["b3"])


def loss64(p, x0, y, pos_weight):
    z = forward64(p, x0)
    y = np.asarray(y, bool)
    sp = np.maximum(z, 0) + np.log1p(np.exp(-np.abs(z)))
    return float(np.mean(np.where(y, pos_weight * (sp - z), sp)))


def grad64(p, x0, y, pos_weight):
    """(loss, gradient vector in the parameter layout)."""
    d = unpack(p)
    x = np.asarray(x0, np.float64)[:, None]
    h1 = _conv(x, d["w1"], d["b1"])
    a1 = np.maximum(h1, 0)
    h2 = _conv(a1, d["w2"], d["b2"])
    a2 = np.maximum(h2, 0)
    z = _conv(a2, d["w3"], d["b3"])
    y = np.asarray(y, bool)
    wts = np.where(y, pos_weight, 1.0)
    sp = np.maximum(z, 0) + np.log1p(np.exp(-np.abs(z)))
    loss = float(np.mean(np.where(y, pos_weight * (sp - z), sp)))
    dz = wts * (1 / (1 + np.exp(-z)) - y) / z.size
    g2, gw3, gb3 = _conv_bwd(a2, d["w3"], dz)
    g2 = g2 * (h2 > 0)
    g1, gw2, gb2 = _conv_bwd(a1, d["w2"], g2)
    g1 = g1 * (h1 > 0)
    _, gw1, gb1 = _conv_bwd(x, d["w1"], g1)
    return loss, pack(dict(w1=gw1, b1=gb1, w2=gw2, b2=gb2, w3=gw3, b3=gb3))


def init_params(seed):
    """He-style initial weights and zero biases (the training run's starting point)."""
    rng = np.random.default_rng(seed)
    d = {}
    for name, shp in SHAPES:
        if name.startswith("w"):
            fan_in = shp[1] * 9
            d[name] = rng.normal(0, np.sqrt(2.0 / fan_in), shp)
        else:
            d[name] = np.zeros(shp)
    d["b3"] = np.full(S, -2.0)
    return pack(d)
