"""FP64 oracle of the cascade networks: the textbook restatement of one network's
forward pass in float64, from the model's weights as the binary format lays them out
(cnn_simd_cntk.cpp:42-404) and the reference's layer definitions:

  L1, L2, L3   valid cross-correlation (cnnpp_cplusplus.cpp:36-321), + conv_b, then
               w1*c + w2*max(0, c) + bn_b (the batch norm folded into w1 / w2, :593-645),
               2x2 floor max-pool after L1 and L2;
  sums         each map plus its neighbours (cnn_simd_cntk.cpp:705-725, 758-778), sum i
               feeding convolutions 2i and 2i+1 of the next layer;
  hidden       unit u = S*g + t: bn_w[t] * (0.5 * T(tanh_w[g] * (b[u] + sum_j a3[g+j] w[u,j]))
               + 0.5) + bn_b[t] (cnnpp_cplusplus.cpp:646-672);
  output       T(tanh_w * (sum_{g,t} h[S*g+t] * w[t*G+g] + b)), scaled by af_scale
               (cnnpp_cplusplus.cpp:673-690, 952-984);
  T            the reference's rational tanh sgn(c) (1 - 1/(1 + |c| + c^2 + 1.41645 c^4))
               (cnnpp_cplusplus.cpp:662-666), evaluated exactly.

TEST INFRASTRUCTURE: the reduced-precision paths (FP16 stage 1, BF16 / FP8 stages 2/3)
state their per-map error against this function; the exact FP32 path is checked against it
as well. Pinned by tests/test_oracle.py against the C oracle (within FP32 rounding) and,
through it, against the compiled reference.
"""
import ctypes

import numpy as np


def _arr(p, n):
    return np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ctypes.c_float)), shape=(n,)).astype(np.float64)


def _conv(img, k):
    H, W = img.shape
    r, c = k.shape
    out = np.zeros((H - r + 1, W - c + 1))
    for i in range(r):
        for j in range(c):
            out += k[i, j] * img[i:i + H - r + 1, j:j + W - c + 1]
    return out


def _pool(v):
    h, w = v.shape[0] // 2 * 2, v.shape[1] // 2 * 2
    v = v[:h, :w]
    return np.maximum(np.maximum(v[0::2, 0::2], v[0::2, 1::2]), np.maximum(v[1::2, 0::2], v[1::2, 1::2]))


def _sums(maps):
    n = len(maps)
    return [sum(maps[max(0, i - 1):min(n, i + 2)]) for i in range(n)]


def _rtanh(c):
    return np.sign(c) * (1 - 1 / (1 + np.abs(c) + c * c + 1.41645 * c ** 4))


def forward64(m, x):
    """float64 forward of model m (an oracle.bind.OcModel) on image x (H x W): the
    response map (index_output >= 0) or the stacked outputs (index_output < 0)."""
    M = list(m.maps)
    kw, kh = list(m.kw), list(m.kh)
    K = [_arr(m.kern[l], M[l] * kw[l] * kh[l]).reshape(M[l], kh[l], kw[l]) for l in range(3)]
    B = [_arr(m.conv_b[l], M[l]) for l in range(3)]
    W1 = [_arr(m.lrelu_w1[l], M[l]) for l in range(3)]
    W2 = [_arr(m.lrelu_w2[l], M[l]) for l in range(3)]
    BB = [_arr(m.bn_b[l], M[l]) for l in range(3)]

    def act(v, l, i):
        v = v + B[l][i]
        return W1[l][i] * v + W2[l][i] * np.maximum(v, 0) + BB[l][i]

    x = np.asarray(x, np.float64)
    p1 = [_pool(act(_conv(x, K[0][i]), 0, i)) for i in range(M[0])]
    s1 = _sums(p1)
    p2 = [_pool(act(_conv(s1[i // 2], K[1][i]), 1, i)) for i in range(M[1])]
    s2 = _sums(p2)
    a3 = [act(_conv(s2[i // 2], K[2][i]), 2, i) for i in range(M[2])]
    S, G, N = m.hl_scale, m.hl // m.hl_scale, m.connect
    # weights per hidden unit: `connect` of them, or one per L3 map when fully connected
    hw = _arr(m.hl_w, m.hl * m.hl_nin).reshape(m.hl, m.hl_nin)[:, :N]
    hb, ht = _arr(m.hl_b, m.hl), _arr(m.hl_tanh_w, G)
    bw, bb = _arr(m.hl_bn_w, S), _arr(m.hl_bn_b, S)
    outs = range(m.ol) if m.index_output < 0 else [m.index_output]
    h = []
    for g in range(G):
        for t in range(S):
            u = S * g + t
            c = hb[u] + sum(a3[g + j] * hw[u, j] for j in range(N))
            h.append(bw[t] * (0.5 * _rtanh(c * ht[g]) + 0.5) + bb[t])
    res = []
    for k in outs:
        ow = _arr(m.ol_w, m.ol * m.hl).reshape(m.ol, m.hl)[k]
        ob = _arr(m.ol_b, m.ol)[k]
        o = sum(h[S * g + t] * ow[t * G + g] for g in range(G) for t in range(S))
        v = _rtanh((o + ob) * m.ol_tanh_w)
        res.append(0.5 * v + 0.5 if m.af_scale == 0 else m.af_scale * v)
    return res[0] if m.index_output >= 0 else np.stack(res)


def max_abs_error(got, exp64):
    """Per-map error statement: the largest |got - exp| over the map."""
    return float(np.max(np.abs(np.asarray(got, np.float64) - exp64))) if np.size(exp64) else 0.0
