"""Synthetic CNTK node dumps of cascade networks, for testing the dump -> binary converter.

A dump is the text that CNTK's "dumpnode" action writes for a trained network
(cntk/stage_1.cntk:299-304 configures it): one node per header line
"<name>=<operation> ...", a parameter node's values on the lines after its header, and a
line of '#' closing each node. The converter (cnn_models_converter.cpp:79-268) reads
only what this module writes: the "features" input shape "[ w x h x 3 ]", the line after
the first "Pooling" header (its second field: 1 for max pooling), the "Kernel:" fields of
the z.c<l>1.c and z.h1.c convolution headers, and the values of the parameter nodes of
the z.c<l><m> / z.bn<l><m> maps, the z.h<g> hidden-layer groups, z.bn_h and the output
layer z.z. The network shapes follow cntk/stage_1.cntk-stage_3.cntk; the values are
seeded random numbers.
"""
import numpy as np

# (maps per conv layer, (kw, kh) per conv layer, hidden groups, hidden units per group,
#  inputs per hidden unit, outputs, input w x h, full_connect)
SHAPES = {
    # cntk/stage_1.cntk: 4/8/16 maps, 4x4, 3x3, 4x5 kernels, 13 groups of 4 units over 4 maps
    "stage1": ((4, 8, 16), ((4, 4), (3, 3), (4, 5)), 13, 4, 4, 3, (23, 27), False),
    # cntk/stage_2.cntk: 6/12/24 maps, 7x8 third-layer kernels, 21 groups of 4 units over 4 maps
    "stage2": ((6, 12, 24), ((4, 4), (3, 3), (7, 8)), 21, 4, 4, 3, (47, 51), False),
    # a fully connected hidden layer with 13 outputs, as the landmarks network has
    # (one group of units over all third-layer maps; more than three outputs)
    "fc13": ((4, 8, 12), ((4, 4), (3, 3), (5, 6)), 1, 20, 12, 13, (40, 44), True),
}


def _fmt(rng, v):
    # CNTK prints plain decimals; some values in exponent form exercise the float parser
    return ("%.9g" % v) if rng.random() < 0.8 else ("%.7e" % v)


class _Writer:
    def __init__(self, rng, sep, dot_underscore):
        self.rng = rng
        self.sep = sep
        self.du = dot_underscore
        self.lines = []

    def name(self, node, param):
        # CNTK's BrainScript names carry "._" pieces, which the converter drops
        if self.du and self.rng.random() < 0.5:
            return "z.%s._.%s" % (node, param)
        return "z.%s.%s" % (node, param)

    def header(self, text):
        self.lines.append(text)

    def param(self, node, param, values, per_line):
        v = np.asarray(values, dtype=np.float64).ravel()
        self.lines.append("%s=LearnableParameter [%d x 1]   learningRateMultiplier=1.000000  NeedsGradient=true"
                          % (self.name(node, param), v.size))
        for i in range(0, v.size, per_line):
            self.lines.append(" " + " ".join(_fmt(self.rng, x) for x in v[i:i + per_line]) + " ")
        self.lines.append(" " + "#" * 20)

    def text(self):
        return self.sep.join(self.lines) + self.sep


def make_dump(shape, seed, max_pool=True, sep="\n", dot_underscore=True, drop=None, n_out=None):
    """Text of a dump of network `shape` (a SHAPES key) with seeded values.

    max_pool False writes average pooling (kind 0); sep "\\r\\n" writes CRLF line ends;
    drop names a parameter node ("<node>.<param>") to leave out; n_out overrides the
    output count."""
    maps, ks, groups, hl_scale, connect, outs, (iw, ih), full = SHAPES[shape]
    if n_out is not None:
        outs = n_out
    rng = np.random.default_rng(seed)
    w = _Writer(rng, sep, dot_underscore)
    w.header("features=InputValue [ %d x %d x 3 ]" % (iw, ih))
    w.header("labels=InputValue [ %d ]" % outs)
    w.header("z.p11.p=Pooling (z.bn11) : [%d x %d x 1] -> [%d x %d x 1]"
             % (iw - ks[0][0] + 1, ih - ks[0][1] + 1, (iw - ks[0][0] + 1) // 2, (ih - ks[0][1] + 1) // 2))
    w.header(" poolKind: %d  window: 2 x 2  stride: 2 x 2" % (1 if max_pool else 0))

    def emit(node, param, values, per_line):
        if drop != "%s.%s" % (node, param):
            w.param(node, param, values, per_line)

    for l in range(3):
        kw, kh = ks[l]
        ch = 3 if l == 0 else 1
        for m in range(maps[l]):
            node = "c%d%d" % (l + 1, m + 1)
            w.header("%s=Convolution (%s, %s) : Kernel: %d x %d x %d, Map: 1, Stride: 1 x 1 x %d, Sharing: (1), "
                     "AutoPadding: (0), LowerPad: 0, UpperPad: 0"
                     % (w.name(node, "c"), w.name(node, "W"), "features" if l == 0 else "z.s%d" % l, kw, kh, ch, ch))
            emit(node, "W", rng.normal(0.0, 0.3, size=kw * kh * ch), kw * ch)
            emit(node, "b", rng.normal(0.0, 0.2, size=1), 1)
            emit(node, "res.PlusArgs[0].ElementTimesArgs[0]", rng.uniform(0.05, 0.3, size=1), 1)
            emit(node, "res.PlusArgs[1].ElementTimesArgs[0]", rng.uniform(0.6, 1.2, size=1), 1)
        for m in range(maps[l]):
            node = "bn%d%d" % (l + 1, m + 1)
            emit(node, "scale", rng.uniform(0.5, 1.5, size=1), 1)
            emit(node, "bias", rng.normal(0.0, 0.3, size=1), 1)
            emit(node, "runMean", rng.normal(0.0, 0.5, size=1), 1)
            emit(node, "runVariance", rng.uniform(0.05, 2.0, size=1), 1)
    for g in range(groups):
        node = "h%d" % (g + 1)
        w.header("%s=Convolution (%s, z.concat_%d) : Kernel: %d x 1 x 1 x %d, Map: %d, Stride: 1 x 1 x 1, "
                 "Sharing: (1), AutoPadding: (0)" % (w.name(node, "c"), w.name(node, "W"), g + 1, connect,
                                                    hl_scale, hl_scale))
        emit(node, "W", rng.normal(0.0, 0.5, size=connect * hl_scale), hl_scale)
        emit(node, "b", rng.normal(0.0, 0.2, size=hl_scale), 1)
        emit(node, "res.PlusArgs[0].ElementTimesArgs[1].x.ElementTimesArgs[0]", rng.uniform(0.3, 1.0, size=1), 1)
    # batch normalisation of the hidden layer: one row per unit of a group
    emit("bn_h", "scale", rng.uniform(0.5, 1.5, size=hl_scale), 1)
    emit("bn_h", "bias", rng.normal(0.0, 0.3, size=hl_scale), 1)
    emit("bn_h", "runMean", rng.normal(0.0, 0.5, size=hl_scale), 1)
    emit("bn_h", "runVariance", rng.uniform(0.05, 2.0, size=hl_scale), 1)
    hl = groups * hl_scale
    emit("z", "arrayOfFunctions[0].W", rng.normal(0.0, 0.4, size=outs * hl), hl)
    emit("z", "arrayOfFunctions[0].b", rng.normal(0.0, 0.2, size=outs), 1)
    emit("z", "PlusArgs[0].ElementTimesArgs[1].x.ElementTimesArgs[0]", rng.uniform(0.3, 1.0, size=1), 1)
    if full:
        assert connect == maps[2]
    return w.text()
