"""CPU reference of the learned pressure warm start's network (TEST
INFRASTRUCTURE ONLY: the gradient and inference checks of the GPU kernels use
it; nothing in the library calls it).

The network has no counterpart in the reference solver: it is the learned
warm start the benchmark configuration adds in front of the plain SOR
solve. Architecture, a 3-level 3-D U-Net on channels-last grids:

    in (5) -> c0 16 -> c1 16 = s1 -> pool -> c2 32 -> c3 32 = s2 -> pool
           -> c4 64 -> c5 64 -> up, concat s2 (96) -> c6 32 -> c7 32
           -> up, concat s1 (48) -> c8 16 -> c9 16 -> head 1x1x1 -> out (1)

3x3x3 convolutions with zero padding, bias and ReLU; average pooling 2x2x2;
nearest-neighbour upsampling x2; a linear 1x1x1 head.

Parameters are one flat FP32 vector, per layer: weights [Cout][tap][Cin]
(tap = (dz*3 + dy)*3 + dx, offsets -1..1), then bias [Cout]. The same order
is what the GPU library stores.

`bf16=True` rounds weights, stored activations and the gradients that feed
products to bfloat16 (round to nearest even) at the points where the GPU
kernels store bf16, so the GPU result can be compared tightly; with
`bf16=False` everything is FP64 (the finite-difference checks use that).
"""
import numpy as np

LAYERS = [  # name, Cin, Cout, kernel (3 or 1), relu
    ("c0", 5, 16, 3, True), ("c1", 16, 16, 3, True),
    ("c2", 16, 32, 3, True), ("c3", 32, 32, 3, True),
    ("c4", 32, 64, 3, True), ("c5", 64, 64, 3, True),
    ("c6", 96, 32, 3, True), ("c7", 32, 32, 3, True),
    ("c8", 48, 16, 3, True), ("c9", 16, 16, 3, True),
    ("head", 16, 1, 1, False),
]


def layer_sizes():
    out = []
    for name, ci, co, k, _ in LAYERS:
        taps = k * k * k
        out.append((name, co * taps * ci, co))
    return out


def param_count():
    return sum(w + b for _, w, b in layer_sizes())


def unpack(theta):
    """flat vector -> {name: (W[Cout, taps, Cin], b[Cout])}"""
    p, off = {}, 0
    for (name, ci, co, k, _), (_, nw, nb) in zip(LAYERS, layer_sizes()):
        W = theta[off:off + nw].reshape(co, k ** 3, ci)
        off += nw
        b = theta[off:off + nb]
        off += nb
        p[name] = (W, b)
    return p


def bf16(a):
    """round to bfloat16 (nearest even), returned as float64"""
    a32 = np.ascontiguousarray(a, dtype=np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    u = ((u + 0x7FFF + lsb) >> 16) << 16
    return (u.astype(np.uint32)).view(np.float32).astype(np.float64)


def init_params(seed):
    """He-normal weights, zero bias (head: small weights)"""
    rng = np.random.default_rng(seed)
    parts = []
    for name, ci, co, k, relu in LAYERS:
        fan = ci * k ** 3
        std = np.sqrt(2.0 / fan) if relu else 0.1 / np.sqrt(fan)
        parts.append(rng.normal(0, std, size=co * k ** 3 * ci))
        parts.append(np.zeros(co))
    return np.concatenate(parts)


OFFS = [(dz, dy, dx) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]


def shift(x, dz, dy, dx):
    """y[z, y, x] = x[z+dz, y+dy, x+dx] with zeros outside"""
    Z, Y, X = x.shape[:3]
    y = np.zeros_like(x)
    zs, ze = max(0, -dz), min(Z, Z - dz)
    ys, ye = max(0, -dy), min(Y, Y - dy)
    xs, xe = max(0, -dx), min(X, X - dx)
    y[zs:ze, ys:ye, xs:xe] = x[zs + dz:ze + dz, ys + dy:ye + dy, xs + dx:xe + dx]
    return y


def conv(x, W, b, k):
    if k == 1:
        return np.einsum("zyxc,oc->zyxo", x, W[:, 0, :]) + b
    out = np.zeros(x.shape[:3] + (W.shape[0],))
    for t, (dz, dy, dx) in enumerate(OFFS):
        out += np.einsum("zyxc,oc->zyxo", shift(x, dz, dy, dx), W[:, t, :])
    return out + b


def conv_back(x, W, k, gy):
    """gradients of conv wrt its input and weights, bias"""
    gb = gy.sum(axis=(0, 1, 2))
    if k == 1:
        gW = np.einsum("zyxo,zyxc->oc", gy, x)[:, None, :]
        gx = np.einsum("zyxo,oc->zyxc", gy, W[:, 0, :])
        return gx, gW, gb
    gW = np.zeros_like(W)
    gx = np.zeros_like(x)
    for t, (dz, dy, dx) in enumerate(OFFS):
        xs = shift(x, dz, dy, dx)
        gW[:, t, :] = np.einsum("zyxo,zyxc->oc", gy, xs)
        gx += shift(np.einsum("zyxo,oc->zyxc", gy, W[:, t, :]), -dz, -dy, -dx)
    return gx, gW, gb


def pool(x):
    Z, Y, X, C = x.shape
    return x.reshape(Z // 2, 2, Y // 2, 2, X // 2, 2, C).mean(axis=(1, 3, 5))


def pool_back(g):
    return np.repeat(np.repeat(np.repeat(g, 2, 0), 2, 1), 2, 2) / 8.0


def up(x):
    return np.repeat(np.repeat(np.repeat(x, 2, 0), 2, 1), 2, 2)


def up_back(g):
    Z, Y, X, C = g.shape
    return g.reshape(Z // 2, 2, Y // 2, 2, X // 2, 2, C).sum(axis=(1, 3, 5))


def forward(theta, inp, bf=False):
    """inp [Z, Y, X, 5] -> (out [Z, Y, X], cache)"""
    r = bf16 if bf else (lambda a: a)
    P = {k: (r(W), b) for k, (W, b) in unpack(theta).items()}
    cache = {}

    def layer(name, x):
        _, ci, co, k, relu = next(L for L in LAYERS if L[0] == name)
        W, b = P[name]
        z = conv(x, W, b, k)
        y = np.maximum(z, 0) if relu else z
        y = r(y) if relu else y
        cache[name] = (x, z)
        return y

    x0 = r(inp)
    a0 = layer("c0", x0)
    s1 = layer("c1", a0)
    p1 = r(pool(s1))
    a2 = layer("c2", p1)
    s2 = layer("c3", a2)
    p2 = r(pool(s2))
    a4 = layer("c4", p2)
    a5 = layer("c5", a4)
    u1 = np.concatenate([up(a5), s2], axis=3)
    a6 = layer("c6", u1)
    a7 = layer("c7", a6)
    u2 = np.concatenate([up(a7), s1], axis=3)
    a8 = layer("c8", u2)
    a9 = layer("c9", a8)
    out = layer("head", a9)[..., 0]
    cache["P"] = P
    return out, cache


def loss(out, target, mask):
    """masked squared error normalised by the target's energy"""
    m = mask.astype(bool)
    den = max(float((target[m] ** 2).sum()), 1e-30)
    d = np.where(m, out - target, 0.0)
    return float((d * d).sum() / den), 2.0 * d / den


def backward(theta, cache, gout, bf=False):
    """gradient of a scalar wrt theta given d scalar / d out"""
    r = bf16 if bf else (lambda a: a)
    P = cache["P"]
    grads = {}

    def layer_back(name, gy):
        _, ci, co, k, relu = next(L for L in LAYERS if L[0] == name)
        x, z = cache[name]
        if relu:
            gy = np.where(z > 0, gy, 0.0)
        gy = r(gy)
        gx, gW, gb = conv_back(x, P[name][0], k, gy)
        grads[name] = (gW, gb)
        return gx

    g = gout[..., None]
    g = layer_back("head", g)
    g = layer_back("c9", g)
    g = layer_back("c8", g)
    g_up2, g_s1 = g[..., :32], g[..., 32:]
    g = layer_back("c7", up_back(g_up2))
    g = layer_back("c6", g)
    g_up1, g_s2 = g[..., :64], g[..., 64:]
    g = layer_back("c5", up_back(g_up1))
    g = layer_back("c4", g)
    g = layer_back("c3", pool_back(g) + g_s2)
    g = layer_back("c2", g)
    g = layer_back("c1", pool_back(g) + g_s1)
    layer_back("c0", g)
    parts = []
    for name, *_ in LAYERS:
        gW, gb = grads[name]
        parts.append(gW.ravel())
        parts.append(gb)
    return np.concatenate(parts)


def inputs(r0, cell_type, q_prev):
    """the five input channels from the initial residual, the cell types
    (0 air, 1 fluid, 2 solid) and the previous pressure, each scaled by its
    RMS over fluid cells; returns (x [Z,Y,X,5], s) with s the residual scale"""
    fl = cell_type == 1
    s = float(np.sqrt((r0[fl] ** 2).mean())) if fl.any() else 1.0
    s = s if s > 0 else 1.0
    sq = float(np.sqrt((q_prev[fl] ** 2).mean())) if fl.any() else 1.0
    sq = sq if sq > 0 else 1.0
    x = np.zeros(r0.shape + (5,))
    x[..., 0] = np.where(fl, r0 / s, 0.0)
    x[..., 1] = cell_type == 0
    x[..., 2] = fl
    x[..., 3] = cell_type == 2
    x[..., 4] = np.where(fl, q_prev / sq, 0.0)
    return x, s
