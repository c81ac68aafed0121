"""CPU ORACLE of the learned coarse solver (test infrastructure only).

numpy restatement of the convolutional multilevel network used as the coarse
solve of the learned two-grid preconditioner, with its forward pass and its
backward pass (gradients of a loss with respect to every parameter and to the
input).  The device kernels store activations and weights in BF16 and
accumulate in FP32; this oracle rounds exactly where the device rounds (BF16
round-to-nearest-even of every stored activation and of the weights) and
accumulates in FP64, so the two agree to FP32 accumulation order.

Nothing in the product imports this file; only tests/ may.

Network (grid n^3 at level 0, n/2 at 1, n/4 at 2; C = (c0, c1, c2) channels):
    x1 = relu(conv(in, W0)) ; x1 = relu(conv(x1, W1))                 level 0
    x2 = relu(conv(pool(x1), W2)) ; x2 = relu(conv(x2, W3))            level 1
    x3 = relu(conv(pool(x2), W4)) ; x3 = relu(conv(x3, W5))            level 2
    y2 = relu(conv(up(x3), W6) + x2) ; y2 = relu(conv(y2, W7))         level 1
    y1 = relu(conv(up(y2), W8) + x1) ; out = conv(y1, W9)              level 0
conv = 3x3x3, stride 1, zero padding, with bias; pool = 2x2x2 mean;
up = nearest-neighbour x2.  Arrays are [C][z][y][x].
"""
import numpy as np


def bf16(a):
    """Round float32 values to bfloat16 (round to nearest even), returned as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    r = ((u + 0x7FFF + lsb) >> 16) << 16
    nan = np.isnan(a)
    out = r.astype(np.uint32).view(np.float32).copy()
    out[nan] = np.nan
    return out


def layer_shapes(cin, c):
    c0, c1, c2 = c
    return [(c0, cin), (c0, c0), (c1, c0), (c1, c1), (c2, c1), (c2, c2), (c1, c2), (c1, c1), (c0, c1), (1, c0)]


def num_params(cin, c):
    return sum(co * ci * 27 + co for co, ci in layer_shapes(cin, c))


def unpack(theta, cin, c):
    """Flat parameter vector -> [(W[co][ci][27], b[co])] in layer order (W then b per layer)."""
    out, k = [], 0
    for co, ci in layer_shapes(cin, c):
        W = theta[k:k + co * ci * 27].reshape(co, ci, 27)
        k += co * ci * 27
        b = theta[k:k + co]
        k += co
        out.append((W, b))
    assert k == len(theta)
    return out


TAPS = [(dz, dy, dx) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]


def shift(x, dz, dy, dx):
    """y[.., z, y, x] = x[.., z+dz, y+dy, x+dx] with zero outside."""
    C, Z, Y, X = x.shape
    out = np.zeros_like(x)
    zs, ze = max(0, -dz), min(Z, Z - dz)
    ys, ye = max(0, -dy), min(Y, Y - dy)
    xs, xe = max(0, -dx), min(X, X - dx)
    out[:, zs:ze, ys:ye, xs:xe] = x[:, zs + dz:ze + dz, ys + dy:ye + dy, xs + dx:xe + dx]
    return out


def conv(x, W, b, acc=np.float64):
    """3x3x3 zero-padded convolution: y[co] = b[co] + sum_ci sum_t W[co,ci,t] x[ci](. + t),
    accumulated in acc (FP64; FP32 to measure how FP32 sums move the BF16-stored
    activations that follow)."""
    co = W.shape[0]
    y = np.zeros((co,) + x.shape[1:], dtype=acc)
    for t, (dz, dy, dx) in enumerate(TAPS):
        xs = shift(x.astype(acc), dz, dy, dx)
        y += np.einsum("oc,czyx->ozyx", W[:, :, t].astype(acc), xs)
    return (y + b.astype(acc)[:, None, None, None]).astype(np.float64)


def conv_backward(x, W, dy):
    """Gradients of sum(dy * conv(x, W, b)) with respect to x, W and b."""
    dy = dy.astype(np.float64)
    dx = np.zeros(x.shape, dtype=np.float64)
    dW = np.zeros(W.shape, dtype=np.float64)
    for t, (dz, dyy, dxx) in enumerate(TAPS):
        xs = shift(x.astype(np.float64), dz, dyy, dxx)
        dW[:, :, t] = np.einsum("ozyx,czyx->oc", dy, xs)
        g = np.einsum("oc,ozyx->czyx", W[:, :, t].astype(np.float64), dy)
        dx += shift(g, -dz, -dyy, -dxx)
    db = dy.sum(axis=(1, 2, 3))
    return dx, dW, db


def pool(x):
    C, Z, Y, X = x.shape
    return x.reshape(C, Z // 2, 2, Y // 2, 2, X // 2, 2).mean(axis=(2, 4, 6))


def pool_backward(dy):
    return np.repeat(np.repeat(np.repeat(dy, 2, 1), 2, 2), 2, 3) / 8.0


def up(x):
    return np.repeat(np.repeat(np.repeat(x, 2, 1), 2, 2), 2, 3)


def up_backward(dy):
    C, Z, Y, X = dy.shape
    return dy.reshape(C, Z // 2, 2, Y // 2, 2, X // 2, 2).sum(axis=(2, 4, 6))


def forward(inp, theta, c, round_bf16=True, acc=np.float64):
    """Forward pass; returns (out, cache).  inp: [cin][n][n][n].
    round_bf16=False evaluates the same network exactly in FP64 (used to pin
    the backward pass by finite differences).  acc: the convolutions'
    accumulation type (FP32: the spread that FP32 sums, as the device forms
    them, give the BF16-rounded activations; a one-ulp BF16 difference where a
    sum lies within FP32 rounding of a BF16 rounding boundary)."""
    cin = inp.shape[0]
    if round_bf16:
        L = unpack(bf16(theta), cin, c)
        st = lambda v: bf16(v)  # every stored activation is BF16
    else:
        L = unpack(np.asarray(theta, dtype=np.float64), cin, c)
        st = lambda v: np.asarray(v, dtype=np.float64)
    relu = lambda v: np.maximum(v, 0.0)
    a0 = st(inp)
    x1a = st(relu(conv(a0, *L[0], acc=acc)))
    x1 = st(relu(conv(x1a, *L[1], acc=acc)))
    p1 = st(pool(x1))
    x2a = st(relu(conv(p1, *L[2], acc=acc)))
    x2 = st(relu(conv(x2a, *L[3], acc=acc)))
    p2 = st(pool(x2))
    x3a = st(relu(conv(p2, *L[4], acc=acc)))
    x3 = st(relu(conv(x3a, *L[5], acc=acc)))
    u2 = up(x3)
    y2a = st(relu(conv(u2, *L[6], acc=acc) + x2))
    y2 = st(relu(conv(y2a, *L[7], acc=acc)))
    u1 = up(y2)
    y1a = st(relu(conv(u1, *L[8], acc=acc) + x1))
    out = conv(y1a, *L[9], acc=acc)
    out = out.astype(np.float32) if round_bf16 else out
    cache = dict(L=L, a0=a0, x1a=x1a, x1=x1, p1=p1, x2a=x2a, x2=x2, p2=p2, x3a=x3a, x3=x3, u2=u2, y2a=y2a, y2=y2,
                 u1=u1, y1a=y1a)
    return out, cache


def backward(dout, cache):
    """Gradient of sum(dout * out) with respect to the flat parameters and the input."""
    L = cache["L"]
    grads = [None] * 10
    mask = lambda g, y: g * (y > 0)
    dy1a, grads[9] = _bw(cache["y1a"], L[9], dout)
    g = mask(dy1a, cache["y1a"])
    du1, grads[8] = _bw(cache["u1"], L[8], g)
    dx1_skip = g
    dy2 = up_backward(du1)
    g = mask(dy2, cache["y2"])
    dy2a, grads[7] = _bw(cache["y2a"], L[7], g)
    g = mask(dy2a, cache["y2a"])
    du2, grads[6] = _bw(cache["u2"], L[6], g)
    dx2_skip = g
    dx3 = up_backward(du2)
    g = mask(dx3, cache["x3"])
    dx3a, grads[5] = _bw(cache["x3a"], L[5], g)
    g = mask(dx3a, cache["x3a"])
    dp2, grads[4] = _bw(cache["p2"], L[4], g)
    dx2 = pool_backward(dp2) + dx2_skip
    g = mask(dx2, cache["x2"])
    dx2a, grads[3] = _bw(cache["x2a"], L[3], g)
    g = mask(dx2a, cache["x2a"])
    dp1, grads[2] = _bw(cache["p1"], L[2], g)
    dx1 = pool_backward(dp1) + dx1_skip
    g = mask(dx1, cache["x1"])
    dx1a, grads[1] = _bw(cache["x1a"], L[1], g)
    g = mask(dx1a, cache["x1a"])
    dinp, grads[0] = _bw(cache["a0"], L[0], g)
    flat = np.concatenate([np.concatenate([dW.ravel(), db]) for dW, db in grads])
    return flat, dinp


def _bw(x, layer, dy):
    W, _ = layer
    dx, dW, db = conv_backward(x, W, dy)
    return dx, (dW, db)


def init_params(cin, c, seed):
    """He-normal weights (fan-in 27 cin), zero biases, last layer scaled down."""
    rng = np.random.default_rng(seed)
    parts = []
    shapes = layer_shapes(cin, c)
    for k, (co, ci) in enumerate(shapes):
        std = np.sqrt(2.0 / (27 * ci)) * (0.1 if k == len(shapes) - 1 else 1.0)
        parts.append(rng.normal(0.0, std, size=co * ci * 27))
        parts.append(np.zeros(co))
    return np.concatenate(parts).astype(np.float32)
